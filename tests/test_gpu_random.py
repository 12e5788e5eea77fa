"""The seeded random programs of tests/programs.py through the device
library, against the oracle: C-ABI programs (the same ones run against the
reference library in test_oracle.py) and programs of the library's own
device-text, batch and host-text entry points."""
import pytest

import kache_hash_amd as KA
import programs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api():
    return KA.api()


@pytest.mark.parametrize("case", range(programs.CASES))
def test_random_program(api, case):
    programs.run_program(api, case)


@pytest.mark.parametrize("case", range(programs.DEVICE_CASES))
def test_random_device_program(api, case):
    programs.run_device_program(api, case)
