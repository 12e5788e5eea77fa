"""The FASTA / FASTQ reader of the library (kache_amd_read_fastx), on the CPU,
over the input shapes of the reference's reader test (test/counter.cpp:64-125,
240-300): one-line and multi-line FASTA with empty lines, "\\n" and "\\r\\n"
line ends, no final line end, FASTQ with quality lines starting with '@' or
'+', header-only records, an empty file, and the malformed inputs the
reference refuses.  The expected sequences follow the format rules
(counter/Fastx_Reader.hpp:17-22), with known answers for small cases."""
import os

import numpy as np
import pytest

import kache_hash_amd as KA
from kache_hash_amd import capi


@pytest.fixture(scope="module")
def api():
    return KA.api()


def records(rng, c, max_len, long_len):
    alpha = np.frombuffer(b"ACGTACGTACGTACGTACGTacgtNRY", np.uint8)
    out = []
    for i in range(c):
        n = long_len if rng.integers(0, 100) == 0 else int(rng.integers(0, max_len + 1))
        out.append((b"r%d description" % i, alpha[rng.integers(0, len(alpha), n)].tobytes()))
    return out


def fasta(R, rng, width, eol, last_eol=True):
    s = []
    for name, seq in R:
        s.append(b">" + name + eol)
        if width == 0:
            s.append(seq + eol)
            continue
        w = 1 + int(rng.integers(0, width))
        for i in range(0, len(seq), w):
            s.append(seq[i:i + w] + eol)
            if rng.integers(0, 50) == 0:
                s.append(eol)
    t = b"".join(s)
    return t[:-len(eol)] if (not last_eol and t) else t


def fastq(R, rng, eol, last_eol=True):
    s = []
    qa = np.frombuffer(b"!#%+5:@AEIJ", np.uint8)
    for name, seq in R:
        q = bytearray(qa[rng.integers(0, len(qa), len(seq))].tobytes())
        if q and rng.integers(0, 4) == 0:
            q[0] = ord("@") if rng.integers(0, 2) else ord("+")
        s.append(b"@" + name + eol + seq + eol + b"+" + eol + bytes(q) + eol)
    t = b"".join(s)
    return t[:-len(eol)] if (not last_eol and t) else t


def expected(R):
    return [seq for _, seq in R if seq]


def read(api, tmp_path, name, text):
    p = tmp_path / name
    p.write_bytes(text)
    return capi.read_fastx(api, str(p))


def test_known_answers(api, tmp_path):
    assert read(api, tmp_path, "a.fa", b">a\nACG\nTT\r\n>b\n\n>c x\r\nGG\nNN") == [b"ACGTT", b"GGNN"]
    assert read(api, tmp_path, "b.fq", b"@r\nACGT\n+\n@III\n@s\n\n+\n\n@t\r\nGA\r\n+x\r\n+I\r\n") == [b"ACGT", b"GA"]
    assert read(api, tmp_path, "c.fa", b"") == []
    assert read(api, tmp_path, "d.fa", b">only\n>headers\n") == []


def test_reader_shapes(api, tmp_path):
    rng = np.random.default_rng(1)
    R = records(rng, 3000, 400, 70_000)
    Q = records(rng, 3000, 300, 20_000)
    cases = {
        "fasta": (fasta(R, rng, 0, b"\n"), R),
        "fasta-lines": (fasta(R, rng, 80, b"\n"), R),
        "fasta-lines-crlf": (fasta(R, rng, 80, b"\r\n"), R),
        "fasta-no-last-eol": (fasta(R, rng, 0, b"\n", False), R),
        "fastq": (fastq(Q, rng, b"\n"), Q),
        "fastq-crlf": (fastq(Q, rng, b"\r\n"), Q),
        "fastq-no-last-eol": (fastq(Q, rng, b"\n", False), Q),
    }
    H = records(rng, 100, 0, 0)
    cases["headers"] = (fasta(H, rng, 0, b"\n"), H)
    for name, (text, recs) in cases.items():
        assert read(api, tmp_path, name, text) == expected(recs), name


def test_reader_blocks_of_a_large_file(api, tmp_path):
    """A file over several read and text blocks (16 MiB reads, 64 MiB blocks),
    with a sequence longer than a read block."""
    rng = np.random.default_rng(2)
    R = records(rng, 200, 300_000, 20 << 20)
    R.append((b"long", b"ACGT" * (5 << 20)))
    text = fasta(R, rng, 1000, b"\n")
    assert read(api, tmp_path, "big.fa", text) == expected(R)


@pytest.mark.parametrize("what,text", [
    ("short quality", b"@r\nACGT\n+\nII\n"),
    ("no '+' line", b"@r\nACGT\nIIII\n@s\nAC\n+\nII\n"),
    ("cut record", b"@r\nACGT\n+\nIIII\n@s\nAC\n"),
    ("unknown format", b"ACGT\n"),
    ("gzip", b"\x1f\x8b\x08\x00"),
])
def test_malformed_inputs_fail(api, tmp_path, what, text):
    with pytest.raises(capi.KacheError) as e:
        read(api, tmp_path, "bad", text)
    assert e.value.status == capi.KACHE_EINVAL
    assert api.last_error()


def test_absent_file_fails(api, tmp_path):
    with pytest.raises(capi.KacheError):
        capi.read_fastx(api, str(tmp_path / "absent"))


@pytest.mark.parametrize("kind", ["fasta", "fastq"])
def test_reader_named_pipe(api, tmp_path, kind):
    """Input from a named pipe, written in small pieces by another thread while
    the reader reads (as test/counter.cpp:198-225 feeds the reference's
    reader): the same sequences as from a regular file."""
    import threading
    import time
    rng = np.random.default_rng(3)
    R = records(rng, 2000, 500, 50_000)
    text = fasta(R, rng, 70, b"\n") if kind == "fasta" else fastq(R, rng, b"\r\n")
    p = tmp_path / ("in." + kind)
    os.mkfifo(p)

    def writer():
        with open(p, "wb", buffering=0) as f:
            for i in range(0, len(text), 65_537):
                f.write(text[i:i + 65_537])
                if i % (16 * 65_537) == 0:
                    time.sleep(0.001)

    th = threading.Thread(target=writer)
    th.start()
    try:
        got = capi.read_fastx(api, str(p))
    finally:
        th.join(timeout=60)
    assert got == expected(R)
