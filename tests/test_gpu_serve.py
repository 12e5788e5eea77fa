"""The serving loop (kache_amd_server_*): requests of every kind run in waves
on one worker, and each request's outputs must be bit-exact those of running
the requests alone, one after another, in served order (wave, phase,
ticket) -- checked against the oracle, and the final table (keys, values,
capacity) against a second device table fed the same calls serially."""
import threading
import time

import numpy as np
import pytest

import kache_hash_amd as KA
import oracle as O
import seqgen as G
from kache_hash_amd import capi

pytestmark = pytest.mark.gpu
DT = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}


@pytest.fixture(scope="module")
def api():
    return KA.api()


def random_seqs(rng, g, m, max_len=400):
    """m pieces of genome g: random lengths (some shorter than k, some
    empty), strands, substitutions, N's and lowercase runs."""
    ga = np.frombuffer(g, np.uint8)
    out = []
    for _ in range(m):
        L = int(rng.integers(0, max_len))
        s0 = int(rng.integers(0, len(g) - L))
        s = ga[s0:s0 + L].copy()
        if rng.random() < 0.5:
            s = np.frombuffer(G.revcomp(s.tobytes()), np.uint8).copy()
        if L:
            sub = rng.random(L) < 0.01
            s[sub] = G.ACGT[rng.integers(0, 4, int(sub.sum()))]
            s[rng.random(L) < 0.002] = ord("N")
            if rng.random() < 0.2:
                a = int(rng.integers(0, L))
                s[a:a + 40] |= 0x20
        out.append(s.tobytes())
    return out


OPS_INSERT = [(capi.ADD, 1), (capi.ADD, 1), (capi.MAX, 7), (capi.ASSIGN, 5)]
OPS_UPDATE = [(capi.ADD, 3), (capi.MIN, 2), (capi.OR, 16)]


def make_requests(seed, g, n, width, p=(0.45, 0.2, 0.1, 0.25), max_seqs=40):
    rng = np.random.default_rng(seed)
    reqs = []
    for _ in range(n):
        kind = int(rng.choice(4, p=p))
        seqs = random_seqs(rng, g, int(rng.integers(0, max_seqs)))
        op, x = 0, 0
        if kind == capi.REQ_INSERT:
            op, x = OPS_INSERT[int(rng.integers(0, len(OPS_INSERT)))]
        elif kind == capi.REQ_UPDATE:
            op, x = OPS_UPDATE[int(rng.integers(0, len(OPS_UPDATE)))]
        if not width and kind == capi.REQ_INSERT:
            op, x = capi.KEEP, 0
        reqs.append((kind, seqs, op, x))
    return reqs


def plain_call(t, user, kind, seqs, op, x):
    """The request as one plain C-ABI call: (status, outputs)."""
    try:
        if kind == capi.REQ_COUNT:
            return 0, {"counts": t.count(seqs)}
        if kind == capi.REQ_LOOKUP:
            f, v, fw = t.lookup(seqs)
            return 0, {"found": f, "values": v, "fwd": fw}
        if kind == capi.REQ_UPDATE:
            return 0, {"counts": np.array([t.update(seqs, op, x)], np.uint64)}
        return 0, {"counts": np.array([t.insert(seqs, op, x, user)], np.uint64)}
    except capi.KacheError as e:
        return e.status, {}


def oracle_call(o, width, kind, seqs, op, x):
    if kind == capi.REQ_COUNT:
        return {"counts": o.counts(seqs)[0]}
    if kind == capi.REQ_LOOKUP:
        f, v, fw = o.lookup(seqs)
        return {"found": f, "values": v if width else None, "fwd": fw}
    if kind == capi.REQ_UPDATE:
        return {"counts": np.array([o.update(seqs, op, x)], np.uint64)}
    return {"counts": np.array([o.insert(seqs, op if width else 0, x if width else 0)], np.uint64)}


def served_outputs(width, kind, seqs, out):
    """A served request's outputs in the plain call's form."""
    n = len(seqs)
    p = capi.Batch(seqs).positions()
    if kind == capi.REQ_COUNT:
        return {"counts": out["counts"][:n]}
    if kind == capi.REQ_LOOKUP:
        v = out["values"][:p * width].view(DT[width]).astype(np.uint64) if width else None
        return {"found": out["found"][:p], "values": v, "fwd": out["fwd"][:p]}
    return {"counts": out["counts"][:1]}


def same(a, b):
    for k in a:
        if a[k] is None or b.get(k) is None:
            assert a[k] is None and b.get(k) is None, k
            continue
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k


def serve(api, width, reqs, wave_bytes, max_wave, threads=1, expected=20_000, prefill=None, gaps=None, pinned=False):
    """Serves `reqs` (one submitter: request i after a pause of gaps[i]
    seconds, if given), then checks every output against the oracle and a
    serial device table run in served order.  With `pinned`, outputs are
    page-locked (lookups copied into them by DMA directly).  Returns the
    served results."""
    t = capi.Table(api, width, expected_keys=expected)
    u = t.register()
    o = O.OTable(width)
    ser = capi.Table(api, width, expected_keys=expected)
    su = ser.register()
    if prefill:
        t.insert(prefill, capi.ADD if width else capi.KEEP, 1 if width else 0, u)
        ser.insert(prefill, capi.ADD if width else capi.KEEP, 1 if width else 0, su)
        o.insert(prefill, capi.ADD if width else 0, 1 if width else 0)
    with capi.Server(t, u, wave_bytes, max_wave) as s:
        made = [s.request(k, seqs, op, x, pinned=pinned) for k, seqs, op, x in reqs]
        tickets = [None] * len(reqs)
        if threads == 1:
            for i, r in enumerate(made):
                if gaps is not None and gaps[i] > 0:
                    time.sleep(gaps[i])
                tickets[i] = s.submit(r)
        else:
            def submitter(w):
                for i in range(w, len(made), threads):
                    tickets[i] = s.submit(made[i])
            th = [threading.Thread(target=submitter, args=(w,)) for w in range(threads)]
            for x in th:
                x.start()
            for x in th:
                x.join()
        results = [s.wait(tk) for tk in tickets]
        stats = s.stats()
    assert stats["requests"] == len(reqs)
    order = sorted(range(len(reqs)), key=lambda i: (results[i][0]["wave"], results[i][0]["phase"], results[i][0]["ticket"]))
    for i in order:
        kind, seqs, op, x = reqs[i]
        res, out = results[i]
        st, want = plain_call(ser, su, kind, seqs, op, x)
        assert res["status"] == st, (i, kind, res)
        assert res["submit_s"] <= res["start_s"] <= res["done_s"]
        if st:
            continue
        got = served_outputs(width, kind, seqs, out)
        same(got, want)
        same(got, oracle_call(o, width, kind, seqs, op, x))
    ka, va, _ = t.for_each()
    kb, vb, _ = ser.for_each()
    kc, vc = o.dump()
    assert np.array_equal(ka, kb) and np.array_equal(va, vb)
    oc = np.argsort(kc)
    assert np.array_equal(ka, kc[oc]) and (not width or np.array_equal(va, vc[oc]))
    assert t.stats()["capacity"] == ser.stats()["capacity"] and t.size() == ser.size()
    return [r for r, _ in results], stats


def test_serve_mixed_wide_waves(api):
    g = G.genome(200_000, 51)
    reqs = make_requests(1, g, 300, 4)
    print("serve u32, 300 requests, one wave budget of 64 MiB", flush=True)
    res, stats = serve(api, 4, reqs, 64 << 20, 1 << 20, prefill=[g[:50_000]])
    assert stats["waves"] < len(reqs)          # requests merged


def test_serve_page_locked_outputs(api):
    """Served lookups into page-locked outputs (copied by DMA directly), in
    waves mixing every kind, with the table doubling meanwhile."""
    g = G.genome(300_000, 57)
    reqs = make_requests(8, g, 400, 8, p=(0.2, 0.45, 0.1, 0.25), max_seqs=60)
    print("serve u64, 400 requests, page-locked outputs", flush=True)
    res, stats = serve(api, 8, reqs, 64 << 20, 1 << 20, expected=2_000, prefill=[g[:40_000]], pinned=True)
    assert stats["waves"] < len(reqs)


def test_serve_small_waves_and_oversized_requests(api):
    g = G.genome(200_000, 52)
    reqs = make_requests(2, g, 200, 4)
    print("serve u32, 200 requests, 4 KiB waves", flush=True)
    res, stats = serve(api, 4, reqs, 4096, 64)
    assert stats["waves"] > 20 and stats["max_wave_bytes"] > 4096   # a large request runs alone


def test_serve_one_request_per_wave_is_plain_calls(api):
    g = G.genome(100_000, 53)
    reqs = make_requests(3, g, 80, 8)
    print("serve u64, 80 requests, one per wave", flush=True)
    res, stats = serve(api, 8, reqs, 64 << 20, 1)
    assert stats["waves"] == len(reqs)
    assert [r["ticket"] for r in sorted(res, key=lambda r: r["wave"])] == sorted(r["ticket"] for r in res)


def test_serve_spaced_arrivals_run_lone_requests_unstaged(api):
    """Requests 3 ms apart reach an idle server alone and run unstaged by
    the plain calls; a burst after them is merged into staged waves; both
    kinds of wave interleave in served order."""
    g = G.genome(150_000, 57)
    reqs = make_requests(7, g, 120, 4)
    gaps = [0.003 if i < 40 or i % 10 == 0 else 0.0 for i in range(len(reqs))]
    print("serve u32, 120 requests, spaced then bursts", flush=True)
    res, stats = serve(api, 4, reqs, 64 << 20, 1 << 20, prefill=[g[:40_000]], gaps=gaps)
    assert 0 < stats["direct_waves"] < stats["waves"] < len(reqs)


def test_serve_set_table_rejects_updates_per_request(api):
    g = G.genome(100_000, 54)
    reqs = make_requests(4, g, 120, 0, p=(0.35, 0.2, 0.15, 0.3))
    print("serve set, 120 requests (updates are EINVAL on a set)", flush=True)
    res, _ = serve(api, 0, reqs, 1 << 20, 1 << 20)
    bad = [r for r, q in zip(res, reqs) if q[0] == capi.REQ_UPDATE]
    assert bad and all(r["status"] == capi.KACHE_EINVAL for r in bad)


def test_serve_concurrent_submitters_and_growth(api):
    g = G.genome(400_000, 55)
    reqs = make_requests(5, g, 400, 2, p=(0.3, 0.1, 0.1, 0.5), max_seqs=60)
    print("serve u16, 400 requests from 4 threads, table grows", flush=True)
    serve(api, 2, reqs, 256 << 10, 1 << 20, threads=4, expected=2_000)


def test_replay_open_loop(api):
    g = G.genome(100_000, 56)
    reqs = make_requests(6, g, 60, 4)
    t = capi.Table(api, 4, expected_keys=50_000)
    u = t.register()
    with capi.Server(t, u, 1 << 20, 1 << 20) as s:
        made = [s.request(k, seqs, op, x) for k, seqs, op, x in reqs]
        arrivals = np.cumsum(np.full(len(reqs), 0.002))
        print("replay 60 requests over 120 ms", flush=True)
        t0 = api.lib.kache_amd_clock()
        res = s.replay(made, arrivals)
    assert all(r["status"] == 0 for r in res)
    sub = np.array([r["submit_s"] for r in res]) - t0
    assert (sub >= arrivals - 1e-4).all() and (sub <= arrivals + 0.5).all()
    assert all(r["submit_s"] <= r["start_s"] <= r["done_s"] for r in res)
    assert sorted(r["ticket"] for r in res) == [r["ticket"] for r in res]


def test_lookup_groups_equals_separate_lookups(api):
    """kache_amd_lookup_groups: one pass, each group's outputs laid out as a
    lookup of its own sequences (empty groups and groups without k-mers
    included), with and without values / fwd."""
    import ctypes as C
    g = G.genome(100_000, 57)
    rng = np.random.default_rng(57)
    t = capi.Table(api, 2, expected_keys=200_000)
    u = t.register()
    t.insert(random_seqs(rng, g, 300), capi.ADD, 1, u)
    groups = [random_seqs(rng, g, int(m)) for m in (5, 0, 17, 1, 0, 40, 3)]
    groups[3] = [b"ACGTN" * 4]          # a group without k-mers
    seqs = [s for grp in groups for s in grp]
    b = capi.Batch(seqs)
    first = np.cumsum([0] + [len(x) for x in groups[:-1]]).astype(np.uint64)
    for with_values, with_fwd in ((True, True), (False, True), (True, False)):
        outs = [capi.make_request(2, capi.REQ_LOOKUP, grp, values=with_values, fwd=with_fwd)[1] for grp in groups]
        P = C.c_void_p * len(groups)
        found = P(*[o["found"].ctypes.data for o in outs])
        values = P(*[o["values"].ctypes.data if "values" in o else None for o in outs])
        fwd = P(*[o["fwd"].ctypes.data if "fwd" in o else None for o in outs])
        api.check(api.lib.kache_amd_lookup_groups(t.h, b.c_seqs, b.c_lens, b.n, first.ctypes.data, len(groups), found,
                                                  values if with_values else None, fwd if with_fwd else None), "lookup_groups")
        for grp, o in zip(groups, outs):
            f, v, fw = t.lookup(grp)
            p = len(f)
            assert np.array_equal(o["found"][:p], f)
            if with_values:
                assert np.array_equal(o["values"][:2 * p].view(np.uint16).astype(np.uint64), v)
            if with_fwd:
                assert np.array_equal(o["fwd"][:p], fw)
    bad = np.array([1, 0], np.uint64)
    one = (C.c_void_p * 2)(1, 1)
    assert api.lib.kache_amd_lookup_groups(t.h, b.c_seqs, b.c_lens, b.n, bad.ctypes.data, 2, one, None, None) == capi.KACHE_EINVAL


def test_device_text_queries_equal_host_queries(api):
    """kache_amd_count_seqs_device / kache_amd_lookup_groups_device on a device
    text give kache_count / kache_lookup of the same sequences."""
    import ctypes as C
    g = G.genome(80_000, 58)
    rng = np.random.default_rng(58)
    t = capi.Table(api, 4, expected_keys=100_000)
    u = t.register()
    t.insert(random_seqs(rng, g, 200), capi.ADD, 3, u)
    seqs = random_seqs(rng, g, 60)
    text = b"".join(x + b"\n" for x in seqs)
    off = np.concatenate([[0], np.cumsum([len(x) + 1 for x in seqs])]).astype(np.uint64)
    d = capi.DeviceBuffer(api, len(text) + capi.TEXT_PAD)
    d.upload(text + b"\n" * capi.TEXT_PAD)
    out = np.zeros(len(seqs), np.uint64)
    api.check(api.lib.kache_amd_count_seqs_device(t.h, d.ptr, len(text), off.ctypes.data, len(seqs), out.ctypes.data), "count")
    assert np.array_equal(out, t.count(seqs))
    groups = [seqs[:10], seqs[10:10], seqs[10:45], seqs[45:]]
    first = np.array([0, 10, 10, 45], np.uint64)
    outs = [capi.make_request(4, capi.REQ_LOOKUP, grp)[1] for grp in groups]
    P = C.c_void_p * 4
    api.check(api.lib.kache_amd_lookup_groups_device(t.h, d.ptr, len(text), off.ctypes.data, len(seqs), first.ctypes.data, 4,
                                                     P(*[o["found"].ctypes.data for o in outs]),
                                                     P(*[o["values"].ctypes.data for o in outs]),
                                                     P(*[o["fwd"].ctypes.data for o in outs])), "lookup")
    for grp, o in zip(groups, outs):
        f, v, fw = t.lookup(grp)
        p = len(f)
        assert np.array_equal(o["found"][:p], f) and np.array_equal(o["fwd"][:p], fw)
        assert np.array_equal(o["values"][:4 * p].view(np.uint32).astype(np.uint64), v)
    bad = off.copy()
    bad[3] = bad[2]                       # an empty span: no room for the separator
    assert api.lib.kache_amd_count_seqs_device(t.h, d.ptr, len(text), bad.ctypes.data, len(seqs), out.ctypes.data) == capi.KACHE_EINVAL


def test_serve_hbm_budget_refuses_an_oversized_request(api):
    """Admission by HBM budget: with the server's staging budget at its two
    slots, a request larger than a slot completes with KACHE_ENOMEM without
    running, whether it arrives alone or among others; every other request
    runs, with the outputs of a serial run in served order."""
    g = G.genome(200_000, 71)
    rng = np.random.default_rng(71)
    prefill = random_seqs(rng, g, 300)
    wave = 64 << 10
    t = capi.Table(api, 4, expected_keys=100_000)
    u = t.register()
    t.insert(prefill, capi.ADD, 1, u)
    o = O.OTable(4)
    o.insert(prefill, capi.ADD, 1)
    big = [g[i:i + 2000] for i in range(0, 200_000 - 2000, 1500)]       # ~ 265 KB staged: over the slot
    smalls = [random_seqs(rng, g, 20, 300) for _ in range(12)]
    with capi.Server(t, u, wave, 1 << 20) as s:
        slots = 2 * (wave + 4096 + capi.TEXT_PAD)
        s.set_hbm_budget(slots)
        alone = s.request(capi.REQ_LOOKUP, big)
        r_alone, _ = s.wait(s.submit(alone))
        made = [s.request(capi.REQ_COUNT, x) for x in smalls[:6]] + [s.request(capi.REQ_LOOKUP, big)] + \
               [s.request(capi.REQ_COUNT, x) for x in smalls[6:]]
        results = [s.wait(tk) for tk in [s.submit(r) for r in made]]
        stats = s.stats()
    assert r_alone["status"] == capi.KACHE_ENOMEM
    assert results[6][0]["status"] == capi.KACHE_ENOMEM
    assert stats["refused"] == 2
    for i, (res, out) in enumerate(results):
        if i == 6:
            continue
        assert res["status"] == capi.KACHE_OK
        seqs = smalls[i if i < 6 else i - 1]
        assert np.array_equal(out["counts"][:len(seqs)], o.counts(seqs)[0])
    # Without the budget the same request runs.
    with capi.Server(t, u, wave, 1 << 20) as s:
        r, out = s.wait(s.submit(s.request(capi.REQ_LOOKUP, big)))
    assert r["status"] == capi.KACHE_OK
    f, v, fw = o.lookup(big)
    p = len(f)
    assert np.array_equal(out["found"][:p], f) and np.array_equal(out["fwd"][:p], fw)
    t.close()


def test_lookups_into_page_locked_outputs(api):
    """Lookup outputs in page-locked memory (kache_amd_host_alloc) are copied
    by DMA directly, others through pinned chunks: both give the outputs of
    kache_lookup, for requests above and below the chunked path's 16 MiB,
    served and through kache_amd_lookup_groups."""
    import ctypes as C
    rng = np.random.default_rng(31)
    g = G.genome(3_000_000, 32)
    t = capi.Table(api, 4, expected_keys=4_000_000)
    u = t.register()
    t.insert([g[i:i + 100_000] for i in range(0, len(g), 100_000)], capi.ADD, 1, u)
    big = [g[int(s):int(s) + 100_000] for s in rng.integers(0, len(g) - 100_000, 60)]      # ~6M positions: 24 MB of values
    small = random_seqs(rng, g, 50)
    for seqs in (big, small):
        f, v, fw = t.lookup(seqs)
        p = len(f)
        for pinned in (False, True):
            r, o, b = capi.make_request(4, capi.REQ_LOOKUP, seqs, api=api if pinned else None)
            P = C.c_void_p * 1
            first = np.zeros(1, np.uint64)
            api.check(api.lib.kache_amd_lookup_groups(t.h, b.c_seqs, b.c_lens, b.n, first.ctypes.data, 1,
                                                      P(o["found"].ctypes.data), P(o["values"].ctypes.data),
                                                      P(o["fwd"].ctypes.data)), "lookup_groups")
            assert np.array_equal(o["found"][:p], f) and np.array_equal(o["fwd"][:p], fw)
            assert np.array_equal(o["values"][:4 * p].view(np.uint32).astype(np.uint64), v)
            r, o, b = capi.make_request(4, capi.REQ_LOOKUP, seqs, api=api if pinned else None)
            with capi.Server(t, u, 64 << 20, 1 << 20) as s:
                res = s.replay([(r, o, b)], np.zeros(1))
            assert res[0]["status"] == 0
            assert np.array_equal(o["found"][:p], f) and np.array_equal(o["fwd"][:p], fw)
            assert np.array_equal(o["values"][:4 * p].view(np.uint32).astype(np.uint64), v)
    t.close()

