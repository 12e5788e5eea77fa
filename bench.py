"""Benchmark: the k=31 count map on one MI355X.

Workload (the count-map configuration): a <31, u64, l=22> map sized for
100M keys; a 100 Mbp seeded uniform genome; 20M reads of 150 bp (uniform
starts, random strand, 0.5% substitutions, 0.05% N, lowercase runs in half
the reads), upserted with +1 (absent keys start at 1) in batches of 65,536
reads; then lookups of 2M fresh reads.  A step is one batch upsert.  Reads
are generated on the device and resident in HBM before the timed region.

Prints one JSON line: k-mer upserts/s over the timed steps, with lookups/s,
the insert kernel's roofline position, and the reference CPU path timed on a
bounded sample of the same reads on this host's cores.
"""
import argparse
import collections
import ctypes as C
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

READ_LEN = 150
PEAK_HBM_GBS = 8000.0
# Algorithmic bytes per k-mer of a u64 count-map upsert/lookup over 150-bp reads:
# input 150/120 = 1.25 B, the bucket metadata line 64 B / 5.29 k-mers per
# minimizer run = 12.10 B, and the 16-byte key + value slot.
ALG_BYTES_PER_KMER = 150 / 120 + 64 / 5.29 + 16


# Bucket lines touched per k-mer of a u64 count map (the 64-byte metadata line
# per 5.29-k-mer minimizer run, a quarter of a 64-byte line of 16-byte slots),
# as the random-access roofline counts them.
LINES_PER_KMER = 1 / 5.29 + 16 / 64

# Timed passes whose launches are bracketed by HIP events: one in PROFILE_EVERY.
PROFILE_EVERY = 4
# Timed lookup passes over the lookup reads (the median is reported).
LOOKUP_PASSES = 5
GROWTH = {"auto": 0, "in-place": 1, "out-of-place": 2}    # KACHE_AMD_GROW_*


def transaction_roof(kmers_per_s):
    """The insert kernel against the device's random-transaction capacity:
    its memory reads and memory-side atomics per k-mer (rocprofv3 counters,
    profiles/insert_count_map_traffic.json) at the measured k-mers/s, each
    over its measured roof (random_read of 128-byte lines, random_atomic;
    profiles/memory_roof_mi355x.json), summed."""
    try:
        tr = json.load(open(os.path.join(ROOT, "profiles", "insert_count_map_traffic.json")))
        roof = json.load(open(os.path.join(ROOT, "profiles", "memory_roof_mi355x.json")))
        reads, atomics = tr["memory_reads_per_kmer"], tr["memory_atomics_per_kmer"]
        r_roof, a_roof = roof["random_read_128_lines_per_s"], roof["random_atomic_64_lines_per_s"]
    except (OSError, KeyError, ValueError):
        return None
    r, a = kmers_per_s * reads, kmers_per_s * atomics
    return {"reads_per_kmer": reads, "atomics_per_kmer": atomics, "reads_per_s": r, "atomics_per_s": a,
            "read_roof_per_s": r_roof, "atomic_roof_per_s": a_roof, "frac": r / r_roof + a / a_roof}


def random_line_roof(kmers_per_s):
    """The insert kernel against the measured random read-modify-write roof of
    64-byte lines (profiles/memory_roof_mi355x.json, made by
    kache_hash_amd/memory_roof on an MI355X)."""
    path = os.path.join(ROOT, "profiles", "memory_roof_mi355x.json")
    try:
        roof = json.load(open(path))["random_update_64_lines_per_s"]
    except (OSError, KeyError, ValueError):
        return None
    lines = kmers_per_s * LINES_PER_KMER
    return {"lines_per_kmer": LINES_PER_KMER, "achieved_lines_per_s": lines, "roof": "random_update_64",
            "roof_lines_per_s": roof, "frac": lines / roof}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=None,
                   help="timed steps (count map: full passes over all reads, default 5; other workloads: their own)")
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--reads", type=int, default=20_000_000)
    p.add_argument("--batch", type=int, default=65_536)
    p.add_argument("--genome", type=int, default=100_000_000)
    p.add_argument("--expected", type=int, default=100_000_000)
    p.add_argument("--lookup-reads", type=int, default=2_000_000)
    p.add_argument("--cpu-sample-reads", type=int, default=20_000_000)
    p.add_argument("--cpu-threads", type=int, default=None)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--served-requests", type=int, default=600,
                   help="count-map: requests of the served-latency probe on the built table (0: none)")
    p.add_argument("--blocks-per-cu", type=int, default=0)
    p.add_argument("--growth", choices=["auto", "in-place", "out-of-place"], default="auto",
                   help="how tables double (kache_amd_set_growth)")
    p.add_argument("--seed", type=int, default=None)
    p.add_argument("--workload", choices=["count-map", "serve", "counter", "batched"], default="count-map")
    p.add_argument("--batched-gbp", type=float, default=8.0)
    p.add_argument("--batched-genome", type=int, default=500_000_000)
    p.add_argument("--batched-expected", type=int, default=187_500_000)
    p.add_argument("--cpu-sample-bytes", type=int, default=1 << 30)
    p.add_argument("--counter-genome", type=int, default=950_000_000)
    p.add_argument("--counter-read-gbp", type=float, default=30.0)
    p.add_argument("--requests", type=int, default=2000, help="serve: requests per run")
    p.add_argument("--serve-genome", type=int, default=2_000_000_000)
    p.add_argument("--serve-expected", type=int, default=2_000_000_000)
    p.add_argument("--cpu-requests", type=int, default=600)
    p.add_argument("--cpu-genome", type=int, default=200_000_000)
    return p.parse_args()


def gen_text(api, capi, n_reads, genome_seed, genome_len, seed, sub=0.005, nrate=0.0005, lower=0.5):
    """n_reads reads as a device text (each read followed by '\\n'), padded."""
    L1 = READ_LEN + 1
    buf = capi.DeviceBuffer(api, n_reads * L1 + capi.TEXT_PAD)
    api.check(api.lib.kache_amd_synth_reads(buf.ptr, n_reads, READ_LEN, genome_seed, genome_len, seed, sub, nrate, lower), "synth_reads")
    api.check(api.lib.kache_amd_device_fill(buf.ptr + n_reads * L1, ord("\n"), capi.TEXT_PAD), "fill")
    return buf


def valid_positions(host_rows):
    """Number of 31-mer positions without a non-ACGT character, per read."""
    bad = ~np.isin(host_rows | 0x20, np.frombuffer(b"acgt", np.uint8))
    cs = np.concatenate([np.zeros((len(host_rows), 1), np.int32), np.cumsum(bad, axis=1, dtype=np.int32)], axis=1)
    win = cs[:, 31:] - cs[:, :-31]
    return (win == 0).sum(axis=1)


def host_cores():
    """Threads for the CPU baselines: every core this process may run on (its
    affinity set, within the cgroup's CPU quota when one is set), at most the
    reference's 128 users (include/Streaming_Kmer_Hash_Table.hpp:59); with the
    figures they come from."""
    cpus = os.cpu_count() or 1
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = cpus
    quota = None
    try:
        q, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = float(q) / float(period)
    except (OSError, ValueError):
        pass
    usable = avail if quota is None else max(1, min(avail, int(quota)))
    return min(128, usable), {"host_cpus": cpus, "affinity_cpus": avail, "cgroup_cpu_quota": quota}


def cpu_baseline(host_text, n_reads, expected, threads, look_text=None, look_reads=0):
    """The reference's C ABI (oracle/_ref, built from its sources) inserting
    the same reads from `threads` threads, each its own user, in calls of
    1,024 reads (timed over the inserts only); then, on the table built,
    kache_count of the lookup reads from the same threads in calls of 1,024
    reads (timed over the lookups only)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from kache_hash_amd import capi
    path = O.ref_library()
    if path is None:
        return None
    api = capi.Api(path)
    t = capi.Table(api, 8, expected, 0.8, False, threads)
    users = [t.register() for _ in range(threads)]
    L1 = READ_LEN + 1
    per_call = 1024
    lens = np.full(per_call, READ_LEN, dtype=np.uint64)

    def calls_of(text, n):
        base = text.ctypes.data
        calls = [(a, min(per_call, n - a)) for a in range(0, n, per_call)]
        return calls, [(np.arange(c, dtype=np.uint64) * L1 + np.uint64(base + a * L1)).astype(np.uint64) for a, c in calls]

    def parallel(work):
        th = [threading.Thread(target=work, args=(i,)) for i in range(threads)]
        t0 = time.perf_counter()
        for x in th:
            x.start()
        for x in th:
            x.join()
        return time.perf_counter() - t0

    calls, ptr_tabs = calls_of(host_text, n_reads)
    totals = [0] * threads

    def insert(i):
        n = C.c_uint64()
        for j in range(i, len(calls), threads):
            st = api.lib.kache_insert(t.h, users[i], ptr_tabs[j].ctypes.data, lens.ctypes.data, calls[j][1], capi.ADD, 1, C.byref(n))
            assert st == 0
            totals[i] += n.value

    dt = parallel(insert)
    rows = host_text[: n_reads * L1].reshape(n_reads, L1)[:, :READ_LEN]
    kmers = int(valid_positions(rows).sum())
    size = t.size()
    out = {"value": kmers / dt, "unit": "k-mer upserts/s", "cores": threads, "kind": "reference",
           "sample": f"all {n_reads} reads ({kmers} k-mers, {size} distinct) upserted into a fresh 100M-key u64 map "
                     f"by {threads} threads, one user each, in calls of {per_call} reads; {os.path.basename(path)}",
           "seconds": dt, "resizes_capacity": t.stats()["capacity"]}
    if look_text is not None and look_reads:
        lcalls, lptrs = calls_of(look_text, look_reads)
        hits = [0] * threads

        def lookup(i):
            cnt = np.zeros(per_call, np.uint64)
            for j in range(i, len(lcalls), threads):
                st = api.lib.kache_count(t.h, lptrs[j].ctypes.data, lens.ctypes.data, lcalls[j][1], cnt.ctypes.data)
                assert st == 0
                hits[i] += int(cnt[:lcalls[j][1]].sum())

        dtl = parallel(lookup)
        lk = int(valid_positions(look_text[: look_reads * L1].reshape(look_reads, L1)[:, :READ_LEN]).sum())
        out.update(lookups_per_s=lk / dtl, lookup_kmers=lk, lookup_hits=sum(hits), lookup_seconds=dtl,
                   lookup_sample=f"kache_count of {look_reads} fresh reads ({lk} k-mers) by {threads} threads in calls of "
                                 f"{per_call} reads, on the table built")
    t.close()
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}): there is no multi-GPU path; the table "
              "would shard by contiguous bucket ranges, one all-to-all of super-k-mers per batch", file=sys.stderr)
        sys.exit(2)

    import kache_hash_amd as KA
    from kache_hash_amd import capi

    api = KA.api()
    if a.workload == "serve":
        return serve_main(a, KA, capi, api)
    if a.workload == "counter":
        return counter_main(a, KA, capi, api)
    if a.workload == "batched":
        return batched_main(a, KA, capi, api)
    if a.seed is None:
        a.seed = 0xC2
    steps = a.steps if a.steps is not None else 5
    warmup = a.warmup
    n_reads = a.reads
    t_gen = time.perf_counter()
    # Reads come straight from the seeded genome's hash (the genome itself is
    # never materialized: kache_amd_synth_reads derives each window from it).
    text = gen_text(api, capi, n_reads, a.seed, a.genome, a.seed + 1)
    look = gen_text(api, capi, a.lookup_reads, a.seed, a.genome, a.seed + 2)
    print(f"bench: generated {n_reads} reads + {a.lookup_reads} lookup reads in {time.perf_counter() - t_gen:.1f}s", file=sys.stderr, flush=True)

    L1 = READ_LEN + 1
    base = text.ptr
    # Batch bounds of the whole run: 65,536 reads each (the last one shorter).
    edges = np.array([min(r, n_reads) * L1 for r in range(0, n_reads + a.batch, a.batch)], np.uint64)
    edges = np.unique(edges)
    n_batches = len(edges) - 1

    def one_pass(profile):
        """One step: a fresh table sized for the expected keys, every batch
        upserted (+1, new keys at 1) as one pipelined call; returns the table."""
        t = KA.table(width=8, expected_keys=a.expected)
        t.set_growth(GROWTH[a.growth])
        if a.blocks_per_cu:
            t.set_blocks_per_cu(a.blocks_per_cu)
        t.profile(profile)
        u = t.register()
        t.insert_batches(base, edges, KA.ADD, 1, u)
        return t, u

    # Warm-up passes (untimed); the last one also counts the k-mer positions
    # of the reads (every one is present after the inserts).
    kmers_per_pass = None
    for i in range(max(warmup, 1)):
        t, u = one_pass(False)
        if kmers_per_pass is None:
            kmers_per_pass = t.count_device(base, n_reads * L1, positions=True)[1]
        t.close()
    print(f"bench: {n_batches} batches, {kmers_per_pass} k-mer positions per pass; timing {steps} passes", file=sys.stderr, flush=True)

    # Kernel times come from HIP events bracketing every launch of the profiled
    # passes: every PROFILE_EVERY-th timed pass, the first and the last
    # included.  Events around all 306 launches of a pass cost ~3.5% of it.
    ins_ms = split_ms = re_ms = 0.0
    ins_launches = profiled = 0
    api.check(api.lib.kache_amd_device_sync(), "sync")
    t0 = time.perf_counter()
    for i in range(steps):
        prof = i % PROFILE_EVERY == 0 or i + 1 == steps
        t, u = one_pass(prof)
        if prof:
            m, n = t.kernel_time("insert")
            ins_ms, ins_launches = ins_ms + m, ins_launches + n
            split_ms += t.kernel_time("split")[0]
            re_ms += t.kernel_time("reinsert")[0]
            profiled += 1
        if i + 1 < steps:
            t.close()
    api.check(api.lib.kache_amd_device_sync(), "sync")
    dt = time.perf_counter() - t0    # each call returns after its device work
    t.profile(False)
    info = t.info()
    stats = t.stats()
    timed_kmers = kmers_per_pass * steps

    # Lookups: 2M fresh reads, in batches of 65,536, back to back, on the
    # last pass's table.  LOOKUP_PASSES timed passes without kernel timing
    # events (the median pass is reported), then one profiled pass for the
    # query kernel's own time.
    lb = np.unique(np.array([min(r, a.lookup_reads) * L1 for r in range(0, a.lookup_reads + a.batch, a.batch)], np.uint64))
    t.count_batches(look.ptr, lb)       # untimed warm-up of the query path
    look_dts = []
    for _ in range(LOOKUP_PASSES):
        t1 = time.perf_counter()
        h, p = t.count_batches(look.ptr, lb)
        look_dts.append(time.perf_counter() - t1)
    dt_look = float(np.median(look_dts))
    hits, look_kmers = int(h.sum()), int(p.sum())
    t.profile(True)
    h2, _ = t.count_batches(look.ptr, lb)
    q_ms, q_launches = t.kernel_time("query")
    t.profile(False)
    assert int(h2.sum()) == hits, "lookup passes disagree"

    per_launch_kmers = kmers_per_pass * profiled / max(ins_launches, 1)
    avg_ms = ins_ms / max(ins_launches, 1)
    achieved = per_launch_kmers * ALG_BYTES_PER_KMER / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    traffic = None
    tf = os.path.join(ROOT, "profiles", "insert_count_map_traffic.json")
    if os.path.exists(tf):
        try:
            traffic = json.load(open(tf)).get("bytes_per_launch")
        except Exception:
            traffic = None

    served = served_probe(a, KA, capi, api, t, u) if a.served_requests > 0 else None
    t.close()

    cpu = None
    if not a.no_cpu:
        threads, cores = host_cores()
        if a.cpu_threads:
            threads = a.cpu_threads
        m = min(a.cpu_sample_reads, n_reads)
        host = text.download(m * L1)
        lhost = look.download(a.lookup_reads * L1)
        print(f"bench: reference CPU path on {m} reads + {a.lookup_reads} lookup reads, {threads} threads ({cores})",
              file=sys.stderr, flush=True)
        cpu = cpu_baseline(host, m, a.expected, threads, lhost, a.lookup_reads)
        if cpu:
            cpu.update(cores)

    value = timed_kmers / dt
    out = {
        "metric": "k-mer upserts/s + lookups/s (k=31 count map); p50/p99 served-batch latency",
        "value": value, "unit": "k-mer upserts/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": dt * 1e3 / steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u64", "data": "synthetic (seeded on device)",
        "config": {"workload": "count map <31,u64,22>: 100 Mbp genome, 20M 150-bp reads (0.5% subst, 0.05% N, mixed case), "
                               "upsert +1 in batches of 65,536 reads into a fresh 100M-key table per step (all doublings "
                               "included), lookups of 2M fresh reads",
                   "step": "one full pass: table creation, every batch, the doublings, table release",
                   "reads": n_reads, "batches": n_batches, "batch_reads": a.batch, "expected_keys": a.expected,
                   "genome_bp": a.genome, "growth": a.growth},
        "lookups_per_s": look_kmers / dt_look,
        "lookup_hits": hits, "lookup_kmers": look_kmers, "lookup_pass_ms": [x * 1e3 for x in look_dts],
        "timed_kmers": timed_kmers, "kmers_per_step": kmers_per_pass,
        "keys": stats["keys"], "capacity": stats["capacity"], "resizes": info["resizes"],
        "in_place_doublings": info["in_place_doublings"], "store_mapped_bytes": info["store_mapped_bytes"],
        "last_doubling_away_keys": info["last_away_keys"], "away_list_regrows": info["away_list_regrows"],
        "placement_pct": [stats["primary_pct"], stats["secondary_pct"], stats["overflow_pct"]],
        "kernel_ms": {"insert": ins_ms, "split": split_ms, "reinsert": re_ms, "query": q_ms,
                      "insert_launches": ins_launches, "query_launches": q_launches, "profiled_steps": profiled},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "traffic": traffic,
                     "kernel": "k_insert<8, ADD> (8-byte values)", "alg_bytes_per_kmer": ALG_BYTES_PER_KMER,
                     "kmers_per_launch": per_launch_kmers, "avg_launch_ms": avg_ms,
                     "random_line_roof": random_line_roof(per_launch_kmers / (avg_ms * 1e-3) if avg_ms > 0 else 0.0),
                     "transaction_roof": transaction_roof(per_launch_kmers / (avg_ms * 1e-3) if avg_ms > 0 else 0.0)},
        "cpu_baseline": cpu,
        "served": served,
    }
    if cpu:
        out["speedup_vs_cpu"] = value / cpu["value"]
        if cpu.get("lookups_per_s"):
            out["lookup_speedup_vs_cpu"] = out["lookups_per_s"] / cpu["lookups_per_s"]
    print(json.dumps(out), flush=True)


def served_probe(a, KA, capi, api, t, u):
    """Served-batch latency on the count map just built: requests of the
    serving workload's mix and shapes drawn from the same genome (count,
    lookup, update +1, upsert +1), run by the serving loop and by plain
    per-request calls, closed loop and at Poisson arrivals of a quarter, half
    and three quarters of the plain path's capacity (four alternating runs of
    each per rate, latencies pooled: a run's p99 is its sixth-slowest of 600
    requests and moves by a few percent from run to run).  After the timed region; its upserts
    change the table."""
    g = capi.DeviceBuffer(api, a.genome)
    api.check(api.lib.kache_amd_synth_genome(g.ptr, a.genome, a.seed), "synth_genome")
    genome = g.download(a.genome)
    g.free()
    reqs = serve_requests(capi, genome, a.served_requests, a.seed + 5, a.genome)
    # Output arrays in page-locked memory (kache_amd_host_alloc), as a server's
    # response buffers would be: lookups are copied into them by DMA directly.
    made = [capi.make_request(8, k, b, capi.ADD, 1, api=api) for k, b in reqs]
    for m in made:
        for key, arr in m[1].items():
            if not key.startswith("_"):
                arr.fill(1)
    WAVE = 64 << 20
    serve_run(capi, t, u, reqs, made, np.zeros(len(reqs)), WAVE, 1 << 20)     # warm-up
    ro = [m for m, (k, _) in zip(made, reqs) if k in (capi.REQ_COUNT, capi.REQ_LOOKUP)]
    cap_s = serve_run(capi, t, u, reqs, made, np.zeros(len(reqs)), WAVE, 1 << 20, ro)
    cap_p = serve_run(capi, t, u, reqs, made, np.zeros(len(reqs)), WAVE, 1, ro)
    rng = np.random.default_rng(a.seed + 6)
    rows = []
    for i, frac in enumerate((0.25, 0.5, 0.75)):
        rate = frac * cap_p["requests_per_s"]
        arr = np.cumsum(rng.exponential(1.0 / rate, len(reqs)))
        if i == 0:
            serve_run(capi, t, u, reqs, made, arr, WAVE, 1 << 20, ro)       # untimed open-loop warm-up of each
            serve_run(capi, t, u, reqs, made, arr, WAVE, 1, ro)
        runs_s, runs_p = [], []
        for _ in range(4):      # alternating, on the same arrivals; latencies pooled
            runs_s.append(serve_run(capi, t, u, reqs, made, arr, WAVE, 1 << 20, ro))
            runs_p.append(serve_run(capi, t, u, reqs, made, arr, WAVE, 1, ro))
        at_s, at_p = pooled(runs_s), pooled(runs_p)
        print(f"bench: served {cap_s['requests_per_s']:.0f} req/s vs plain {cap_p['requests_per_s']:.0f} req/s; at {rate:.0f} req/s "
              f"served p50 {at_s['p50_ms']:.3f} p99 {at_s['p99_ms']:.3f} ms, plain p50 {at_p['p50_ms']:.3f} p99 {at_p['p99_ms']:.3f} ms",
              file=sys.stderr, flush=True)
        rows.append({"rate_req_per_s": rate, "of_plain_capacity": frac, "served_p50_ms": at_s["p50_ms"],
                     "served_p99_ms": at_s["p99_ms"], "plain_p50_ms": at_p["p50_ms"], "plain_p99_ms": at_p["p99_ms"],
                     "served_p99_ms_per_run": at_s["p99_ms_per_run"], "plain_p99_ms_per_run": at_p["p99_ms_per_run"],
                     "served_req_per_s": at_s["requests_per_s"], "plain_req_per_s": at_p["requests_per_s"]})
    return {"requests": len(reqs), "kmer_positions": sum(b.positions() for _, b in reqs),
            "mix": "35% count, 35% lookup, 5% update +1, 25% upsert +1; 1-10,000 sequences, lognormal median 300 bp",
            "closed_loop": {"served_req_per_s": cap_s["requests_per_s"], "served_kmers_per_s": cap_s["kmers_per_s"],
                            "plain_req_per_s": cap_p["requests_per_s"], "plain_kmers_per_s": cap_p["kmers_per_s"],
                            "served_p50_ms": cap_s["p50_ms"], "served_p99_ms": cap_s["p99_ms"]},
            "poisson": rows}


# --- Serving workload ------------------------------------------------------

SERVE_MIX = (0.35, 0.35, 0.05, 0.25)     # count, lookup, update (+1 if present), upsert (+1, new keys start at 1)
_COMP = np.arange(256, dtype=np.uint8)
for _a, _b in (b"AT", b"TA", b"CG", b"GC"):
    _COMP[_a] = _b
_CODE = np.zeros(256, np.uint8)
for _i, _c in enumerate(b"ACGT"):
    _CODE[_c] = _i
_ACGT = np.frombuffer(b"ACGT", np.uint8)


def serve_requests(capi, genome, n, seed, region):
    """n requests: kind by SERVE_MIX; 1..10,000 sequences (log-uniform), of
    lengths ~ lognormal(median 300, sigma 1) clamped to [1, 50,000], from
    uniform starts in genome[:region], random strand, 1% substitutions.
    Returns [(kind, batch)]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        kind = int(rng.choice(4, p=SERVE_MIX))
        m = min(int(np.exp(rng.uniform(0, np.log(10_001)))), 10_000)
        lens = np.clip(np.rint(rng.lognormal(np.log(300), 1.0, m)), 1, 50_000).astype(np.int64)
        starts = rng.integers(0, region - lens)
        rc = np.repeat(rng.random(m) < 0.5, lens)
        cum = np.zeros(m + 1, np.int64)
        np.cumsum(lens, out=cum[1:])
        j = np.arange(cum[-1]) - np.repeat(cum[:-1], lens)
        L = np.repeat(lens, lens)
        src = np.repeat(starts, lens) + np.where(rc, L - 1 - j, j)
        b = genome[src]
        b = np.where(rc, _COMP[b], b)
        sub = np.flatnonzero(rng.random(len(b)) < 0.01)
        b[sub] = _ACGT[(_CODE[b[sub]] + rng.integers(1, 4, len(sub))) % 4]
        out.append((kind, capi.Batch.from_buffer(np.ascontiguousarray(b, np.uint8), cum[:-1], lens)))
    return out


def _pct(x, q):
    return float(np.percentile(np.asarray(x), q)) if len(x) else None


def serve_run(capi, table, user, reqs, made, arrivals, wave_bytes, max_wave, warm=None):
    """One server over the table, fed `reqs` (as `made` requests) at
    `arrivals`; summary dict.  `warm` (made read-only requests) run first on
    the same server, closed loop and untimed: a new server's first requests
    pay one-time costs (its threads' first calls) that are not its steady
    state.  Wave figures count the timed requests' waves only."""
    with capi.Server(table, user, wave_bytes, max_wave) as s:
        if warm:
            s.replay(warm, np.zeros(len(warm)))
        st0 = s.stats()
        res = s.replay(made, arrivals)
        st = s.stats()
    bad = [r for r in res if r["status"] != 0]
    if bad:
        raise RuntimeError(f"served request failed: {bad[0]}")
    lat = [r["done_s"] - r["submit_s"] for r in res]
    span = max(r["done_s"] for r in res) - min(r["submit_s"] for r in res)
    pos = sum(b.positions() for _, b in reqs)
    sizes = collections.Counter(r["wave"] for r in res)
    return {"kmers_per_s": pos / span, "requests_per_s": len(reqs) / span, "p50_ms": _pct(lat, 50) * 1e3,
            "p99_ms": _pct(lat, 99) * 1e3, "mean_ms": float(np.mean(lat)) * 1e3, "waves": st["waves"] - st0["waves"],
            "direct_waves": st["direct_waves"] - st0["direct_waves"], "max_wave_requests": max(sizes.values()),
            "seconds": span, "_lat": lat, "_pos": pos}


def pooled(runs):
    """Runs of the same requests and arrivals as one sample: latencies
    pooled, rates over the summed spans."""
    lat = [x for r in runs for x in r["_lat"]]
    span = sum(r["seconds"] for r in runs)
    return {"kmers_per_s": sum(r["_pos"] for r in runs) / span, "requests_per_s": len(lat) / span,
            "p50_ms": _pct(lat, 50) * 1e3, "p99_ms": _pct(lat, 99) * 1e3, "mean_ms": float(np.mean(lat)) * 1e3,
            "waves": sum(r["waves"] for r in runs), "direct_waves": sum(r["direct_waves"] for r in runs),
            "max_wave_requests": max(r["max_wave_requests"] for r in runs), "seconds": span, "runs": len(runs),
            "p99_ms_per_run": [r["p99_ms"] for r in runs]}


def public(summary):
    return {k: v for k, v in summary.items() if not k.startswith("_")}


def cpu_serve(capi, reqs, genome, region, threads):
    """The reference C ABI (oracle/_ref) serving requests from `threads`
    workers under its threading contract: a request starts only when no
    request of another class (reads, updates, upserts) runs.
    Table: u32 map of genome[:region] (+1 per position), sized 1.3x.  Closed loop: all
    requests queued at the start, run phased by class and in arrival order."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    path = O.ref_library()
    if path is None:
        return None
    api = capi.Api(path)
    t = capi.Table(api, 4, int(region * 1.3), 0.8, False, threads)    # headroom: no doubling while timed
    users = [t.register() for _ in range(threads)]
    # Prefill from all workers: the region in pieces of 1 Mbp (overlapping by k - 1).
    piece = 1 << 20
    starts = list(range(0, region, piece - 30))

    def prefill(w):
        n = C.c_uint64()
        for i in range(w, len(starts), threads):
            a0 = starts[i]
            seg = np.ascontiguousarray(genome[a0:min(region, a0 + piece)])
            ptr = np.array([seg.ctypes.data], np.uint64)
            ln = np.array([len(seg)], np.uint64)
            assert api.lib.kache_insert(t.h, users[w], ptr.ctypes.data, ln.ctypes.data, 1, capi.ADD, 1, C.byref(n)) == 0

    th = [threading.Thread(target=prefill, args=(w,)) for w in range(threads)]
    for x in th:
        x.start()
    for x in th:
        x.join()
    cls = [0 if k in (capi.REQ_COUNT, capi.REQ_LOOKUP) else (1 if k == capi.REQ_UPDATE else 2) for k, _ in reqs]

    def run(order):
        cv = threading.Condition()
        state = {"next": 0, "cls": None, "inflight": 0}

        def worker(w):
            n = C.c_uint64()
            while True:
                with cv:
                    while True:
                        if state["next"] >= len(order):
                            return
                        i = order[state["next"]]
                        if state["inflight"] == 0 or state["cls"] == cls[i]:
                            break
                        cv.wait()
                    state["next"] += 1
                    state["cls"] = cls[i]
                    state["inflight"] += 1
                k, b = reqs[i]
                if k == capi.REQ_COUNT:
                    out = np.zeros(max(b.n, 1), np.uint64)
                    st = api.lib.kache_count(t.h, b.c_seqs, b.c_lens, b.n, out.ctypes.data)
                elif k == capi.REQ_LOOKUP:
                    p = max(b.positions(), 1)
                    f, v, fw = np.zeros(p, np.uint8), np.zeros(4 * p, np.uint8), np.zeros(p, np.uint8)
                    st = api.lib.kache_lookup(t.h, b.c_seqs, b.c_lens, b.n, f.ctypes.data, v.ctypes.data, fw.ctypes.data)
                elif k == capi.REQ_UPDATE:
                    st = api.lib.kache_update(t.h, b.c_seqs, b.c_lens, b.n, capi.ADD, 1, C.byref(n))
                else:
                    st = api.lib.kache_insert(t.h, users[w], b.c_seqs, b.c_lens, b.n, capi.ADD, 1, C.byref(n))
                assert st == 0
                with cv:
                    state["inflight"] -= 1
                    cv.notify_all()

        th = [threading.Thread(target=worker, args=(w,)) for w in range(threads)]
        t0 = time.perf_counter()
        for x in th:
            x.start()
        for x in th:
            x.join()
        return time.perf_counter() - t0

    pos = sum(b.positions() for _, b in reqs)
    # Arrival order: the plain calls, as concurrent as the contract allows.
    dt_plain = run(list(range(len(reqs))))
    # Phased: all reads, then all updates, then all upserts (the reordering a
    # served form would make; the CPU path's best case).
    dt = run(sorted(range(len(reqs)), key=lambda i: cls[i]))
    t.close()
    return {"value": pos / dt, "unit": "k-mer operations/s", "cores": threads, "kind": "reference",
            "requests_per_s": len(reqs) / dt, "arrival_order_kmers_per_s": pos / dt_plain,
            "sample": f"{len(reqs)} requests of the same mix and shapes from the first {region // 1_000_000} Mbp of the genome, "
                      f"run closed-loop by {threads} threads through the reference C ABI, phased (reads, then updates, then "
                      f"upserts; classes never overlap), against a resident u32 map of those {region // 1_000_000} Mbp "
                      f"(far beyond the caches); arrival_order_kmers_per_s: the same requests in arrival order; "
                      f"{os.path.basename(path)}",
            "seconds": dt}


def serve_main(a, KA, capi, api):
    """Serving: a resident u32 count map of a 2 Gbp genome (sized for 2G keys
    exactly); requests arrive as a Poisson process; served (waves) vs plain
    per-request calls (one request per wave) at the same arrival rates."""
    seed = 0xC5 if a.seed is None else a.seed
    t_gen = time.perf_counter()
    n = a.serve_genome
    g = capi.DeviceBuffer(api, n + 1 + capi.TEXT_PAD)
    api.check(api.lib.kache_amd_synth_genome(g.ptr, n, seed), "synth_genome")
    api.check(api.lib.kache_amd_device_fill(g.ptr + n, ord("\n"), 1 + capi.TEXT_PAD), "fill")
    t = KA.table(width=4, expected_keys=a.serve_expected)
    t.set_growth(GROWTH[a.growth])
    u = t.register()
    t.insert_device(g.ptr, n + 1, KA.ADD, 1, u)
    genome = g.download(n)
    g.free()
    print(f"serve: genome {n} bp resident as {t.size()} keys in {time.perf_counter() - t_gen:.1f}s", file=sys.stderr, flush=True)
    reqs = serve_requests(capi, genome, a.requests, seed + 1, n)
    warm = serve_requests(capi, genome, max(a.requests // 10, 50), seed + 2, n)
    pos = sum(b.positions() for _, b in reqs)
    print(f"serve: {len(reqs)} requests, {pos} k-mer positions", file=sys.stderr, flush=True)
    WAVE = 64 << 20
    # Requests (and their output arrays) are made once and reused by every run;
    # the outputs are page-locked (kache_amd_host_alloc), as a server's response
    # buffers would be, so lookups are copied into them by DMA directly.
    made = [capi.make_request(4, k, b, capi.ADD, 1, api=api) for k, b in reqs]
    wmade = [capi.make_request(4, k, b, capi.ADD, 1, api=api) for k, b in warm]
    for m in made + wmade:
        for key, arr in m[1].items():
            if not key.startswith("_"):
                arr.fill(1)
    serve_run(capi, t, u, warm, wmade, np.zeros(len(warm)), WAVE, 1 << 20)       # warm-up: the first upserts double the table
    serve_run(capi, t, u, warm, wmade, np.zeros(len(warm)), WAVE, 1)
    ro = [m for m, (k, _) in zip(wmade, warm) if k in (capi.REQ_COUNT, capi.REQ_LOOKUP)]
    cap_served = serve_run(capi, t, u, reqs, made, np.zeros(len(reqs)), WAVE, 1 << 20, ro)
    cap_plain = serve_run(capi, t, u, reqs, made, np.zeros(len(reqs)), WAVE, 1, ro)
    print(f"serve: capacity served {cap_served['requests_per_s']:.0f} req/s, plain {cap_plain['requests_per_s']:.0f} req/s",
          file=sys.stderr, flush=True)
    rng = np.random.default_rng(seed + 3)
    sweep = []
    fracs = (0.25, 0.5, 0.75, 0.9, 1.5, 2.5)
    arrivals = [np.cumsum(rng.exponential(1.0 / (f * cap_plain["requests_per_s"]), len(reqs))) for f in fracs]
    # Untimed, one of each: the first open-loop served run after the
    # closed-loop ones has shown a one-off stall of about a millisecond early
    # in the run.
    serve_run(capi, t, u, reqs, made, arrivals[0], WAVE, 1 << 20, ro)
    serve_run(capi, t, u, reqs, made, arrivals[0], WAVE, 1, ro)
    for frac, arr in zip(fracs, arrivals):
        rate = frac * cap_plain["requests_per_s"]
        row = {"rate_req_per_s": rate, "of_plain_capacity": frac}
        if frac <= 0.9:
            # Served and plain alternate, twice each, on the same arrivals.
            runs = {"served": [], "plain": []}
            for _ in range(2):
                runs["served"].append(serve_run(capi, t, u, reqs, made, arr, WAVE, 1 << 20, ro))
                runs["plain"].append(serve_run(capi, t, u, reqs, made, arr, WAVE, 1, ro))
            row["served"], row["plain"] = pooled(runs["served"]), pooled(runs["plain"])
        else:
            row["served"] = public(serve_run(capi, t, u, reqs, made, arr, WAVE, 1 << 20, ro))
        sweep.append(row)
        print(f"serve: rate {rate:.0f}/s served p50 {row['served']['p50_ms']:.2f} p99 {row['served']['p99_ms']:.2f} ms"
              + (f", plain p50 {row['plain']['p50_ms']:.2f} p99 {row['plain']['p99_ms']:.2f} ms" if "plain" in row else ""),
              file=sys.stderr, flush=True)
    info = t.info()
    cpu = None
    if not a.no_cpu:
        threads = a.cpu_threads or host_cores()[0]
        creqs = serve_requests(capi, genome, a.cpu_requests, seed + 4, a.cpu_genome)
        print(f"serve: reference CPU path, {len(creqs)} requests, {threads} threads", file=sys.stderr, flush=True)
        cpu = cpu_serve(capi, creqs, genome, a.cpu_genome, threads)
    out = {
        "metric": "k-mer upserts/s + lookups/s (k=31 count map); p50/p99 served-batch latency",
        "value": cap_served["kmers_per_s"], "unit": "k-mer operations/s (served, closed loop)", "n_gpus": 1,
        "steps": len(reqs), "warmup": len(warm), "ms_per_step": cap_served["seconds"] * 1e3 / len(reqs),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
        "data": "synthetic (seeded genome on device, requests on host)",
        "config": {"workload": "serving: resident <31,u32,22> map of a 2 Gbp genome (sized for 2G keys); requests "
                               "35% count / 35% lookup / 5% update +1 / 25% upsert +1, 1-10,000 sequences "
                               "(log-uniform), lengths lognormal median 300 (max 50k), 1% substitutions; Poisson arrivals",
                   "requests": len(reqs), "kmer_positions": pos, "wave_bytes": WAVE, "genome_bp": n},
        "p50_ms": cap_served["p50_ms"], "p99_ms": cap_served["p99_ms"],
        "capacity": {"served": public(cap_served), "plain_per_request": public(cap_plain)},
        "sweep": sweep,
        "table": {"keys": t.size(), "buckets": info["buckets"], "resizes": info["resizes"], "device_bytes": info["device_bytes"]},
        "cpu_baseline": cpu,
    }
    if cpu:
        out["speedup_vs_cpu"] = out["value"] / cpu["value"]
    print(json.dumps(out), flush=True)


# --- Counter workload -------------------------------------------------------

def counter_inputs(capi, api, genome_len, read_gbp, seed, donor_frac=0.6):
    """The counter's inputs, in host memory: segments covering a seeded
    genome (lengths uniform in [1, 50,000]) plus 5% more bases of segment
    variants (copies of random places with one substitution per 500 bases),
    newline-separated; and reads of 150 bp (0.5% substitutions, 0.05% N)
    from the first `donor_frac` of the genome, newline-separated."""
    rng = np.random.default_rng(seed)
    g = capi.DeviceBuffer(api, genome_len)
    api.check(api.lib.kache_amd_synth_genome(g.ptr, genome_len, seed), "synth_genome")
    genome = g.download(genome_len)
    g.free()
    cuts = np.cumsum(rng.integers(1, 50_001, genome_len // 25_000 + 10))
    cuts = cuts[cuts < genome_len]
    parts = np.split(genome, cuts)
    extra, want = [], genome_len // 20
    while want > 0:
        n = int(min(want, rng.integers(1, 50_001)))
        a0 = int(rng.integers(0, genome_len - n))
        piece = genome[a0:a0 + n].copy()
        sub = np.arange(int(rng.integers(0, 500)), n, 500)
        piece[sub] = _ACGT[(_CODE[piece[sub]] + 1) % 4]
        extra.append(piece)
        want -= n
    nl = np.frombuffer(b"\n", np.uint8)
    segs = parts + extra
    prime = np.concatenate([x for p in segs for x in (p, nl)])
    L1 = READ_LEN + 1
    n_reads = int(read_gbp * 1e9) // READ_LEN
    reads = np.empty(n_reads * L1, np.uint8)
    chunk = 10_000_000
    buf = capi.DeviceBuffer(api, chunk * L1)
    donor = int(genome_len * donor_frac)
    for r0 in range(0, n_reads, chunk):
        c = min(chunk, n_reads - r0)
        # Each chunk is its own seeded read set from the donor part of the genome.
        api.check(api.lib.kache_amd_synth_reads(buf.ptr, c, READ_LEN, seed, donor, seed + 1 + r0 // chunk, 0.005, 0.0005, 0.0), "synth_reads")
        reads[r0 * L1:(r0 + c) * L1] = buf.download(c * L1)
    buf.free()
    return prime, reads, len(segs)


def counter_cpu(capi, prime, reads, n_prime_bytes, n_read_bytes, threads):
    """The reference C ABI (oracle/_ref) as the reference's counter drives its
    table (counter/Kmer_Set_Counter.cpp:64-94): a pow2 u32 table at load
    0.9; prime = upsert KEEP 0 of the segments, count = update +1 of the
    first reads, from `threads` threads in batches of 1,024 sequences.
    Timed: the count.  Returns k-mer windows/s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    path = O.ref_library()
    if path is None:
        return None
    api = capi.Api(path)
    p = prime[:n_prime_bytes]
    p = p[:int(np.flatnonzero(p == 10)[-1]) + 1]
    t = capi.Table(api, 4, len(p), 0.9, True, threads)
    users = [t.register() for _ in range(threads)]

    def batches(text, per=1024):
        ends = np.flatnonzero(text == 10)
        starts = np.concatenate([[0], ends[:-1] + 1])
        out = []
        for i in range(0, len(ends), per):
            s, e = starts[i:i + per], ends[i:i + per]
            out.append(((s + np.uint64(text.ctypes.data)).astype(np.uint64), (e - s).astype(np.uint64)))
        return out

    def run(bs, fn):
        def work(w):
            n = C.c_uint64()
            for j in range(w, len(bs), threads):
                ptrs, lens = bs[j]
                assert fn(w, ptrs, lens, n) == 0
        th = [threading.Thread(target=work, args=(w,)) for w in range(threads)]
        t0 = time.perf_counter()
        for x in th:
            x.start()
        for x in th:
            x.join()
        return time.perf_counter() - t0

    run(batches(p), lambda w, ptrs, lens, n: api.lib.kache_insert(t.h, users[w], ptrs.ctypes.data, lens.ctypes.data, len(lens), capi.KEEP, 0, C.byref(n)))
    r = reads[:n_read_bytes]
    r = r[:int(np.flatnonzero(r == 10)[-1]) + 1]
    rb = batches(r)
    dt = run(rb, lambda w, ptrs, lens, n: api.lib.kache_update(t.h, ptrs.ctypes.data, lens.ctypes.data, len(lens), capi.ADD, 1, C.byref(n)))
    pos = sum(int(np.maximum(l.astype(np.int64) - 30, 0).sum()) for _, l in rb)
    size = t.size()
    t.close()
    return {"value": pos / dt, "unit": "k-mer count updates/s", "cores": threads, "kind": "reference",
            "sample": f"prime {len(p) >> 20} MiB of the segments ({size} k-mers), then count {len(r) >> 20} MiB of the reads "
                      f"({pos} positions) by {threads} threads in batches of 1,024 sequences; {os.path.basename(path)}",
            "seconds": dt}


def counter_main(a, KA, capi, api):
    """Set counting as the reference's Kmer_Set_Counter does it, on one GPU:
    prime a u32 set (pow2 table, load 0.9) with 1 Gbp of segments, count
    30 Gbp of reads streamed from host memory, histogram(10000)."""
    seed = 0xC4 if a.seed is None else a.seed
    t0 = time.perf_counter()
    prime, reads, n_segs = counter_inputs(capi, api, a.counter_genome, a.counter_read_gbp, seed)
    print(f"counter: {len(prime) >> 20} MiB of {n_segs} segments, {len(reads) >> 20} MiB of reads in host memory "
          f"({time.perf_counter() - t0:.1f}s)", file=sys.stderr, flush=True)
    t = KA.table(width=4, expected_keys=a.counter_genome, load_factor=0.9, pow2=True)
    t.set_growth(GROWTH[a.growth])
    u = t.register()
    t.profile(True)
    t1 = time.perf_counter()
    t.insert_text(prime, KA.KEEP, 0, u)
    dt_prime = time.perf_counter() - t1
    size = t.size()
    # Counting: one call per ~4 GiB of whole reads, as a reader hands them over.
    step = (4 << 30) // (READ_LEN + 1) * (READ_LEN + 1)
    found = 0
    t2 = time.perf_counter()
    for b0 in range(0, len(reads), step):
        b1 = min(len(reads), b0 + step)
        found += t.update_text(reads[b0:b1], KA.ADD_WRAP, 1)
    dt_count = time.perf_counter() - t2
    t3 = time.perf_counter()
    hist = t.histogram(10_000)
    dt_hist = time.perf_counter() - t3
    upd_ms, upd_launches = t.kernel_time("update")
    t.profile(False)
    # k-mer windows of the reads (those with an N, about 1.5%, are skipped by the update)
    positions = (len(reads) // (READ_LEN + 1)) * (READ_LEN - 30)
    h = hist.astype(np.float64)
    zero = size - int(hist.sum())
    checks = {"sum_c_hist_equals_found": int((np.arange(len(hist), dtype=np.uint64) * hist).sum()) == found,
              "keys_at_zero_pct": 100.0 * zero / max(size, 1)}
    stats = t.stats()
    cpu = None
    if not a.no_cpu:
        threads = a.cpu_threads or host_cores()[0]
        print(f"counter: reference CPU path, {threads} threads", file=sys.stderr, flush=True)
        cpu = counter_cpu(capi, prime, reads, len(prime), 1 << 30, threads)
    # The update kernel's roofline: 1.25 B of input + the 64-byte metadata line per
    # 5.29 k-mers + the 12-byte u32 slot, per k-mer position.
    alg = 150 / 120 + 64 / 5.29 + 12
    achieved = (positions / max(upd_launches, 1)) * alg / (upd_ms / max(upd_launches, 1) * 1e-3) / 1e9 if upd_ms > 0 else 0.0
    out = {
        "metric": "k-mer upserts/s + lookups/s (k=31 count map); p50/p99 served-batch latency",
        "value": positions / dt_count, "unit": "k-mer count updates/s", "n_gpus": 1,
        "steps": (len(reads) + step - 1) // step, "warmup": 0, "ms_per_step": dt_count * 1e3 / ((len(reads) + step - 1) // step),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
        "data": "synthetic (seeded genome and reads made on device, held in host memory)",
        "config": {"workload": "set counter <31,u32,22> (pow2, load 0.9): prime with segments of a 950 Mbp genome (lengths "
                               "U[1, 50k]) plus 5% variant segments; count 30 Gbp of 150-bp reads (0.5% subst, 0.05% N) "
                               "from the first 60% of the genome, streamed from host memory; histogram(10000)",
                   "genome_bp": a.counter_genome, "read_bytes": len(reads), "prime_bytes": len(prime)},
        "prime_kmers_per_s": (len(prime) - n_segs * 31) / dt_prime, "prime_seconds": dt_prime,
        "count_seconds": dt_count, "histogram_seconds": dt_hist, "found": found, "set_size": size,
        "capacity": stats["capacity"], "checks": checks,
        "host_to_device_GBps": len(reads) / dt_count / 1e9,
        "kernel_ms": {"update": upd_ms, "update_launches": upd_launches},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "traffic": None, "kernel": "k_update<ADD_WRAP>",
                     "alg_bytes_per_kmer": alg},
        "cpu_baseline": cpu,
    }
    if cpu:
        out["speedup_vs_cpu"] = out["value"] / cpu["value"]
    print(json.dumps(out), flush=True)


# --- Batched uneven sequences -------------------------------------------------

def batched_text(capi, api, gbp, genome_len, seed):
    """Sequences of lognormal lengths (median 2 kbp, sigma 1.3, clamped to
    [50, 100,000]) totalling `gbp` Gbp, from a seeded genome with 1%
    substitutions, 0.05% N and lowercase runs in 20% of them, as one device
    text.  Returns (buffer, text bytes, number of sequences)."""
    rng = np.random.default_rng(seed)
    target = int(gbp * 1e9)
    lens = []
    total = 0
    while total < target:
        l = np.clip(np.rint(rng.lognormal(np.log(2000), 1.3, 1 << 20)), 50, 100_000).astype(np.uint64)
        c = np.cumsum(l)
        k = int(np.searchsorted(c, target - total)) + 1
        lens.append(l[:k])
        total += int(c[min(k, len(c)) - 1])
    lens = np.concatenate(lens)
    n = len(lens)
    off = np.zeros(n + 1, np.uint64)
    np.cumsum(lens + 1, out=off[1:])
    text_bytes = int(off[-1])
    buf = capi.DeviceBuffer(api, text_bytes + capi.TEXT_PAD)
    d_off = capi.DeviceBuffer(api, 8 * (n + 1))
    d_off.upload(off.view(np.uint8))
    api.check(api.lib.kache_amd_synth_seqs(buf.ptr, d_off.ptr, n, int(lens.max()), seed, genome_len, seed + 1, 0.01, 0.0005, 0.2), "synth_seqs")
    api.check(api.lib.kache_amd_device_fill(buf.ptr + text_bytes, ord("\n"), capi.TEXT_PAD), "fill")
    d_off.free()
    return buf, text_bytes, n


def batched_cpu(capi, host, expected, threads, oriented):
    """The reference C ABI (oracle/_ref) inserting the same sequences (a
    prefix of the text) from `threads` threads, each its own user, in calls
    of about 1 Mbp (as the reference's reader batches them); the table
    starts at `expected` keys and doubles as it fills."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    path = O.ref_library()
    if path is None:
        return None
    api = capi.Api(path)
    t = capi.Table(api, 1 if oriented else 4, expected, 0.8, False, threads)
    users = [t.register() for _ in range(threads)]
    ends = np.flatnonzero(host == 10)
    starts = np.concatenate([[0], ends[:-1] + 1])
    calls, a = [], 0
    csum = np.cumsum(ends - starts)
    while a < len(ends):
        b = int(np.searchsorted(csum, (csum[a - 1] if a else 0) + (1 << 20))) + 1
        b = min(max(b, a + 1), len(ends))
        calls.append(((starts[a:b] + np.uint64(host.ctypes.data)).astype(np.uint64), (ends[a:b] - starts[a:b]).astype(np.uint64)))
        a = b

    def work(w):
        n = C.c_uint64()
        for j in range(w, len(calls), threads):
            ptrs, lens = calls[j]
            if oriented:
                st = api.lib.kache_insert_oriented(t.h, users[w], ptrs.ctypes.data, lens.ctypes.data, len(lens), capi.OR, 1, 2, C.byref(n))
            else:
                st = api.lib.kache_insert(t.h, users[w], ptrs.ctypes.data, lens.ctypes.data, len(lens), capi.ADD, 1, C.byref(n))
            assert st == 0

    th = [threading.Thread(target=work, args=(w,)) for w in range(threads)]
    t0 = time.perf_counter()
    for x in th:
        x.start()
    for x in th:
        x.join()
    dt = time.perf_counter() - t0
    pos = sum(int(np.maximum(l.astype(np.int64) - 30, 0).sum()) for _, l in calls)
    stats = t.stats()
    t.close()
    return {"value": pos / dt, "unit": "k-mer upserts/s", "cores": threads, "kind": "reference",
            "sample": f"the first {len(host) >> 20} MiB of the text ({pos} k-mer windows) inserted by {threads} threads in calls "
                      f"of ~1 Mbp into a {'u8 oriented OR' if oriented else 'u32 +1'} map starting at {expected} keys "
                      f"(capacity after: {stats['capacity']}); {os.path.basename(path)}",
            "seconds": dt}


def batched_main(a, KA, capi, api):
    """Uneven sequences in persistent launches with growth: a u32 count map
    (+1) and a u8 oriented map (OR 1 forward / 2 reverse), each from 1/16 of
    the final size, over 8 Gbp of lognormal-length sequences; the launch
    shape is chosen online by the autotuner."""
    seed = 0xC3 if a.seed is None else a.seed
    t0 = time.perf_counter()
    text, nbytes, nseq = batched_text(capi, api, a.batched_gbp, a.batched_genome, seed)
    print(f"batched: {nseq} sequences, {nbytes} bytes on the device ({time.perf_counter() - t0:.1f}s)", file=sys.stderr, flush=True)
    results = {}
    for name, width, oriented in (("u32_count", 4, False), ("u8_oriented", 1, True)):
        t = KA.table(width=width, expected_keys=a.batched_expected)
        t.set_growth(GROWTH[a.growth])
        u = t.register()
        t.autotune(2)
        t.profile(True)
        t1 = time.perf_counter()
        if oriented:
            t.insert_oriented_device(text.ptr, nbytes, KA.OR, 1, 2, u)
        else:
            t.insert_device(text.ptr, nbytes, KA.ADD, 1, u)
        dt = time.perf_counter() - t1
        ins_ms, ins_n = t.kernel_time("insert")
        split_ms, _ = t.kernel_time("split")
        re_ms, _ = t.kernel_time("reinsert")
        t.profile(False)
        _, kmers = t.count_device(text.ptr, nbytes, positions=True)
        info, stats = t.info(), t.stats()
        results[name] = {"kmers_per_s": kmers / dt, "seconds": dt, "kmers": kmers, "keys": stats["keys"],
                         "capacity": stats["capacity"], "resizes": info["resizes"], "device_bytes": info["device_bytes"],
                         "kernel_ms": {"insert": ins_ms, "insert_launches": ins_n, "split": split_ms, "reinsert": re_ms},
                         "tuning": t.tuning(), "trials": t.tuning_trials()}
        print(f"batched {name}: {kmers / dt / 1e9:.2f} G k-mers/s, {info['resizes']} doublings, tuned {t.tuning()}", file=sys.stderr, flush=True)
        t.close()
    cpu = None
    if not a.no_cpu:
        threads = a.cpu_threads or host_cores()[0]
        host = text.download(min(nbytes, a.cpu_sample_bytes))
        host = host[:int(np.flatnonzero(host == 10)[-1]) + 1]
        print(f"batched: reference CPU path on {len(host) >> 20} MiB, {threads} threads", file=sys.stderr, flush=True)
        cpu = batched_cpu(capi, host, max(len(host) // 16, 1 << 20), threads, False)
    u32 = results["u32_count"]
    alg = 1.0 + 64 / 5.49 + 12      # long sequences: input, the metadata line per minimizer run, the 12-byte u32 slot
    ms, n = u32["kernel_ms"]["insert"], u32["kernel_ms"]["insert_launches"]
    achieved = (u32["kmers"] / max(n, 1)) * alg / (ms / max(n, 1) * 1e-3) / 1e9 if ms > 0 else 0.0
    out = {
        "metric": "k-mer upserts/s + lookups/s (k=31 count map); p50/p99 served-batch latency",
        "value": u32["kmers_per_s"], "unit": "k-mer upserts/s (u32 count map, doublings included)", "n_gpus": 1,
        "steps": n, "warmup": 0, "ms_per_step": u32["seconds"] * 1e3 / max(n, 1), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u32", "data": "synthetic (seeded on device)",
        "config": {"workload": "batched uneven sizes with growth: 8 Gbp of lognormal-length sequences (median 2 kbp, "
                               "sigma 1.3, 50 bp-100 kbp) from a 500 Mbp genome, 1% subst; <31,u32,22> count map and "
                               "<31,u8,22> oriented OR map, each from 1/16 of the final size; autotuned launch shape",
                   "sequences": nseq, "text_bytes": nbytes, "expected_keys": a.batched_expected},
        "tables": results,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS,
                     "traffic": None, "kernel": "k_insert<4, ADD> (u32 values in 4-byte words)", "alg_bytes_per_kmer": alg},
        "cpu_baseline": cpu,
    }
    if cpu:
        out["speedup_vs_cpu"] = out["value"] / cpu["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
