"""Seeded synthetic sequences for the tests, shaped like the reference's test
inputs (test/capi.cpp:483-512, test/oriented.cpp:240-281): pieces of a
genome, some reverse-complemented, with lowercase runs and non-ACGT
characters, minimizer groups that overflow their buckets, and edge cases.
"""
import numpy as np

K, L = 31, 22
ACGT = np.frombuffer(b"ACGT", np.uint8)
_COMP = np.arange(256, dtype=np.uint8)
for a, b in (b"AT", b"TA", b"CG", b"GC", b"at", b"ta", b"cg", b"gc"):
    _COMP[a] = b


def genome(n, seed):
    rng = np.random.default_rng(seed)
    return ACGT[rng.integers(0, 4, n)].tobytes()


def revcomp(s):
    a = np.frombuffer(s, np.uint8)[::-1]
    return _COMP[a].tobytes()


def reads(g, n, length, seed, sub=0.01, nrate=0.001, lower=0.0):
    """n reads of `length` from genome g: uniform starts, random strand,
    substitutions at rate `sub`, N at rate `nrate`, and with probability
    `lower` a lowercase run."""
    rng = np.random.default_rng(seed)
    ga = np.frombuffer(g, np.uint8)
    starts = rng.integers(0, len(g) - length + 1, n)
    idx = starts[:, None] + np.arange(length)[None, :]
    r = ga[idx].copy()
    m = rng.random(r.shape) < sub
    codes = np.searchsorted(ACGT, r)
    r[m] = ACGT[(codes[m] + rng.integers(1, 4, m.sum())) % 4]
    r[rng.random(r.shape) < nrate] = ord("N")
    flip = rng.random(n) < 0.5
    r[flip] = _COMP[r[flip][:, ::-1]]
    if lower > 0:
        for i in np.nonzero(rng.random(n) < lower)[0]:
            a = rng.integers(0, length)
            b = min(length, a + rng.integers(1, length))
            r[i, a:b] |= 0x20
    return [bytes(x) for x in r]


def pieces(g, size=4096):
    """Pieces of `size` overlapping by k - 1 (test/capi.cpp:94-108)."""
    return [g[i:i + size] for i in range(0, len(g) - K + 1, size - (K - 1))]


def mixed_inserts(g, seed, max_positions=200_000):
    """test/capi.cpp:483-499: pieces, every third reverse-complemented, the
    reverse complement of every seventh beside it, and lowercase runs and
    non-ACGT characters in every eleventh."""
    S, pos = [], 0
    for i, p in enumerate(pieces(g, 1024)):
        if pos >= max_positions:
            break
        s = bytearray(revcomp(p) if i % 3 == 0 else p)
        if i % 11 == 5 and len(s) > 300:
            for j in range(100, 160):
                s[j] |= 0x20
            s[40], s[200], s[230], s[260] = ord("N"), ord("n"), ord("R"), ord("-")
        S.append(bytes(s))
        pos += max(len(s) - K + 1, 0)
        if i % 7 == 0:
            S.append(revcomp(bytes(s)))
    return S


def edge_queries(s0):
    return [b"", b"ACGT", b"N" * 40, s0[:K], s0[:K - 1], b"acgtnACGTACGTACGTACGTACGTACGTACGTACGTAC",
            b"A" * (K - 1), b"a" * K, b"NNNNACGTACGTACGTACGTACGTACGTACGTACGTNNNN"]


def _wang(x):
    M = (1 << 64) - 1
    x = ((~x) + (x << 21)) & M
    x ^= x >> 24
    x = (x + (x << 3) + (x << 8)) & M
    x ^= x >> 14
    x = (x + (x << 2) + (x << 4)) & M
    x ^= x >> 28
    return (x + (x << 31)) & M


def _lmer_hash(s):
    f = r = 0
    for i, ch in enumerate(s):
        b = ((ch >> 2) ^ (ch >> 1)) & 3
        f = (f << 2) | b
        r |= (3 - b) << (2 * i)
    return _wang(min(f, r))


def minimizer_groups(G, n, seed):
    """test/oriented.cpp:121-152: G groups of n distinct k-mers sharing a
    minimizer l-mer, so only 96 of a group fit in the main table."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(G):
        while True:
            X = ACGT[rng.integers(0, 4, L)].tobytes()
            if _lmer_hash(X) <= ((1 << 64) - 1) >> 10:
                break
        target = _lmer_hash(X)
        grp, seen = [], set()
        while len(grp) < n:
            s = ACGT[rng.integers(0, 4, K)].tobytes()
            o = int(rng.integers(0, K - L + 1))
            s = s[:o] + X + s[o + L:]
            if min(_lmer_hash(s[i:i + L]) for i in range(K - L + 1)) != target:
                continue
            if s in seen or revcomp(s) in seen:
                continue
            seen.add(s)
            grp.append(s)
        out.append(grp)
    return out


def ragged_batches(S, seed=0):
    """Batches of 1..13 sequences (test/capi.cpp:156-171)."""
    out, i, b = [], 0, 0
    while i < len(S):
        c = min(1 + b % 13, len(S) - i)
        out.append(S[i:i + c])
        i += c
        b += 1
    return out
