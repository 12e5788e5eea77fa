/* Oracle: plain-C, single-threaded restatement of the streaming k-mer table's
 * semantics at k = 31, l = 22.  TEST INFRASTRUCTURE ONLY: tests/, the smoke
 * check and bench.py's CPU-baseline leg may call it, as a checker; the product
 * never links it.  Each function cites the reference lines it restates. */
#include "kache_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define K 31
#define L 22
#define W (K - L + 1)
#define B 32
#define KMER_MASK ((UINT64_C(1) << (2 * K)) - 1)
#define LMER_MASK ((UINT64_C(1) << (2 * L)) - 1)
#define MAX_CAP ((UINT64_C(1) << 32) - 1)

/* include/Kmer_Batch.hpp:24-35 (Thomas Wang's 64-bit mix). */
uint64_t ko_wang_u64(uint64_t x)
{
    x = (~x) + (x << 21);
    x = x ^ (x >> 24);
    x = (x + (x << 3)) + (x << 8);
    x = x ^ (x >> 14);
    x = (x + (x << 2)) + (x << 4);
    x = x ^ (x >> 28);
    x = x + (x << 31);
    return x;
}

/* include/Kmer_Batch.hpp:44-67: the top byte of the checksum hash, 0 -> 1.
 * AVX2 builds multiply by 0x9E3779B97F4A7C15; AVX-512 builds take wang_u64. */
static uint8_t top_byte_nonzero(uint64_t x)
{
    const uint8_t c = (uint8_t)(x >> 56);
    return c ? c : 1;
}
uint8_t ko_checksum_mul(uint64_t key) { return top_byte_nonzero(key * UINT64_C(0x9E3779B97F4A7C15)); }
uint8_t ko_checksum_wang(uint64_t key) { return top_byte_nonzero(ko_wang_u64(key)); }

/* Reverse complement of a 31-mer in its 2-bit encoding, first base most
 * significant (include/Kmer_Lanes.hpp:258-263 does the same by pair, nibble
 * and byte swaps; here base by base). */
uint64_t ko_revcomp31(uint64_t x)
{
    uint64_t r = 0;
    for(int i = 0; i < K; ++i)
    {
        r = (r << 2) | (3 - (x & 3));
        x >>= 2;
    }
    return r;
}

/* include/Kmer_Lanes.hpp:247-274: min over the 10 l-mer offsets p of
 * wang(min(fwd l-mer at p, its reverse complement)). */
uint64_t ko_minimizer_hash(uint64_t key)
{
    const uint64_t r = ko_revcomp31(key);
    uint64_t m = ~UINT64_C(0);
    for(int p = 0; p < W; ++p)
    {
        const uint64_t lf = (key >> (2 * (K - L - p))) & LMER_MASK;
        const uint64_t lr = (r >> (2 * p)) & LMER_MASK;
        const uint64_t hv = ko_wang_u64(lf < lr ? lf : lr);
        if(hv < m)
            m = hv;
    }
    return m;
}

/* include/xxHash/xxhash.h:4273-4285 (XXH3_len_4to8_64b at len 8) and
 * :4206-4213 (XXH3_rrmxmx), secret bytes 8..23 of XXH3_kSecret (:3975). */
static uint64_t rotl64(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }
uint64_t ko_xxh3_8(uint64_t x, uint64_t seed)
{
    static const uint8_t secret_8_24[16] = {0x7c, 0x01, 0x81, 0x2c, 0xf7, 0x21, 0xad, 0x1c,
                                            0xde, 0xd4, 0x6d, 0xe9, 0x83, 0x90, 0x97, 0xdb};
    uint64_t s1 = 0, s2 = 0;
    for(int i = 7; i >= 0; --i)
        s1 = (s1 << 8) | secret_8_24[i], s2 = (s2 << 8) | secret_8_24[8 + i];

    const uint32_t lo = (uint32_t)seed;
    const uint32_t swapped = (lo >> 24) | ((lo >> 8) & 0xFF00u) | ((lo << 8) & 0xFF0000u) | (lo << 24);
    seed ^= (uint64_t)swapped << 32;
    const uint64_t input1 = (uint32_t)x, input2 = x >> 32;    /* little-endian 32-bit reads */
    const uint64_t bitflip = (s1 ^ s2) - seed;
    const uint64_t input64 = input2 + (input1 << 32);
    uint64_t h = input64 ^ bitflip;
    const uint64_t prime_mx2 = UINT64_C(0x9FB21C651E98DF25);
    h ^= rotl64(h, 49) ^ rotl64(h, 24);
    h *= prime_mx2;
    h ^= (h >> 35) + 8;
    h *= prime_mx2;
    return h ^ (h >> 28);
}

/* include/DNA_Utility.hpp:199-224: a character is a base iff its lower-case
 * form is one of acgt; include/Kmer_Batch.hpp:108: its 2-bit code. */
static int is_acgt(char ch)
{
    const char x = (char)(ch | 0x20);
    return x == 'a' || x == 'c' || x == 'g' || x == 't';
}
static uint64_t code_of(char ch) { return (uint64_t)(((ch >> 2) ^ (ch >> 1)) & 3); }

/* Describers (include/Kmer_Batch.hpp:191-250), restated per position by brute
 * force as test/describers.cpp:54-104 checks them. */
size_t ko_describe(const char* s, size_t len, int wang, uint8_t* valid, uint64_t* key, uint64_t* h, uint8_t* c, uint8_t* fwd)
{
    if(len < K)
        return 0;

    const size_t n = len - (K - 1);
    size_t run = 0; /* ACGT characters ending at the current one */
    uint64_t f = 0, r = 0;
    for(size_t j = 0; j < len; ++j)
    {
        if(!is_acgt(s[j]))
            run = 0;
        else
        {
            const uint64_t b = code_of(s[j]);
            f = ((f << 2) | b) & KMER_MASK;
            r = (r >> 2) | ((3 - b) << (2 * (K - 1)));
            run++;
        }

        if(j + 1 >= K)
        {
            const size_t i = j + 1 - K;
            valid[i] = (run >= K);
            if(run >= K)
            {
                const uint64_t x = f < r ? f : r;
                key[i] = x;
                fwd[i] = (f <= r);
                c[i] = wang ? ko_checksum_wang(x) : ko_checksum_mul(x);
                /* Minimizer hash straight from the characters. */
                uint64_t m = ~UINT64_C(0);
                for(int o = 0; o < W; ++o)
                {
                    uint64_t lf = 0, lr = 0;
                    for(int t = 0; t < L; ++t)
                    {
                        const uint64_t bb = code_of(s[i + o + t]);
                        lf = (lf << 2) | bb;
                        lr |= (3 - bb) << (2 * t);
                    }
                    const uint64_t hv = ko_wang_u64(lf < lr ? lf : lr);
                    if(hv < m)
                        m = hv;
                }
                h[i] = m;
            }
            else
                key[i] = 0, h[i] = 0, c[i] = 0, fwd[i] = 0;
        }
    }

    return n;
}

/* include/utility.hpp:167-182. */
static uint64_t ceil_pow2(uint64_t x)
{
    uint64_t p = 1;
    while(p < x)
        p <<= 1;
    return p;
}

/* include/Streaming_Kmer_Hash_Table.hpp:1003-1011. */
uint64_t ko_cap_for(uint64_t n, double lf, int pow2)
{
    uint64_t cap = (uint64_t)ceil((double)n / lf / B);
    if(cap < 1)
        cap = 1;
    while((uint64_t)((double)(cap * B) * lf) <= n)
        cap++;
    return pow2 ? ceil_pow2(cap) : cap;
}

/* include/Streaming_Kmer_Hash_Table.hpp:799-804. */
int ko_fits(uint64_t n, double lf, int pow2)
{
    return (double)n / lf / B <= (double)MAX_CAP && ko_cap_for(n, lf, pow2) <= MAX_CAP;
}

/* include/Streaming_Kmer_Hash_Table.hpp:1014-1019. */
uint64_t ko_bucket_of(uint64_t h, uint64_t cap) { return ((h & UINT64_C(0xFFFFFFFF)) * cap) >> 32; }

/* include/Streaming_Kmer_Hash_Table.hpp:1063-1068. */
uint32_t ko_filter_pos(uint64_t h)
{
    const uint32_t x = (uint32_t)((h >> 32) & 0xFF);
    return x + (x == 0);
}

/* include/Streaming_Kmer_Hash_Table.hpp:1055-1060. */
uint64_t ko_pick_bucket(uint64_t x, uint64_t cap) { return (uint64_t)(((unsigned __int128)x * cap) >> 64); }

/* include/Streaming_Kmer_Hash_Table.hpp:2111-2115. */
void ko_secondaries(uint64_t h, uint64_t cap, uint64_t* p, uint64_t* q)
{
    *p = ko_pick_bucket(ko_xxh3_8(h, 0), cap);
    *q = ko_pick_bucket(ko_xxh3_8(h, UINT64_C(1) << 63), cap);
}

/* include/Concurrent_Hash_Table.hpp:161 with include/Streaming_Kmer_Hash_Table.hpp:775:
 * slots for room for n keys at load factor 0.8. */
uint64_t ko_ov_slots_for(uint64_t n) { return ceil_pow2((uint64_t)ceil((double)n / 0.8)); }

/* include/Streaming_Kmer_Hash_Table.hpp:383. */
static uint64_t step_for(uint64_t th)
{
    uint64_t s = th / (8 * 128);
    if(s < 1)
        s = 1;
    if(s > 16384)
        s = 16384;
    return s;
}

/* include/Streaming_Kmer_Hash_Table.hpp:770-784 (thresholds at construction). */
void ko_sched_init(ko_sched* s, uint64_t n, double lf, int pow2)
{
    memset(s, 0, sizeof(*s));
    s->lf = lf;
    s->cap = ko_cap_for(n, lf, pow2);
    s->resize_th = (uint64_t)((double)(s->cap * B) * lf);
    s->step = step_for(s->resize_th);
}

/* include/Streaming_Kmer_Hash_Table.hpp:2244-2286 (per-user pending counts,
 * flushes every `step` keys, resize iff the flushed size reaches the
 * threshold), :2509-2513 (the user's checkpoint at the call's end) and
 * :1679-1680 (thresholds after a doubling). */
void ko_sched_call(ko_sched* s, uint32_t user, uint64_t inserted)
{
    uint64_t pending = s->checkpoint[user], flushed = 0;
    uint64_t due = pending < s->step ? s->step - pending : 1;
    while(due <= inserted)
    {
        const uint64_t ins_c = due;
        s->approx_sz += pending + (ins_c - flushed);
        pending = 0, flushed = ins_c;
        due = ins_c + s->step;
        if(s->approx_sz >= s->resize_th)
        {
            s->cap *= 2;
            s->resize_th = (uint64_t)((double)(s->cap * B) * s->lf);
            s->step = step_for(s->resize_th);
            s->resizes++;
        }
    }
    s->checkpoint[user] = pending + (inserted - flushed);
}

/* ---------------------------------------------------------------------------
 * Semantic table: open addressing over canonical keys (< 2^62), empty = ~0. */
struct ko_tab
{
    uint32_t width;
    uint64_t max;
    uint64_t cap;   /* power of 2 */
    uint64_t size;
    uint64_t* keys;
    uint64_t* vals;
};

#define EMPTY (~UINT64_C(0))

static uint64_t slot_hash(uint64_t x) { return ko_wang_u64(x ^ UINT64_C(0x5bd1e995)); }

static void tab_alloc(ko_tab* t, uint64_t cap)
{
    t->cap = cap;
    t->keys = (uint64_t*)malloc(cap * sizeof(uint64_t));
    t->vals = (uint64_t*)calloc(cap, sizeof(uint64_t));
    memset(t->keys, 0xFF, cap * sizeof(uint64_t));
}

ko_tab* ko_tab_new(uint32_t width)
{
    ko_tab* t = (ko_tab*)calloc(1, sizeof(ko_tab));
    t->width = width;
    t->max = width == 8 ? ~UINT64_C(0) : width == 0 ? 0 : (UINT64_C(1) << (8 * width)) - 1;
    tab_alloc(t, 1024);
    return t;
}

void ko_tab_free(ko_tab* t)
{
    if(t)
        free(t->keys), free(t->vals), free(t);
}

uint64_t ko_tab_size(const ko_tab* t) { return t->size; }

static uint64_t* tab_find(const ko_tab* t, uint64_t key)
{
    uint64_t i = slot_hash(key) & (t->cap - 1);
    while(t->keys[i] != EMPTY)
    {
        if(t->keys[i] == key)
            return t->vals + i;
        i = (i + 1) & (t->cap - 1);
    }
    return NULL;
}

static void tab_grow(ko_tab* t)
{
    ko_tab old = *t;
    tab_alloc(t, old.cap * 2);
    for(uint64_t i = 0; i < old.cap; ++i)
        if(old.keys[i] != EMPTY)
        {
            uint64_t j = slot_hash(old.keys[i]) & (t->cap - 1);
            while(t->keys[j] != EMPTY)
                j = (j + 1) & (t->cap - 1);
            t->keys[j] = old.keys[i], t->vals[j] = old.vals[i];
        }
    free(old.keys), free(old.vals);
}

/* Returns the value slot of `key`, inserting it with value `v0` if absent;
 * `*absent` tells which. */
static uint64_t* tab_upsert(ko_tab* t, uint64_t key, uint64_t v0, int* absent)
{
    uint64_t* p = tab_find(t, key);
    if(p)
    {
        *absent = 0;
        return p;
    }
    if(2 * (t->size + 1) > t->cap)
        tab_grow(t);
    uint64_t i = slot_hash(key) & (t->cap - 1);
    while(t->keys[i] != EMPTY)
        i = (i + 1) & (t->cap - 1);
    t->keys[i] = key, t->vals[i] = v0, t->size++;
    *absent = 1;
    return t->vals + i;
}

/* capi/Kmer_Table.hpp:30-43 (the C ABI's operations; ADD saturates) and the
 * C++ API's wrapping `v + x` (test/check.cpp:157-158). */
static uint64_t apply(uint32_t op, uint64_t v, uint64_t x, uint64_t max)
{
    switch(op)
    {
    case KO_KEEP:   return v;
    case KO_ASSIGN: return x;
    case KO_ADD:    return v > max - x ? max : v + x;
    case KO_OR:     return v | x;
    case KO_MIN:    return v < x ? v : x;
    case KO_MAX:    return v > x ? v : x;
    default:        return (v + x) & max;
    }
}

/* Calls `fn` on each k-mer of the sequences in order (the order a single
 * thread's batch describer gives, include/Streaming_Kmer_Hash_Table.hpp:1688-1714). */
typedef void (*kmer_fn)(void* ctx, size_t q, size_t i, uint64_t key, int fwd);
static void for_kmers(const char* const* seqs, const size_t* lens, size_t n, kmer_fn fn, void* ctx)
{
    for(size_t q = 0; q < n; ++q)
    {
        const char* s = seqs[q];
        uint64_t f = 0, r = 0;
        size_t run = 0;
        for(size_t j = 0; j < lens[q]; ++j)
        {
            if(!is_acgt(s[j]))
            {
                run = 0;
                continue;
            }
            const uint64_t b = code_of(s[j]);
            f = ((f << 2) | b) & KMER_MASK;
            r = (r >> 2) | ((3 - b) << (2 * (K - 1)));
            if(++run >= K)
                fn(ctx, q, j + 1 - K, f < r ? f : r, f <= r);
        }
    }
}

typedef struct
{
    ko_tab* t;
    uint32_t op;
    uint64_t x, x_rev;
    int oriented;
    uint64_t count;
} ins_ctx;

/* include/Streaming_Kmer_Hash_Table.hpp:2392-2408 (absent -> val, present -> f(v))
 * and :2290-2300 (oriented: f(T(), fwd) / f(v, fwd)); capi/Kmer_Table.hpp:110-124. */
static void ins_fn(void* vc, size_t q, size_t i, uint64_t key, int fwd)
{
    (void)q, (void)i;
    ins_ctx* c = (ins_ctx*)vc;
    ko_tab* t = c->t;
    const uint64_t x = (c->oriented && !fwd) ? c->x_rev : c->x;
    int absent;
    uint64_t* v = tab_upsert(t, key, 0, &absent);
    c->count += absent;
    if(t->width == 0)
        return;
    if(c->oriented)
        *v = apply(c->op, absent ? 0 : *v, x, t->max);
    else
        *v = absent ? (x & t->max) : apply(c->op, *v, x, t->max);
}

uint64_t ko_tab_insert(ko_tab* t, const char* const* seqs, const size_t* lens, size_t n, uint32_t op, uint64_t x)
{
    ins_ctx c = {t, op, x, 0, 0, 0};
    for_kmers(seqs, lens, n, ins_fn, &c);
    return c.count;
}

uint64_t ko_tab_insert_oriented(ko_tab* t, const char* const* seqs, const size_t* lens, size_t n, uint32_t op, uint64_t x_fwd, uint64_t x_rev)
{
    ins_ctx c = {t, op, x_fwd, x_rev, 1, 0};
    for_kmers(seqs, lens, n, ins_fn, &c);
    return c.count;
}

/* include/Streaming_Kmer_Hash_Table.hpp:2648-2691: present -> f(v), counted with repeats. */
static void upd_fn(void* vc, size_t q, size_t i, uint64_t key, int fwd)
{
    (void)q, (void)i, (void)fwd;
    ins_ctx* c = (ins_ctx*)vc;
    uint64_t* v = tab_find(c->t, key);
    if(v)
    {
        c->count++;
        if(c->t->width)
            *v = apply(c->op, *v, c->x, c->t->max);
    }
}

uint64_t ko_tab_update(ko_tab* t, const char* const* seqs, const size_t* lens, size_t n, uint32_t op, uint64_t x)
{
    ins_ctx c = {t, op, x, 0, 0, 0};
    for_kmers(seqs, lens, n, upd_fn, &c);
    return c.count;
}

typedef struct
{
    const ko_tab* t;
    const uint64_t* off;
    uint8_t* found;
    uint64_t* values;
    uint8_t* fwd;
    uint64_t *counts, *sums, *masked;
    uint64_t m_fwd, m_rev;
} q_ctx;

static void q_fn(void* vc, size_t q, size_t i, uint64_t key, int fwd)
{
    q_ctx* c = (q_ctx*)vc;
    const uint64_t* v = tab_find(c->t, key);
    if(c->found)
    {
        const uint64_t o = c->off[q] + i;
        c->found[o] = v ? 1 : 0;
        if(c->values)
            c->values[o] = v ? *v : 0;
        if(c->fwd)
            c->fwd[o] = (uint8_t)fwd;
    }
    if(c->counts && v)
    {
        c->counts[q]++;
        c->sums[q] += *v;
        c->masked[q] += ((*v & (fwd ? c->m_fwd : c->m_rev)) != 0);
    }
}

/* capi/Kmer_Table.hpp:163-190: outputs at o_q + i, prefilled with NOT_KMER. */
void ko_tab_lookup(const ko_tab* t, const char* const* seqs, const size_t* lens, size_t n, uint8_t* found, uint64_t* values, uint8_t* fwd)
{
    uint64_t* off = (uint64_t*)malloc((n + 1) * sizeof(uint64_t));
    uint64_t p = 0;
    for(size_t q = 0; q < n; ++q)
        off[q] = p, p += (lens[q] >= K ? lens[q] - K + 1 : 0);
    memset(found, 2, p);
    if(values)
        memset(values, 0, p * sizeof(uint64_t));
    if(fwd)
        memset(fwd, 0, p);
    q_ctx c = {t, off, found, values, fwd, NULL, NULL, NULL, 0, 0};
    for_kmers(seqs, lens, n, q_fn, &c);
    free(off);
}

/* capi/Kmer_Table.hpp:134-161. */
void ko_tab_count(const ko_tab* t, const char* const* seqs, const size_t* lens, size_t n, uint64_t* counts, uint64_t* sums, uint64_t mask_fwd, uint64_t mask_rev, uint64_t* masked)
{
    memset(counts, 0, n * sizeof(uint64_t));
    memset(sums, 0, n * sizeof(uint64_t));
    memset(masked, 0, n * sizeof(uint64_t));
    q_ctx c = {t, NULL, NULL, NULL, NULL, counts, sums, masked, mask_fwd, mask_rev};
    for_kmers(seqs, lens, n, q_fn, &c);
}

static int cmp_pair(const void* a, const void* b)
{
    const uint64_t x = ((const uint64_t*)a)[0], y = ((const uint64_t*)b)[0];
    return x < y ? -1 : x > y;
}

uint64_t ko_tab_dump(const ko_tab* t, uint64_t* keys, uint64_t* values, uint64_t cap)
{
    uint64_t* kv = (uint64_t*)malloc((t->size + 1) * 2 * sizeof(uint64_t));
    uint64_t n = 0;
    for(uint64_t i = 0; i < t->cap; ++i)
        if(t->keys[i] != EMPTY)
            kv[2 * n] = t->keys[i], kv[2 * n + 1] = t->vals[i], n++;
    qsort(kv, n, 2 * sizeof(uint64_t), cmp_pair);
    const uint64_t m = n < cap ? n : cap;
    for(uint64_t i = 0; i < m; ++i)
    {
        keys[i] = kv[2 * i];
        if(values)
            values[i] = kv[2 * i + 1];
    }
    free(kv);
    return m;
}

/* capi/Kmer_Table.hpp:203-230. */
uint64_t ko_tab_reduce(const ko_tab* t, uint32_t what, uint64_t mask)
{
    uint64_t r = 0;
    for(uint64_t i = 0; i < t->cap; ++i)
        if(t->keys[i] != EMPTY)
        {
            const uint64_t v = t->vals[i] & mask;
            if(t->width == 0)
                r++;
            else if(what == 0)
                r += (v != 0);
            else if(what == 1)
                r += (uint64_t)__builtin_popcountll(v);
            else
                r += v;
        }
    return r;
}

int ko_tab_get(const ko_tab* t, uint64_t key, uint64_t* value)
{
    const uint64_t* v = tab_find(t, key);
    if(value)
        *value = v ? *v : 0;
    return v != NULL;
}
