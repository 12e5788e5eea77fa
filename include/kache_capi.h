#ifndef KACHE_HASH_CAPI_H
#define KACHE_HASH_CAPI_H

/*
 * kache-hash C ABI, served by an MI355X (gfx950) device table.
 *
 * Drop-in for the reference's capi/kache_capi.h: the same types, enums, struct
 * layouts and 18 entry points, with the same argument checks and status codes,
 * so that callers (the Rust crate rust/kache-hash-sys, K2Rmini-style tools)
 * link unchanged.  Each declaration names the reference declaration it
 * replaces (capi/kache_capi.h:line) and the implementation it mirrors
 * (capi/kache_capi.cpp:line).
 *
 * Tables hold canonical 31-mers (minimizers of length 22) in HBM.  Sequences
 * are host arrays: every call copies them to the device, runs the fused
 * describe-and-probe kernels, and copies results back before returning.
 * Threading: inserts may overlap one another (each user on one thread at a
 * time), updates may overlap one another, queries may overlap one another; a
 * table serializes its device work, so overlapping calls are safe and run in
 * turn (capi/kache_capi.h:14-19).
 */

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#define KACHE_NOEXCEPT noexcept
#else
#define KACHE_NOEXCEPT
#endif

/* Status codes (capi/kache_capi.h:34-41). */
enum
{
    KACHE_OK = 0,
    KACHE_EINVAL = 1,
    KACHE_EUSERS = 2,
    KACHE_ENOMEM = 3,
    KACHE_EINTERNAL = 4
};

#define KACHE_MAX_USERS 128 /* capi/kache_capi.h:43 */

/* How an insert or update combines a value v with an operand x (capi/kache_capi.h:46-54). */
enum
{
    KACHE_KEEP = 0,     /* v */
    KACHE_ASSIGN = 1,   /* x */
    KACHE_ADD = 2,      /* v + x, saturating at the width's maximum */
    KACHE_OR = 3,       /* v | x */
    KACHE_MIN = 4,      /* min(v, x) */
    KACHE_MAX = 5       /* max(v, x) */
};

/* Reductions over the keys, with a mask m of their values v (capi/kache_capi.h:58-63). */
enum
{
    KACHE_REDUCE_KEYS = 0,      /* keys with v & m nonzero; all keys of a set */
    KACHE_REDUCE_POPCOUNT = 1,  /* sum of popcount(v & m) */
    KACHE_REDUCE_SUM = 2        /* sum of v & m, modulo 2^64 */
};

/* Per-position lookup results (capi/kache_capi.h:66-71). */
enum
{
    KACHE_ABSENT = 0,
    KACHE_PRESENT = 1,
    KACHE_NOT_KMER = 2
};

typedef struct kache_table kache_table;     /* capi/kache_capi.h:75 */

/* capi/kache_capi.h:78-87; the layout is fixed (the Rust crate redeclares it). */
typedef struct kache_options
{
    uint32_t k;
    uint32_t value_bytes;
    uint64_t expected_keys;
    double load_factor;
    uint32_t resize_workers;
    uint32_t pow2_capacity;
} kache_options;

/* capi/kache_capi.h:90-99.  `bytes` is the device bucket store's size. */
typedef struct kache_stats
{
    uint64_t capacity;
    uint64_t bytes;
    uint64_t keys;
    uint64_t overflow_keys;
    double primary_pct;
    double secondary_pct;
    double overflow_pct;
} kache_stats;

/* capi/kache_capi.h:101-106: keys in 2-bit encoding (A,C,G,T = 0..3, first
 * base most significant); values of the table's width, NULL in a set. */
typedef int (*kache_keys_fn)(void* ctx, const uint64_t* keys, const void* values, size_t n);

/* capi/kache_capi.h:111 / capi/kache_capi.cpp:68-71 */
int kache_supports(uint32_t k, uint32_t value_bytes) KACHE_NOEXCEPT;

/* capi/kache_capi.h:116 / capi/kache_capi.cpp:74-77 */
kache_options kache_default_options(uint32_t k, uint32_t value_bytes) KACHE_NOEXCEPT;

/* capi/kache_capi.h:120 / capi/kache_capi.cpp:80-108.  ENOMEM when HBM cannot hold it. */
int kache_new(const kache_options* options, kache_table** out) KACHE_NOEXCEPT;

/* capi/kache_capi.h:123 / capi/kache_capi.cpp:111-114 */
void kache_free(kache_table* t) KACHE_NOEXCEPT;

/* capi/kache_capi.h:128 / capi/kache_capi.cpp:117-135 */
int kache_register(kache_table* t, uint32_t* user) KACHE_NOEXCEPT;

/* capi/kache_capi.h:135 / capi/kache_capi.cpp:138-151 */
int kache_insert(kache_table* t, uint32_t user, const char* const* seqs, const size_t* lens, size_t n, uint32_t op, uint64_t x, uint64_t* inserted) KACHE_NOEXCEPT;

/* capi/kache_capi.h:141 / capi/kache_capi.cpp:154-168 */
int kache_insert_oriented(kache_table* t, uint32_t user, const char* const* seqs, const size_t* lens, size_t n, uint32_t op, uint64_t x_fwd, uint64_t x_rev, uint64_t* inserted) KACHE_NOEXCEPT;

/* capi/kache_capi.h:148 / capi/kache_capi.cpp:171-184 */
int kache_update(kache_table* t, const char* const* seqs, const size_t* lens, size_t n, uint32_t op, uint64_t x, uint64_t* present) KACHE_NOEXCEPT;

/* capi/kache_capi.h:154 / capi/kache_capi.cpp:187-193 */
int kache_count(const kache_table* t, const char* const* seqs, const size_t* lens, size_t n, uint64_t* counts) KACHE_NOEXCEPT;

/* capi/kache_capi.h:161 / capi/kache_capi.cpp:196-202 */
int kache_count_masked(const kache_table* t, const char* const* seqs, const size_t* lens, size_t n, uint64_t mask_fwd, uint64_t mask_rev, uint64_t* counts) KACHE_NOEXCEPT;

/* capi/kache_capi.h:166 / capi/kache_capi.cpp:205-211 */
int kache_sum(const kache_table* t, const char* const* seqs, const size_t* lens, size_t n, uint64_t* sums) KACHE_NOEXCEPT;

/* capi/kache_capi.h:176 / capi/kache_capi.cpp:214-220 */
int kache_lookup(const kache_table* t, const char* const* seqs, const size_t* lens, size_t n, uint8_t* found, void* values, uint8_t* fwd) KACHE_NOEXCEPT;

/* capi/kache_capi.h:182 / capi/kache_capi.cpp:223-237 */
int kache_get(const kache_table* t, const char* kmer, uint8_t* found, uint64_t* value) KACHE_NOEXCEPT;

/* capi/kache_capi.h:186 / capi/kache_capi.cpp:240-246 */
int kache_size(const kache_table* t, uint64_t* keys) KACHE_NOEXCEPT;

/* capi/kache_capi.h:192 / capi/kache_capi.cpp:249-255.  `workers` is checked
 * (at least 1) but the reduction runs as one device kernel. */
int kache_reduce(const kache_table* t, uint32_t what, uint64_t mask, uint32_t workers, uint64_t* out) KACHE_NOEXCEPT;

/* capi/kache_capi.h:198 / capi/kache_capi.cpp:258-264 */
int kache_for_each(const kache_table* t, uint32_t shards, uint32_t shard, kache_keys_fn f, void* ctx) KACHE_NOEXCEPT;

/* capi/kache_capi.h:201 / capi/kache_capi.cpp:267-273 */
int kache_get_stats(const kache_table* t, kache_stats* out) KACHE_NOEXCEPT;

/* capi/kache_capi.h:204 / capi/kache_capi.cpp:276-287 */
const char* kache_strerror(int code) KACHE_NOEXCEPT;

#ifdef __cplusplus
}
#endif

#endif
