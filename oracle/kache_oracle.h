/* Oracle: a plain-C, single-threaded restatement of kache-hash's streaming k-mer
 * table semantics (k = 31, l = 22), used ONLY as a test checker and as a CPU
 * timing reference.  Nothing in the product links or calls it.
 *
 * Pinning (what checks each function against the reference, see tests/):
 *   ko_wang_u64, ko_checksum_{mul,wang}, ko_minimizer_hash, ko_xxh3_8
 *       -> golden vectors written by the reference's own code (tests/golden/),
 *          the `xxhash` Python package, and brute force over l-mers.
 *   ko_describe -> golden describer records and a brute-force string scan.
 *   ko_cap_for / ko_fits / ko_sched_* -> the C ABI KATs (1000 keys -> 1280 / 2048
 *          slots) and the compiled reference library's capacities after growth.
 *   ko_tab_* (the semantic table) -> the compiled reference library's dumps,
 *          return values and query outputs on identical seeded inputs.
 */
#ifndef KACHE_ORACLE_H
#define KACHE_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* --- L1 encoding and hashes --------------------------------------------- */
uint64_t ko_wang_u64(uint64_t x);
uint8_t ko_checksum_mul(uint64_t key);
uint8_t ko_checksum_wang(uint64_t key);
uint64_t ko_revcomp31(uint64_t x);
uint64_t ko_minimizer_hash(uint64_t key);
uint64_t ko_xxh3_8(uint64_t x, uint64_t seed);

/* --- L2 describer ---------------------------------------------------------
 * For each position i in [0, len - 30) of `s`: valid[i] (all 31 chars ACGT),
 * and for valid positions key[i], h[i], c[i], fwd[i].  Returns the number of
 * positions (0 if len < 31).  `wang` picks the checksum variant. */
size_t ko_describe(const char* s, size_t len, int wang, uint8_t* valid, uint64_t* key, uint64_t* h, uint8_t* c, uint8_t* fwd);

/* --- L3 sizing and bucketing --------------------------------------------- */
uint64_t ko_cap_for(uint64_t n, double lf, int pow2);
int ko_fits(uint64_t n, double lf, int pow2);
uint64_t ko_bucket_of(uint64_t h, uint64_t cap);
uint32_t ko_filter_pos(uint64_t h);
uint64_t ko_pick_bucket(uint64_t x, uint64_t cap);
void ko_secondaries(uint64_t h, uint64_t cap, uint64_t* p, uint64_t* q);
uint64_t ko_ov_slots_for(uint64_t n);

/* Resize schedule: the table's capacity as the reference grows it, from the
 * per-user flush steps of each call's newly inserted keys. */
typedef struct ko_sched
{
    uint64_t cap;           /* buckets */
    uint64_t resize_th;
    uint64_t step;
    uint64_t approx_sz;
    uint64_t checkpoint[128];
    double lf;
    uint32_t resizes;
} ko_sched;

void ko_sched_init(ko_sched* s, uint64_t n, double lf, int pow2);
void ko_sched_call(ko_sched* s, uint32_t user, uint64_t inserted);

/* --- Semantic table --------------------------------------------------------
 * A brute-force map from canonical 31-mers to values of `width` bytes (0: a
 * set), applying the C ABI's operations in sequence order. */
typedef struct ko_tab ko_tab;

enum { KO_KEEP = 0, KO_ASSIGN = 1, KO_ADD = 2, KO_OR = 3, KO_MIN = 4, KO_MAX = 5, KO_ADD_WRAP = 6 };

ko_tab* ko_tab_new(uint32_t width);
void ko_tab_free(ko_tab* t);
uint64_t ko_tab_size(const ko_tab* t);
uint64_t ko_tab_insert(ko_tab* t, const char* const* seqs, const size_t* lens, size_t n, uint32_t op, uint64_t x);
uint64_t ko_tab_insert_oriented(ko_tab* t, const char* const* seqs, const size_t* lens, size_t n, uint32_t op, uint64_t x_fwd, uint64_t x_rev);
uint64_t ko_tab_update(ko_tab* t, const char* const* seqs, const size_t* lens, size_t n, uint32_t op, uint64_t x);
/* Per position, as kache_lookup lays them out; counts/sums/masked per sequence. */
void ko_tab_lookup(const ko_tab* t, const char* const* seqs, const size_t* lens, size_t n, uint8_t* found, uint64_t* values, uint8_t* fwd);
void ko_tab_count(const ko_tab* t, const char* const* seqs, const size_t* lens, size_t n, uint64_t* counts, uint64_t* sums, uint64_t mask_fwd, uint64_t mask_rev, uint64_t* masked);
/* Sorted (key, value) dump; returns the number of keys written (<= cap). */
uint64_t ko_tab_dump(const ko_tab* t, uint64_t* keys, uint64_t* values, uint64_t cap);
uint64_t ko_tab_reduce(const ko_tab* t, uint32_t what, uint64_t mask);
int ko_tab_get(const ko_tab* t, uint64_t key, uint64_t* value);

#ifdef __cplusplus
}
#endif

#endif
