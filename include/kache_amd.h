#ifndef KACHE_AMD_H
#define KACHE_AMD_H

/*
 * MI355X extensions of the kache-hash C ABI (include/kache_capi.h).  None of
 * these has a counterpart in the reference: they serve callers whose
 * sequences already sit in device memory (batch pipelines, the serving loop,
 * benchmarks), measurement, and tests that check device code on the host.
 *
 * Device texts.  A device text is `text_bytes` bytes of concatenated
 * sequences in device memory, each sequence followed by at least one
 * non-ACGT byte, so that no k-mer spans two of them.  The allocation must be
 * 16-byte aligned and readable for KACHE_AMD_TEXT_PAD bytes past text_bytes;
 * those bytes never count as bases.  The text's producer must have finished
 * before the call (a synchronous copy, or a synchronized stream): the calls
 * run on the table's own stream, and return after their work is done.
 */

#include "kache_capi.h"

#ifdef __cplusplus
extern "C" {
#endif

#define KACHE_AMD_TEXT_PAD 512

/* kache_insert / kache_insert_oriented / kache_update over a device text. */
int kache_amd_insert_device(kache_table* t, uint32_t user, const char* text, uint64_t text_bytes, uint32_t op, uint64_t x, uint64_t* inserted) KACHE_NOEXCEPT;
int kache_amd_insert_oriented_device(kache_table* t, uint32_t user, const char* text, uint64_t text_bytes, uint32_t op, uint64_t x_fwd, uint64_t x_rev, uint64_t* inserted) KACHE_NOEXCEPT;
int kache_amd_update_device(kache_table* t, const char* text, uint64_t text_bytes, uint32_t op, uint64_t x, uint64_t* present) KACHE_NOEXCEPT;

/* Batches of one device text: batch i is [bounds[i], bounds[i + 1]), each
 * ending with a non-ACGT byte.  kache_amd_insert_batches inserts them in
 * order, each as one kache_amd_insert_device call of `user` (its count of
 * new keys at inserted[i], unless `inserted` is NULL), with the launches
 * pipelined: a batch starts before the previous one's count is read
 * whenever no doubling can fall due in between.  kache_amd_count_batches
 * looks them up: present positions at hits[i] and, unless NULL, all k-mer
 * positions at positions[i], a k-mer counting only if it lies wholly within
 * its batch (so the counts are those of one kache_amd_count_device call per
 * batch); up to 256 batches run in one launch. */
int kache_amd_insert_batches(kache_table* t, uint32_t user, const char* text, const uint64_t* bounds, uint64_t n_batches, uint32_t op, uint64_t x, uint64_t* inserted) KACHE_NOEXCEPT;
int kache_amd_count_batches(const kache_table* t, const char* text, const uint64_t* bounds, uint64_t n_batches, uint64_t* hits, uint64_t* positions) KACHE_NOEXCEPT;

/* kache_lookup over all n sequences in one pass, for groups of consecutive
 * sequences with outputs of their own: group g holds sequences
 * [group_first[g], group_first[g + 1]) (the last one up to n; group_first[0]
 * is 0, non-decreasing), and its outputs go to found[g], values[g], fwd[g]
 * laid out as kache_lookup of those sequences alone.  `values` / `fwd` (the
 * arrays) or any of their entries may be NULL. */
int kache_amd_lookup_groups(const kache_table* t, const char* const* seqs, const size_t* lens, size_t n, const uint64_t* group_first,
                            size_t n_groups, uint8_t* const* found, void* const* values, uint8_t* const* fwd) KACHE_NOEXCEPT;

/* kache_count and kache_amd_lookup_groups over n sequences already in a
 * device text: sequence i spans [offsets[i], offsets[i + 1] - 1) and a
 * non-ACGT byte ends it at offsets[i + 1] - 1 (offsets: n + 1 host
 * entries, offsets[n] <= text_bytes).  Outputs as the host-sequence calls. */
int kache_amd_count_seqs_device(const kache_table* t, const char* text, uint64_t text_bytes, const uint64_t* offsets, size_t n,
                                uint64_t* counts) KACHE_NOEXCEPT;

/* kache_amd_count_seqs_device without any allocation: no reference
 * counterpart.  For callers that cannot allocate device or pinned memory
 * (code under graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, checks and outputs; its per-sequence counters
 * and segment tables live in a fixed per-device scratch array (about 36 MiB,
 * held until the process ends), so it takes at most 2^20 sequences over at
 * most 2^30 text bytes (offsets[n]) and returns KACHE_EINVAL beyond that
 * (kache_amd_last_error says why).  Calls on one device share its scratch and
 * are serialized by a lock; the scratch keeps the last call's contents. */
int kache_amd_count_seqs_device_static(const kache_table* t, const char* text, uint64_t text_bytes, const uint64_t* offsets, size_t n,
                                       uint64_t* counts) KACHE_NOEXCEPT;
int kache_amd_lookup_groups_device(const kache_table* t, const char* text, uint64_t text_bytes, const uint64_t* offsets, size_t n,
                                   const uint64_t* group_first, size_t n_groups, uint8_t* const* found, void* const* values,
                                   uint8_t* const* fwd) KACHE_NOEXCEPT;

/* The number of k-mer positions of a device text present in the table
 * (`hits`) and, unless NULL, of all its k-mer positions (`positions`). */
int kache_amd_count_device(const kache_table* t, const char* text, uint64_t text_bytes, uint64_t* hits, uint64_t* positions) KACHE_NOEXCEPT;

/* The C++ API's wrapping `v + x` (a map's `upsert(..., [](v){ return v + x; }, x)`,
 * include/Streaming_Kmer_Hash_Table.hpp:672), which the C ABI's saturating
 * KACHE_ADD does not cover: op code for the device-text and host-text calls
 * of this header.  The C ABI's own entry points refuse it (EINVAL), as the
 * reference's do (test/capi.cpp:455). */
#define KACHE_AMD_ADD_WRAP 6

/* kache_insert / kache_update over a host text: sequences separated by
 * non-ACGT bytes (e.g. newline-separated reads), in pageable host memory.
 * The text streams through two pinned buffers in chunks of 16 MiB: while one
 * chunk's kernels run, the next is packed and copied on a second stream.
 * `op` may also be KACHE_AMD_ADD_WRAP.  Validation as kache_insert /
 * kache_update (an update of a set is EINVAL). */
int kache_amd_insert_text(kache_table* t, uint32_t user, const char* text, uint64_t text_bytes, uint32_t op, uint64_t x, uint64_t* inserted) KACHE_NOEXCEPT;
int kache_amd_update_text(kache_table* t, const char* text, uint64_t text_bytes, uint32_t op, uint64_t x, uint64_t* present) KACHE_NOEXCEPT;

/* kache_insert / kache_update that also take KACHE_AMD_ADD_WRAP (the C++
 * API's wrapping v + x); otherwise the same calls. */
int kache_amd_insert_seqs(kache_table* t, uint32_t user, const char* const* seqs, const size_t* lens, size_t n, uint32_t op, uint64_t x, uint64_t* inserted) KACHE_NOEXCEPT;
int kache_amd_update_seqs(kache_table* t, const char* const* seqs, const size_t* lens, size_t n, uint32_t op, uint64_t x, uint64_t* present) KACHE_NOEXCEPT;

/* kache_amd_insert_text / kache_amd_update_text over the sequences of a FASTA
 * or FASTQ file, read as the reference's counter reads its inputs
 * (counter/Fastx_Reader.hpp:17-22): the first byte sets the format, FASTA
 * lines join up to the next header, FASTQ records have four lines with a
 * quality line as long as the sequence, lines end with "\n" or "\r\n".
 * A file that cannot be opened, is gzip-compressed, or is malformed gives
 * EINVAL (kache_amd_last_error says why; records before a malformed one may
 * have been applied).  Blocks of about 64 MiB of sequence each make one
 * text call. */
typedef int (*kache_amd_text_fn)(void* ctx, const char* text, uint64_t bytes);
/* The reader alone (host only): calls fn(ctx, text, bytes) on each block of
 * the file's sequences, each followed by '\n'; a nonzero return stops it
 * (EINVAL). */
int kache_amd_read_fastx(const char* path, kache_amd_text_fn fn, void* ctx) KACHE_NOEXCEPT;
int kache_amd_insert_fastx(kache_table* t, uint32_t user, const char* path, uint32_t op, uint64_t x, uint64_t* inserted) KACHE_NOEXCEPT;
int kache_amd_update_fastx(kache_table* t, const char* path, uint32_t op, uint64_t x, uint64_t* present) KACHE_NOEXCEPT;

/* Persistence of a table: kache_amd_save writes its device image (bucket
 * store and overflow table, as they are) with its host state (size, growth
 * schedule, registered users) to `path`; kache_amd_load makes a new table
 * from such a file, equal to the saved one in every result and in its
 * future growth.  The file is this library's own format (a versioned
 * header with hashes of itself and of the image), not the reference's,
 * which has none.  EINVAL for a file that cannot be opened, is not such a
 * file, or fails its hashes; ENOMEM when the table does not fit. */
int kache_amd_save(const kache_table* t, const char* path) KACHE_NOEXCEPT;
int kache_amd_load(const char* path, kache_table** out) KACHE_NOEXCEPT;

/* Histogram of a map's values, as the reference's counter computes it
 * (counter/Kmer_Set_Counter.cpp:97-120): out[c] = the number of keys with
 * value c for c in [1, max], out[0] = 0; `out` holds max + 1 entries.  One
 * device pass over the table (EINVAL on a set or for max > 2^30). */
int kache_amd_histogram(const kache_table* t, uint64_t max, uint64_t* out) KACHE_NOEXCEPT;

/* Device describer, for tests: for each position p of [0, text_bytes) puts
 * whether a 31-mer starts there (valid), its canonical key, minimizer hash,
 * checksum and orientation.  All pointers are device pointers. */
int kache_amd_describe_device(const char* text, uint64_t text_bytes, uint64_t* key, uint64_t* h, uint8_t* c, uint8_t* fwd, uint8_t* valid) KACHE_NOEXCEPT;

/* Device runs of the hash and placement functions, for tests: for each of the
 * n inputs x[i] writes 10 words at out[10 i]: wang_u64(x), the checksum,
 * the minimizer hash of x & (2^62 - 1), XXH3_64bits_withSeed(&x, 8, 0) and
 * with seed 2^63, bucket_of(x, cap), filter_pos(x), pick_bucket(x, cap), and
 * the secondaries p, q of x as a minimizer hash in a table of cap buckets
 * (include/Kmer_Batch.hpp:24-67, include/Kmer_Lanes.hpp:247-274,
 * include/xxHash/xxhash.h:4273-4285, include/Streaming_Kmer_Hash_Table.hpp:1014-1068,
 * 2111-2115).  x and out are device pointers; 1 <= cap <= 2^32 - 1. */
int kache_amd_hashes_device(const uint64_t* x, uint64_t n, uint64_t cap, uint64_t* out) KACHE_NOEXCEPT;

/* Kernel timing: with profiling on, HIP events bracket every launch on the
 * table's stream; kache_amd_kernel_time puts the summed milliseconds and the
 * number of launches of the kernel family `name` ("insert", "update",
 * "query", "split", "reinsert", "ov_rehash", "collect", "reduce"). */
int kache_amd_profile(kache_table* t, int on) KACHE_NOEXCEPT;
int kache_amd_kernel_time(const kache_table* t, const char* name, double* ms, uint64_t* launches) KACHE_NOEXCEPT;

/* Blocks per compute unit of the persistent launches (0: as many as fit). */
int kache_amd_set_blocks_per_cu(kache_table* t, uint32_t blocks) KACHE_NOEXCEPT;

/* HBM budget of a table (0, the default: only the device's free memory limits
 * it): the most bytes its bucket store and overflow table may hold, the new
 * arrays of a doubling or overflow growth included while both exist.  Growth
 * that would pass the budget or the device's free memory fails before it
 * allocates, with KACHE_ENOMEM: the call's launches so far stay applied (as
 * a call that fails part-way in the reference), the table keeps its capacity
 * and every result stays exact, and the next flush of the growth schedule
 * tries the doubling again (the one exception: an in-place doubling's away
 * list that fills up grows beyond the budget, see kache_amd_set_growth).  A
 * store taken from memory kept by earlier tables
 * may come mapped beyond its buckets (room an in-place doubling grows into);
 * setting a budget returns that room to the device, to within one chunk of at
 * most 1 GiB. */
int kache_amd_set_hbm_budget(kache_table* t, uint64_t bytes) KACHE_NOEXCEPT;

/* How the table doubles (the reference doubles its store in place,
 * include/Streaming_Kmer_Hash_Table.hpp:1473-1681 over include/utility.hpp:
 * 87-116; this entry point has no counterpart there).  In place: the store is
 * an address range reserving room to grow (HIP virtual memory); a doubling
 * maps its second half and splits the buckets from the top, so it needs the
 * new half, the new overflow table, a copy of 1/64 of the store and a list of
 * the keys away from their primary bucket (16 bytes each) beside the table,
 * instead of a second, doubled store.  Out of place: a new store of twice the
 * size, the old one freed after.  Results are the same either way.
 *   KACHE_AMD_GROW_AUTO          (the default) stores of 2 MiB and more can grow in place;
 *                                a doubling runs in place when an out-of-place one would not
 *                                fit the table's budget or would leave less than 1/16 of the
 *                                device's memory free, out of place otherwise (2-3% faster)
 *   KACHE_AMD_GROW_IN_PLACE      in place at every size; a store that cannot
 *                                grow in place moves out of place once, into one that can
 *   KACHE_AMD_GROW_OUT_OF_PLACE  always out of place
 * away_entries: the first size of an in-place doubling's list of away keys
 * (0: the default, 1.25 times the share of keys the table's last in-place
 * doubling listed plus 1% of the keys, 39% at first, and at least 65,536; at
 * least 256 otherwise); a list that fills up is grown and the doubling goes
 * on, beyond the table's budget if need be (the doubling cannot stop there).
 * An in-place doubling that fails once its split has begun (a device error,
 * or no device memory for a list that must grow) leaves the keys partly split:
 * its call fails, and every later call that reads or changes the table's
 * keys (inserts, updates, queries, kache_for_each, kache_get_stats,
 * kache_size, kache_amd_save ...) returns KACHE_EINTERNAL with
 * kache_amd_last_error saying so; kache_amd_get_info and kache_free still
 * work.  Out of place a failed doubling leaves the table as it was. */
enum
{
    KACHE_AMD_GROW_AUTO = 0,
    KACHE_AMD_GROW_IN_PLACE = 1,
    KACHE_AMD_GROW_OUT_OF_PLACE = 2
};
int kache_amd_set_growth(kache_table* t, uint32_t mode, uint64_t away_entries) KACHE_NOEXCEPT;

/* Bucket stores, overflow tables and doubling arrays that tables free are kept
 * for reuse by later tables and growth on the same device (the driver wipes
 * memory it gets back before handing it out again, which for tables of tens
 * of gigabytes costs more than building them): at most half of the device's
 * memory, released when an allocation finds the device full.  This releases
 * the current device's kept arrays now; `released` (may be NULL) receives
 * their bytes. */
int kache_amd_trim_cache(uint64_t* released) KACHE_NOEXCEPT;

/* Launch shape of the streaming inserts of kache_amd_insert_device and
 * kache_amd_insert_oriented_device: blocks per compute unit (0: as many as
 * fit), positions per launch (2^launch_log2, 16..30), and whether a launch
 * and its counter read-back replay a captured hipGraph (graph = 1) instead
 * of being enqueued one by one.  The shape changes speed only, never
 * results. */
typedef struct kache_amd_tuning
{
    uint32_t blocks_per_cu;
    uint32_t launch_log2;
    uint32_t graph;
    uint32_t tuning;            /* get: rounds left to the tuner, 0 when done */
    uint32_t waves_per_simd;    /* occupancy shape of the insert kernel: 6, 7 or 8 (0 on set: the default, 7) */
    uint32_t threads_per_block; /* its tile: threads per block = positions per tile, 64, 128 or 256 (0 on set: 256) */
} kache_amd_tuning;
int kache_amd_set_tuning(kache_table* t, const kache_amd_tuning* cfg) KACHE_NOEXCEPT;
int kache_amd_get_tuning(const kache_table* t, kache_amd_tuning* cfg) KACHE_NOEXCEPT;

/* Online autotuner of the insert launches of device-text calls: a
 * coordinate descent over the launch shape, one dimension per stage in the
 * order waves per SIMD the insert kernel is built for (6, 7, 8), its tile
 * (64, 128, 256 threads per block), blocks per CU (all that fit, 3/4, 1/2),
 * positions per launch (2^23, 2^24, 2^25) and graph replay (off, on).  A stage's launches alternate the best shape so far
 * (the reference) with each candidate, `rounds` times; a candidate's score
 * is the median ratio of its host-timed time per position to the mean of the
 * reference launches either side of it (which cancels the drift of the
 * table's fill and new-key share during tuning), and the best candidate
 * below 0.99 becomes the reference.  The launches are real work: nothing is
 * repeated.  rounds = 0 stops it; at most 16. */
int kache_amd_autotune(kache_table* t, uint32_t rounds) KACHE_NOEXCEPT;
typedef struct kache_amd_tuning_trial
{
    kache_amd_tuning cfg;
    uint32_t launches;
    double median_ns_per_position;
    double median_vs_reference;     /* its score (1 for a stage's reference; 0 if not yet scored) */
} kache_amd_tuning_trial;
int kache_amd_tuning_trials(const kache_table* t, kache_amd_tuning_trial* out, size_t cap, size_t* n) KACHE_NOEXCEPT;

/* Table geometry and history. */
typedef struct kache_amd_info
{
    uint64_t buckets;
    uint64_t bucket_bytes;      /* 320 (set), 448 (values of 1-4 bytes, in 4-byte words), 576 (8-byte values) */
    uint64_t overflow_slots;
    uint64_t overflow_keys;
    uint64_t resizes;
    uint64_t device_bytes;      /* bucket store + overflow table */
    uint64_t value_bytes;       /* 0 for a set */
    uint64_t next_doubling_bytes;   /* HBM the next doubling needs beside the table's arrays */
    uint64_t device_free_bytes;     /* free HBM of the table's device, arrays kept for reuse included */
    uint64_t hbm_budget;            /* kache_amd_set_hbm_budget (0: none) */
    uint64_t store_mapped_bytes;    /* HBM mapped for the bucket store (a store taken from kept memory may hold more than its buckets) */
    uint64_t store_reserved_bytes;  /* address space reserved for it (0: a store that cannot grow in place) */
    uint64_t in_place_doublings;    /* doublings done in place */
    uint64_t away_list_regrows;     /* times an in-place doubling's list of away keys grew */
    uint64_t last_away_keys;        /* keys the last in-place doubling listed (away from their primary bucket) */
    uint32_t growth_mode;           /* kache_amd_set_growth */
    uint32_t reserved0;
} kache_amd_info;
int kache_amd_get_info(const kache_table* t, kache_amd_info* out) KACHE_NOEXCEPT;

/* Device memory for device texts and outputs (hipMalloc'd; synchronous copies). */
int kache_amd_device_alloc(uint64_t bytes, void** ptr) KACHE_NOEXCEPT;
int kache_amd_device_free(void* ptr) KACHE_NOEXCEPT;
int kache_amd_copy(void* dst, const void* src, uint64_t bytes, int to_device) KACHE_NOEXCEPT;
int kache_amd_device_fill(void* ptr, uint8_t value, uint64_t bytes) KACHE_NOEXCEPT;
/* Waits for all work on the current device (hipDeviceSynchronize). */
int kache_amd_device_sync(void) KACHE_NOEXCEPT;
/* Pinned host memory (hipHostMalloc): copies from and to it run at full link
 * speed.  Lookup outputs (kache_lookup, kache_amd_lookup_groups, served
 * lookups) placed in it are written by DMA directly, without the bounce
 * through the library's own pinned chunks that pageable outputs take. */
int kache_amd_host_alloc(uint64_t bytes, void** ptr) KACHE_NOEXCEPT;
int kache_amd_host_free(void* ptr) KACHE_NOEXCEPT;

/* Seeded synthetic workloads, made on the device from a counter-based hash:
 * a genome of n uniform ACGT bases; and n_reads reads of read_len bases from
 * it as a device text (each read followed by '\n'): uniform starts, random
 * strand, substitutions at rate `sub`, N at rate `n_rate`, and a lowercase
 * run in a fraction `lower` of the reads.  kache_amd_host_synth_reads writes
 * reads [first_read, first_read + n_reads) of the same set on the host. */
int kache_amd_synth_genome(char* out, uint64_t n, uint64_t seed) KACHE_NOEXCEPT;
int kache_amd_synth_reads(char* out, uint64_t n_reads, uint64_t read_len, uint64_t genome_seed, uint64_t genome_len,
                          uint64_t seed, double sub, double n_rate, double lower) KACHE_NOEXCEPT;
int kache_amd_host_synth_reads(char* out, uint64_t first_read, uint64_t n_reads, uint64_t read_len, uint64_t genome_seed,
                               uint64_t genome_len, uint64_t seed, double sub, double n_rate, double lower) KACHE_NOEXCEPT;
/* Sequences of uneven lengths as a device text: sequence i fills
 * [offsets[i], offsets[i + 1] - 1) and '\n' ends it; `offsets` (n + 1
 * entries, device memory) must give lengths in [1, max_len], max_len <=
 * genome_len.  Each is a window of the seeded genome at a uniform start,
 * with the read model above. */
int kache_amd_synth_seqs(char* out, const uint64_t* offsets, uint64_t n, uint64_t max_len, uint64_t genome_seed, uint64_t genome_len,
                         uint64_t seed, double sub, double n_rate, double lower) KACHE_NOEXCEPT;
const char* kache_amd_synth_error(void) KACHE_NOEXCEPT;

/* Serving loop over a resident table.  Requests are submitted from any
 * thread and served in waves: pending requests are taken in arrival order,
 * as many as fit `wave_bytes` of staged sequence (a larger request is a wave
 * of its own) and at most `max_wave_requests`.  Two threads pipeline the
 * waves: a stager packs wave k + 1's sequences into one of two pinned
 * buffers and copies them to the device (its HBM budget: two such texts)
 * while an executor runs wave k from the other; a wave is admitted when a
 * buffer is free, and requests arriving meanwhile join it.  A wave runs as
 * device calls one after another, each over requests of one kind (per-
 * sequence counts, lookups, one update, or upserts of one op and operand,
 * one batch per request), its requests in ticket order; a request completes
 * when its call ends.  Unless more requests were waiting when the wave was
 * taken, its calls run shortest first across kinds (requests in size order,
 * consecutive ones of one kind merged while none is larger than the call so
 * far), so that a small request never waits for a large one of its wave;
 * a wave taken from a backlog runs its counts in one call, its lookups in
 * one, its updates, then its upserts in ticket order by runs of equal op and
 * operand.  Each request's results are those of running all requests alone,
 * one after another, in served order: by wave, then call (`phase`), then
 * ticket.  With max_wave_requests = 1
 * there is no pipeline: every request is one plain C-ABI call on one
 * worker thread (the per-request baseline).
 *
 * The request's arrays (sequences, outputs) must stay valid until its wait
 * returns.  Upserts run as `user` (from kache_register); while a server
 * runs, other callers of the table keep the C ABI's threading contract with
 * respect to its calls. */
#define KACHE_AMD_REQ_COUNT 0   /* kache_count: counts[i] per sequence */
#define KACHE_AMD_REQ_LOOKUP 1  /* kache_lookup: found / values / fwd per position (values, fwd may be NULL) */
#define KACHE_AMD_REQ_UPDATE 2  /* kache_update(op, x): *counts = present positions (counts may be NULL) */
#define KACHE_AMD_REQ_INSERT 3  /* kache_insert(op, x): *counts = new keys (counts may be NULL) */

typedef struct kache_amd_request
{
    uint32_t kind;
    uint32_t op;
    uint64_t x;
    const char* const* seqs;
    const size_t* lens;
    size_t n;
    uint64_t* counts;
    uint8_t* found;
    void* values;
    uint8_t* fwd;
} kache_amd_request;

typedef struct kache_amd_request_result
{
    int status;                 /* the call's status */
    uint32_t phase;             /* its call's index within the wave (calls run one after another) */
    uint64_t ticket;
    uint64_t wave;              /* 1, 2, ... */
    double submit_s;            /* kache_amd_clock() at submission, start of its wave, completion */
    double start_s;
    double done_s;
} kache_amd_request_result;

typedef struct kache_amd_server_stats
{
    uint64_t waves;
    uint64_t requests;
    uint64_t max_wave_requests;
    uint64_t max_wave_bytes;
    uint64_t direct_waves;      /* lone requests run at an idle server by the plain calls, unstaged */
    uint64_t refused;           /* requests completed with KACHE_ENOMEM by admission (HBM budget) */
} kache_amd_server_stats;

typedef struct kache_amd_server kache_amd_server;

int kache_amd_server_new(kache_table* t, uint32_t user, uint64_t wave_bytes, uint32_t max_wave_requests, kache_amd_server** out) KACHE_NOEXCEPT;
/* Runs the requests still pending, then stops the worker. */
void kache_amd_server_free(kache_amd_server* s) KACHE_NOEXCEPT;
int kache_amd_server_submit(kache_amd_server* s, const kache_amd_request* r, uint64_t* ticket) KACHE_NOEXCEPT;
/* Blocks until the request is done; each ticket is waited for once. */
int kache_amd_server_wait(kache_amd_server* s, uint64_t ticket, kache_amd_request_result* res) KACHE_NOEXCEPT;
int kache_amd_server_get_stats(kache_amd_server* s, kache_amd_server_stats* out) KACHE_NOEXCEPT;
/* Admission by HBM budget.  A request whose staged sequences exceed its
 * wave slot runs alone in a slot grown for it, admitted only if the growth
 * fits both the server's staging budget (the bytes of its slots together;
 * 0, the default, sets none) and the device's free memory less what the
 * table's next doubling needs (kache_amd_info.next_doubling_bytes), so that
 * staging never takes the memory the table needs to grow; otherwise the
 * request completes with KACHE_ENOMEM without running, and the server goes
 * on with the next (kache_amd_server_stats.refused). */
int kache_amd_server_set_hbm_budget(kache_amd_server* s, uint64_t bytes) KACHE_NOEXCEPT;
/* Open-loop load: submits request i at `arrival_s[i]` seconds (non-decreasing)
 * after the call starts, from the calling thread, while another thread waits
 * for them in order; returns when all are done, with results[i] for each. */
int kache_amd_server_replay(kache_amd_server* s, const kache_amd_request* reqs, const double* arrival_s, size_t n,
                            kache_amd_request_result* results) KACHE_NOEXCEPT;
/* Seconds on the steady clock the results use. */
double kache_amd_clock(void) KACHE_NOEXCEPT;

/* Description of the last failure of a call on this thread. */
const char* kache_amd_last_error(void) KACHE_NOEXCEPT;

/* Host runs of the device's own hash and sizing code, for tests. */
uint64_t kache_amd_host_minimizer_hash(uint64_t key) KACHE_NOEXCEPT;
uint8_t kache_amd_host_checksum(uint64_t key) KACHE_NOEXCEPT;
uint64_t kache_amd_host_xxh3(uint64_t x, uint64_t seed) KACHE_NOEXCEPT;
uint64_t kache_amd_host_cap_for(uint64_t n, double lf, uint32_t pow2) KACHE_NOEXCEPT;
/* Whether a table for n keys at load factor lf (pow2: bucket count rounded up
 * to a power of 2) is within the most buckets supported
 * (include/Streaming_Kmer_Hash_Table.hpp:799-804). */
int kache_amd_host_fits(uint64_t n, double lf, uint32_t pow2) KACHE_NOEXCEPT;
/* Describes a host text the way one device tile does (packing, windows and
 * window minima), position by position; same outputs as the device call. */
int kache_amd_host_describe(const char* text, uint64_t text_bytes, uint64_t* key, uint64_t* h, uint8_t* c, uint8_t* fwd, uint8_t* valid) KACHE_NOEXCEPT;

#ifdef __cplusplus
}
#endif

#endif
