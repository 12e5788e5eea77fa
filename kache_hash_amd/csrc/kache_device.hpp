// Device side of the MI355X k-mer table: the fused describer (2-bit packing,
// rolling canonical 31-mers, 22-mer Wang hashes, window-10 minimum), the
// bucket probes, the value atomics and the kernels built from them.
//
// Layout in HBM (one table):
//   bucket b, at buckets + b * bucket_bytes:
//     [ 0, 32)  checksum of slot j (0 = empty)          include/Streaming_Kmer_Hash_Table.hpp:91-97
//     [32, 64)  flag bits 1..255: the spill filter       (:1063-1093)
//     [64, 320) 32 stored keys (~key, 0 = empty)
//     [320, ..) in a map, the 32 values: 4-byte words for widths 1-4
//               (448-byte buckets), 8-byte words for widths 5-8 (576 bytes)
//   overflow: power-of-two array of slots in blocks of 32, each block its 32
//   stored keys then (in a map) their 32 values, linear probing from
//   XXH3_64bits(key) (include/Concurrent_Hash_Table.hpp:28, 43-52, 205-300).
// A value sits zero-extended in its word, so one atomic of the word's size
// updates it.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

namespace kamd
{

constexpr int K = 31;
constexpr int L = 22;
constexpr int W = K - L + 1;                        // l-mers per k-mer
constexpr uint32_t B = 32;                          // slots per bucket
constexpr uint64_t KMER_MASK = (uint64_t(1) << (2 * K)) - 1;
constexpr uint64_t LMER_MASK = (uint64_t(1) << (2 * L)) - 1;
constexpr int TILE = 256;                           // positions per tile = threads per block
constexpr int TILE_WORDS = 10;                      // 32-byte words readable from a tile start: its last wave stages bytes 192..319
constexpr uint32_t META_BYTES = 64;
constexpr uint64_t TEXT_PAD = 512;                  // readable non-ACGT bytes required past a text's end

enum Op : uint32_t { OP_KEEP = 0, OP_ASSIGN = 1, OP_ADD = 2, OP_OR = 3, OP_MIN = 4, OP_MAX = 5, OP_ADD_WRAP = 6 };

// ---------------------------------------------------------------------------
// Encoding and hashes (host and device, so host tests run the same code).

// include/Kmer_Batch.hpp:24-35.
__host__ __device__ __forceinline__ uint64_t wang_u64(uint64_t x)
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

// include/Kmer_Batch.hpp:44-67, the Wang variant (AVX-512 builds): the top
// byte of wang_u64(key), never 0.
__host__ __device__ __forceinline__ uint8_t checksum(uint64_t key)
{
    const uint8_t c = static_cast<uint8_t>(wang_u64(key) >> 56);
    return c ? c : 1;
}

// Reverses the order of the 32 two-bit groups of x.
__host__ __device__ __forceinline__ uint64_t rev_pairs(uint64_t x)
{
    x = ((x >> 2) & 0x3333333333333333ull) | ((x & 0x3333333333333333ull) << 2);
    x = ((x >> 4) & 0x0F0F0F0F0F0F0F0Full) | ((x & 0x0F0F0F0F0F0F0F0Full) << 4);
    return __builtin_bswap64(x);
}

// Minimizer hash of the canonical 31-mer `key` from its own l-mers
// (include/Kmer_Lanes.hpp:247-274): min over offsets p of
// wang(min(fwd l-mer at p, reverse complement of it)).
__host__ __device__ __forceinline__ uint64_t minimizer_hash(uint64_t key)
{
    const uint64_t rc = rev_pairs(~(key << 2)) & KMER_MASK;
    uint64_t m = ~uint64_t(0);
#pragma unroll
    for(int p = 0; p < W; ++p)
    {
        const uint64_t lf = (key >> (2 * (K - L - p))) & LMER_MASK;
        const uint64_t lr = (rc >> (2 * p)) & LMER_MASK;
        const uint64_t hv = wang_u64(lf < lr ? lf : lr);
        m = hv < m ? hv : m;
    }
    return m;
}

// XXH3_64bits_withSeed(&x, 8, seed) of xxHash 0.8.2: the 4-8 byte path
// (include/xxHash/xxhash.h:4273-4285) and XXH3_rrmxmx (:4206-4213), with the
// secret's bytes 8..23 (:3975) folded into one constant.
__host__ __device__ __forceinline__ uint64_t xxh3_8(uint64_t x, uint64_t seed)
{
    constexpr uint64_t secret_8 = 0x1cad21f72c81017cull, secret_16 = 0xdb979083e96dd4deull;
    const uint32_t lo = static_cast<uint32_t>(seed);
    const uint32_t swapped = (lo >> 24) | ((lo >> 8) & 0xFF00u) | ((lo << 8) & 0xFF0000u) | (lo << 24);
    seed ^= static_cast<uint64_t>(swapped) << 32;
    const uint64_t input64 = (x >> 32) + (static_cast<uint64_t>(static_cast<uint32_t>(x)) << 32);
    uint64_t h = input64 ^ ((secret_8 ^ secret_16) - seed);
    constexpr uint64_t prime_mx2 = 0x9FB21C651E98DF25ull;
    h ^= ((h << 49) | (h >> 15)) ^ ((h << 24) | (h >> 40));
    h *= prime_mx2;
    h ^= (h >> 35) + 8;
    h *= prime_mx2;
    return h ^ (h >> 28);
}

// include/Streaming_Kmer_Hash_Table.hpp:1014-1019.
__host__ __device__ __forceinline__ uint64_t bucket_of(uint64_t h, uint64_t cap) { return ((h & 0xFFFFFFFFull) * cap) >> 32; }

// include/Streaming_Kmer_Hash_Table.hpp:1055-1060: high word of x * cap.
__host__ __device__ __forceinline__ uint64_t pick_bucket(uint64_t x, uint64_t cap)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul64hi(x, cap);
#else
    return static_cast<uint64_t>((static_cast<unsigned __int128>(x) * cap) >> 64);
#endif
}

// include/Streaming_Kmer_Hash_Table.hpp:1063-1068.
__host__ __device__ __forceinline__ uint32_t filter_pos(uint64_t h)
{
    const uint32_t x = static_cast<uint32_t>((h >> 32) & 0xFF);
    return x + (x == 0);
}

// include/Streaming_Kmer_Hash_Table.hpp:2111-2115.
__host__ __device__ __forceinline__ void secondaries(uint64_t h, uint64_t cap, uint64_t& p, uint64_t& q)
{
    p = pick_bucket(xxh3_8(h, 0), cap);
    q = pick_bucket(xxh3_8(h, uint64_t(1) << 63), cap);
}

// Descriptor pieces of the 32-base window `w` (first base most significant)
// starting at a position: its 31-mer's forward and reverse-complement forms
// and the hash of its first l-mer's canonical form.
__host__ __device__ __forceinline__ void window_parts(uint64_t w, uint64_t& kf, uint64_t& kr, uint64_t& lh)
{
    const uint64_t r = rev_pairs(~w);
    kf = w >> 2;
    kr = r & KMER_MASK;
    const uint64_t lf = w >> (2 * (32 - L)), lr = r & LMER_MASK;
    lh = wang_u64(lf < lr ? lf : lr);
}

// Whether a character is ACGT in either case (include/DNA_Utility.hpp:199-224).
__host__ __device__ __forceinline__ uint32_t is_base(uint32_t ch)
{
    const uint32_t x = ch | 0x20u;
    return (x == 'a') | (x == 'c') | (x == 'g') | (x == 't');
}

// Packs 8 characters (char j in byte j of v) into their 2-bit codes
// ((ch >> 2) ^ (ch >> 1)) & 3, A/a 0, C/c 1, G/g 2, T/t 3
// (include/Kmer_Batch.hpp:108), first character most significant (16 bits),
// and a mask of the non-ACGT ones (bit j: char j; is_base), with the bytes
// processed side by side in the 64-bit word.
__host__ __device__ __forceinline__ void pack8(uint64_t v, uint32_t& code, uint32_t& bad)
{
    // Codes of the byte-reversed word, char j's code at bits 8 (7 - j), then
    // gathered pairwise: char j ends at bits 2 (7 - j).
    uint64_t c = __builtin_bswap64(v);
    c = ((c >> 2) ^ (c >> 1)) & 0x0303030303030303ull;
    c = (c | (c >> 6)) & 0x000F000F000F000Full;
    c = (c | (c >> 12)) & 0x000000FF000000FFull;
    code = static_cast<uint32_t>((c | (c >> 24)) & 0xFFFFu);
    // Bytes equal to a, c, g or t after case folding (an exact zero-byte test
    // per candidate), then the high bit of each non-base byte gathered to bit j.
    const uint64_t x = v | 0x2020202020202020ull, lo7 = 0x7F7F7F7F7F7F7F7Full;
    uint64_t good = 0;
    for(const uint64_t b : {0x61ull, 0x63ull, 0x67ull, 0x74ull})
    {
        const uint64_t y = x ^ (b * 0x0101010101010101ull);
        good |= ~(((y & lo7) + lo7) | y | lo7);
    }
    const uint64_t nb = (~good & 0x8080808080808080ull) >> 7;
    bad = static_cast<uint32_t>((nb * 0x0102040810204080ull) >> 56);
}

// ---------------------------------------------------------------------------
// Device memory primitives.

#if defined(__HIP_DEVICE_COMPILE__) || defined(__HIPCC__)

__device__ __forceinline__ uint64_t ld_acq(const uint64_t* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ __forceinline__ bool cas64(uint64_t* p, uint64_t& expected, uint64_t desired)
{
    return __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t add64(uint64_t* p, uint64_t v) { return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ __forceinline__ void or32(uint32_t* p, uint32_t v) { __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// Atomics on a value word (4 or 8 bytes).
template <typename T> __device__ __forceinline__ void at_store(T* p, T v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
template <typename T> __device__ __forceinline__ T at_add(T* p, T v) { return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
template <typename T> __device__ __forceinline__ void at_or(T* p, T v) { __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
template <typename T> __device__ __forceinline__ void at_and(T* p, T v) { __hip_atomic_fetch_and(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
template <typename T> __device__ __forceinline__ void at_min(T* p, T v) { __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
template <typename T> __device__ __forceinline__ void at_max(T* p, T v) { __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
template <typename T> __device__ __forceinline__ void at_xchg(T* p, T v) { __hip_atomic_exchange(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// Orders the wave's LDS writes before its later LDS reads by other lanes.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Masks of the 4 bytes of x equal to the bytes of cc, as bits 0..3.
__device__ __forceinline__ uint32_t byte_eq4(uint32_t x, uint32_t cc)
{
    const uint32_t y = x ^ cc;
    const uint32_t t = ~(((y & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | y | 0x7F7F7F7Fu);
    return ((t >> 7) * 0x01020408u) >> 24;
}

// The slots of a bucket whose checksums equal `c`, and the empty ones.
__device__ __forceinline__ void cs_masks(const uint8_t* meta, uint8_t c, uint32_t& eq, uint32_t& zero)
{
    const uint4* p = reinterpret_cast<const uint4*>(meta);
    const uint4 a = p[0], b = p[1];
    const uint32_t d[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    const uint32_t cc = 0x01010101u * c;
    eq = 0, zero = 0;
#pragma unroll
    for(int i = 0; i < 8; ++i)
    {
        eq |= byte_eq4(d[i], cc) << (4 * i);
        zero |= byte_eq4(d[i], 0u) << (4 * i);
    }
}

// A table as the kernels see it.
struct TableView
{
    uint8_t* buckets;
    uint64_t cap;           // buckets
    uint64_t* ov;           // overflow slots
    uint64_t ov_mask;       // overflow slots - 1
    uint64_t vmax;          // largest value of the width
    uint32_t width;         // value bytes (0 in a set)
};

// Value bytes per slot, VB: 0 (a set), 4 (maps of widths 1-4) or 8 (widths
// 5-8).  V is the value word.
template <int VB> struct Geo
{
    static_assert(VB == 0 || VB == 4 || VB == 8, "value words are 4 or 8 bytes");
    using V = typename std::conditional<VB == 4, uint32_t, uint64_t>::type;
    static constexpr uint32_t block_bytes = B * (8 + VB);               // 32 keys, then their 32 values
    static constexpr uint32_t bucket_bytes = META_BYTES + block_bytes;
};

// The layout of a table whose values are `width` bytes.
__host__ __device__ constexpr uint32_t value_word(uint32_t width) { return width == 0 ? 0 : width <= 4 ? 4 : 8; }

template <int VB>
__device__ __forceinline__ uint8_t* bucket_at(const TableView& T, uint64_t b) { return T.buckets + b * Geo<VB>::bucket_bytes; }

// Keys and values are kept in separate arrays of a block of 32 slots (a
// bucket past its metadata, or 32 overflow slots): the key of slot j at word
// j, the values after the 32 keys.
template <int VB>
__device__ __forceinline__ uint64_t* slot_at(uint8_t* bucket, uint32_t j) { return reinterpret_cast<uint64_t*>(bucket + META_BYTES) + j; }

template <int VB>
__host__ __device__ __forceinline__ uint64_t* ov_slot(uint64_t* ov, uint64_t i)
{
    if constexpr(VB == 0)
        return ov + i;
    else if constexpr(VB == 8)
        return ov + (i & ~uint64_t(B - 1)) * 2 + (i & (B - 1));
    else
        return ov + (i / B) * (Geo<VB>::block_bytes / 8) + (i & (B - 1));
}

// A slot as the probes hand it on: its key word, and where its value is.
// With 8-byte values that is 32 words past the key; with 4-byte values
// (block + 256 + 4j for slot j of a block) it depends on j, which the 4-byte
// layout keeps beside the key pointer.
template <int VB> struct Slot
{
    uint64_t* k;
    Slot() = default;
    __device__ __forceinline__ Slot(uint64_t* key, uint32_t) : k(key) {}
    __device__ __forceinline__ uint32_t j() const { return 0; }
    __device__ __forceinline__ typename Geo<VB>::V* v() const { return k + B; }
};

template <> struct Slot<4>
{
    uint64_t* k;
    uint32_t jj;
    Slot() = default;
    __device__ __forceinline__ Slot(uint64_t* key, uint32_t j) : k(key), jj(j) {}
    __device__ __forceinline__ uint32_t j() const { return jj; }
    __device__ __forceinline__ uint32_t* v() const { return reinterpret_cast<uint32_t*>(k) + 2 * B - jj; }
};

template <int VB>
__device__ __forceinline__ Slot<VB> slot_ref(uint8_t* bucket, uint32_t j) { return Slot<VB>(slot_at<VB>(bucket, j), j); }

template <int VB>
__device__ __forceinline__ Slot<VB> ov_ref(uint64_t* ov, uint64_t i) { return Slot<VB>(ov_slot<VB>(ov, i), static_cast<uint32_t>(i & (B - 1))); }

enum Res : int { RES_FOUND = 0, RES_NEW = 1, RES_FULL = 2 };

// Read-only probe: the slot holding `sk` (a null key pointer if none);
// `full` iff the bucket's last slot is taken (slots fill in order).
template <int VB>
__device__ __forceinline__ Slot<VB> probe_find(uint8_t* bucket, uint64_t sk, uint8_t c, bool& full)
{
    // Checksum word i's equal bytes b flag bit 8b + i of m (slot 4i + b): the
    // zero-byte test leaves 0x80 in each, moved down by 7 - i, with no
    // multiply to gather them.  Candidates come in no particular slot order,
    // which is fine here: a key sits in at most one slot of a bucket.
    const uint4* p = reinterpret_cast<const uint4*>(bucket);
    const uint4 a = p[0], b = p[1];
    const uint32_t d[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    const uint32_t cc = 0x01010101u * c;
    uint32_t m = 0;
#pragma unroll
    for(int i = 0; i < 8; ++i)
    {
        const uint32_t y = d[i] ^ cc;
        m |= (~(((y & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | y | 0x7F7F7F7Fu)) >> (7 - i);
    }
    full = (d[7] >> 24) != 0;       // slots fill in order: the last one taken
    while(m)
    {
        const uint32_t bit = __builtin_ctz(m);
        m &= m - 1;
        const uint32_t j = 4 * (bit & 7) + (bit >> 3);
        uint64_t* s = slot_at<VB>(bucket, j);
        if(*s == sk)
            return Slot<VB>(s, j);
    }
    return Slot<VB>(nullptr, 0);
}

// Overflow table: linear probing from XXH3_64bits(key) (include/Concurrent_Hash_Table.hpp:205-300).
template <int VB>
__device__ __forceinline__ Res ov_insert(const TableView& T, uint64_t key, uint64_t sk, Slot<VB>& slot)
{
    uint64_t i = xxh3_8(key, 0) & T.ov_mask;
    for(uint64_t n = 0; n <= T.ov_mask; ++n, i = (i + 1) & T.ov_mask)
    {
        const Slot<VB> s = ov_ref<VB>(T.ov, i);
        uint64_t cur = ld_acq(s.k);
        if(cur == sk)
            return slot = s, RES_FOUND;
        if(cur == 0)
        {
            if(cas64(s.k, cur, sk))
                return slot = s, RES_NEW;
            if(cur == sk)
                return slot = s, RES_FOUND;
        }
    }
    return RES_FULL;
}

template <int VB>
__device__ __forceinline__ Slot<VB> ov_find(const TableView& T, uint64_t key, uint64_t sk)
{
    uint64_t i = xxh3_8(key, 0) & T.ov_mask;
    for(uint64_t n = 0; n <= T.ov_mask; ++n, i = (i + 1) & T.ov_mask)
    {
        const Slot<VB> s = ov_ref<VB>(T.ov, i);
        const uint64_t cur = *s.k;
        if(cur == sk)
            return s;
        if(cur == 0)
            return Slot<VB>(nullptr, 0);
    }
    return Slot<VB>(nullptr, 0);
}

// Marks the spill of minimizer hash `h` in the bucket's filter (:1079-1085).
__device__ __forceinline__ void mark_spill(uint8_t* bucket, uint64_t h)
{
    const uint32_t i = filter_pos(h);
    or32(reinterpret_cast<uint32_t*>(bucket + 32) + (i >> 5), 1u << (i & 31));
}

__device__ __forceinline__ bool may_spill(const uint8_t* bucket, uint64_t h)
{
    const uint32_t i = filter_pos(h);
    return (bucket[32 + (i >> 3)] >> (i & 7)) & 1u;
}

// ---------------------------------------------------------------------------
// Wave-cooperative inserts.  Slot words hold 0 (empty), a stored key
// (~key: top bits 11), a reservation (stored key with bit 63 flipped: top
// bits 01, the key still readable) or a tombstone (TOMB: top bits 10).
// Reservations and tombstones exist only inside insert launches; tombstones
// carry a nonzero checksum, never match a key, and vanish at the next
// doubling.

constexpr uint64_t RESERVE_FLIP = uint64_t(1) << 63;
constexpr uint64_t TOMB = uint64_t(1) << 63;
constexpr uint32_t CLAIM_ROUNDS = 256;

__device__ __forceinline__ bool is_key(uint64_t v) { return (v >> 62) == 3; }

// The set bits of m in the lanes below this one.
__device__ __forceinline__ uint32_t lanes_below(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32), __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u));
}

__device__ __forceinline__ uint64_t readlane64(uint64_t v, int l)
{
    return (uint64_t(static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v >> 32), l))) << 32) |
           static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), l));
}

__device__ __forceinline__ uint64_t exchange64(uint64_t* p, uint64_t v)
{
    return __hip_atomic_exchange(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Inserts with the reference's policy (include/Streaming_Kmer_Hash_Table.hpp:2431-2490):
// the primary bucket; once it is full, the spill mark, then the secondaries
// and the overflow table.  The secondaries go in a fixed order, p then q, so
// that concurrent inserters of one key walk the same slots.  RES_FULL only if
// the overflow table has no room (or after CLAIM_ROUNDS rounds).  Slots fill
// in order and keys are never removed, so a walk over a bucket's slots from
// its first empty one finds a concurrent insertion of the same key wherever it
// went; the checksums serve as hints only (they may lag their keys).  All lanes
// of a wave insert at once (every lane calls it; `active` lanes have a key).
// Lanes that need a new slot in the same bucket, from the same lowest
// possibly-free slot z, take consecutive slots z, z + 1, ... in one round of
// CASes, ranked by lane, instead of contending for one slot at a time: the
// first claims with its key (nothing below z can hold it: those slots'
// checksums were read), the others with a reservation.  A reserving lane then
// reads the slots from z below its own: all are taken (by the round's other
// lanes or by others before), so the key can only be there as a key or a
// reservation; if it is, the lowest such slot is the key's (its owner sees no
// copy below it) and the lane turns its reservation into a tombstone;
// otherwise it writes its key over the reservation.  Nothing waits on
// anything, so a wave never blocks on another.  A lane whose CAS loses learns
// that the slot is taken and tries again above it.
template <int VB>
__device__ __forceinline__ Res insert_key_wave(const TableView& T, bool active, uint64_t key, uint64_t h, uint8_t c, Slot<VB>& slot, bool& in_ov)
{
    const uint64_t sk = ~key;
    in_ov = false;
    slot = Slot<VB>(nullptr, 0);
    Res r = RES_FOUND;
    // Per-lane state: the policy's stage (0 primary, 1 p, 2 q) and the lowest
    // slot of the current bucket not known to be taken.
    uint32_t stage = 0, lo = 0;
    uint64_t bidx = active ? bucket_of(h, T.cap) : 0;
    bool need = active;
    for(uint32_t round = 0; __ballot(need); ++round)
    {
        if(round == CLAIM_ROUNDS)
        {
            if(need)
                r = RES_FULL, need = false;
            break;
        }
        uint8_t* bk = bucket_at<VB>(T, bidx);
        uint32_t z = B;
        if(need)
        {
            uint32_t eq, zero;
            cs_masks(bk, c, eq, zero);
            while(eq)
            {
                const uint32_t j = __builtin_ctz(eq);
                eq &= eq - 1;
                uint64_t* s = slot_at<VB>(bk, j);
                if(ld_acq(s) == sk)
                {
                    slot = Slot<VB>(s, j), r = RES_FOUND, need = false;
                    break;
                }
            }
            z = zero ? __builtin_ctz(zero) : B;
            z = z > lo ? z : lo;
        }
        const uint64_t claimers = __ballot(need);
        if(!claimers)
            break;
        // Groups: the wave's claimers of the same bucket from the same slot.
        const uint64_t gkey = (bidx << 6) | z;
        uint64_t grp = 0;
        for(uint64_t rem = claimers; rem;)
        {
            const uint64_t gl = readlane64(gkey, __builtin_ctzll(rem));
            const uint64_t g = __ballot(need && gkey == gl);
            if(need && gkey == gl)
                grp = g;
            rem &= ~g;
        }
        const uint64_t below_me = (uint64_t(1) << __lane_id()) - 1;
        // A lane whose key a lower lane of its group also has follows the
        // lowest such lane.  Every lane takes part in each shuffle: a lane
        // that had left a divergent loop would be read as inactive.
        int owner = -1;
        uint64_t below = need ? grp & below_me : 0;
        while(__ballot(below != 0))
        {
            const int l = below ? __builtin_ctzll(below) : static_cast<int>(__lane_id());
            const uint64_t o = static_cast<uint64_t>(__shfl(static_cast<unsigned long long>(sk), l, 64));
            if(below)
            {
                if(o == sk)
                    owner = l, below = 0;
                else
                    below &= below - 1;
            }
        }
        const uint64_t leads = __ballot(need && owner < 0);
        if(need && owner < 0)
        {
            const uint64_t rk = sk ^ RESERVE_FLIP;
            const uint32_t rank = lanes_below(grp & leads);
            const uint32_t j = z + rank;
            // The first of a group has no slot to check below its own (those
            // below z hold keys whose checksums were read): it claims with its
            // key at once.  The others reserve, then check.
            bool mine = false;
            uint64_t occ = 0;
            if(j < B)
                mine = cas64(slot_at<VB>(bk, j), occ, rank == 0 ? sk : rk);
            // Slots z .. j-1 were the targets of this group's lower-ranked
            // lanes: those that won them hold other keys, so only a slot
            // whose CAS lost can hold a concurrent claim of this key.
            const uint64_t lost = __ballot(j < B && !mine);
            if(mine && rank == 0)
            {
                bk[j] = c;
                slot = slot_ref<VB>(bk, j), r = RES_NEW, need = false;
            }
            else
            {
                // The lowest slot from z that holds or will hold this key.
                uint32_t cand = B;
                const uint32_t top = (grp & lost & below_me) ? (j < B ? j : B) : z;
                for(uint32_t t = z; t < top; ++t)
                {
                    const uint64_t v = ld_acq(slot_at<VB>(bk, t));
                    if(v == sk || v == rk)
                    {
                        cand = t;
                        break;
                    }
                }
                if(cand == B && j < B && !mine && (occ == sk || occ == rk))
                    cand = j;
                if(cand < B)
                {
                    if(mine && exchange64(slot_at<VB>(bk, j), TOMB) == rk)
                        bk[j] = c;
                    slot = slot_ref<VB>(bk, cand), r = RES_FOUND, need = false;
                }
                else if(mine)
                {
                    if(exchange64(slot_at<VB>(bk, j), sk) == rk)
                        bk[j] = c;
                    slot = slot_ref<VB>(bk, j), r = RES_NEW, need = false;
                }
                else if(j >= B)
                {
                    // Every slot is taken and none holds the key: the next bucket of the policy.
                    if(stage == 0)
                        mark_spill(bk, h);
                    if(stage < 2)
                    {
                        uint64_t p, q;
                        secondaries(h, T.cap, p, q);
                        bidx = stage == 0 ? p : q;
                        stage++, lo = 0;
                    }
                    else
                    {
                        in_ov = true;
                        r = ov_insert<VB>(T, ~sk, sk, slot);
                        need = false;
                    }
                }
                else
                    lo = j + 1;     // taken since the checksums were read: slots up to j are known taken
            }
        }
        if(__ballot(owner >= 0))
        {
            // Followers take their owner's outcome, or its next step.
            const int src = owner >= 0 ? owner : static_cast<int>(__lane_id());
            // State bits: need 0, r 1-2, in_ov 3, stage 4-7, lo 8-15, the slot's j 16-20.
            const uint32_t o_state = __shfl(static_cast<int>(need) | (static_cast<int>(r) << 1) | (static_cast<int>(in_ov) << 3) |
                                            static_cast<int>(stage << 4) | static_cast<int>(lo << 8) | static_cast<int>(slot.j() << 16), src, 64);
            const uint64_t o_slot = __shfl(static_cast<unsigned long long>(reinterpret_cast<uintptr_t>(slot.k)), src, 64);
            const uint64_t o_b = __shfl(static_cast<unsigned long long>(bidx), src, 64);
            if(owner >= 0)
            {
                if(!(o_state & 1))
                {
                    slot = Slot<VB>(reinterpret_cast<uint64_t*>(static_cast<uintptr_t>(o_slot)), (o_state >> 16) & (B - 1));
                    r = ((o_state >> 1) & 3) == RES_FULL ? RES_FULL : RES_FOUND;
                    in_ov = (o_state >> 3) & 1;
                    need = false;
                }
                else
                    bidx = o_b, stage = (o_state >> 4) & 15, lo = (o_state >> 8) & (VB == 4 ? 255u : ~0u);
            }
        }
    }
    return r;
}

// insert_key_wave for keys known to be absent from the table and distinct
// from every other key being inserted concurrently: a doubling's reinsert,
// whose keys each sit once in the old table and are neither among the
// primaries the split moved nor among one another.  Then no claim can meet
// its own key, so the lanes of a group take their slots z, z + 1, ... with
// their keys at once: no checksum matches to read, no reservations to turn
// into keys, no slots below to scan.  A lost CAS means the slot went to
// another key: the lane tries again above it.  Same policy, same result
// (RES_NEW, or RES_FULL if the overflow table has no room).
template <int VB>
__device__ __forceinline__ Res insert_new_wave(const TableView& T, bool active, uint64_t key, uint64_t h, uint8_t c, Slot<VB>& slot, bool& in_ov)
{
    const uint64_t sk = ~key;
    in_ov = false;
    slot = Slot<VB>(nullptr, 0);
    Res r = RES_NEW;
    uint32_t stage = 0, lo = 0;
    uint64_t bidx = active ? bucket_of(h, T.cap) : 0;
    bool need = active;
    for(uint32_t round = 0; __ballot(need); ++round)
    {
        if(round == CLAIM_ROUNDS)
        {
            if(need)
                r = RES_FULL, need = false;
            break;
        }
        uint8_t* bk = bucket_at<VB>(T, bidx);
        uint32_t z = B;
        if(need)
        {
            uint32_t eq, zero;
            cs_masks(bk, c, eq, zero);
            z = zero ? __builtin_ctz(zero) : B;
            z = z > lo ? z : lo;
        }
        const uint64_t claimers = __ballot(need);
        const uint64_t gkey = (bidx << 6) | z;
        uint64_t grp = 0;
        for(uint64_t rem = claimers; rem;)
        {
            const uint64_t gl = readlane64(gkey, __builtin_ctzll(rem));
            const uint64_t g = __ballot(need && gkey == gl);
            if(need && gkey == gl)
                grp = g;
            rem &= ~g;
        }
        if(need)
        {
            const uint32_t j = z + lanes_below(grp);
            if(j < B)
            {
                uint64_t occ = 0;
                if(cas64(slot_at<VB>(bk, j), occ, sk))
                {
                    bk[j] = c;
                    slot = slot_ref<VB>(bk, j), need = false;
                }
                else
                    lo = j + 1;
            }
            else
            {
                // Every slot is taken: the next bucket of the policy.
                if(stage == 0)
                    mark_spill(bk, h);
                if(stage < 2)
                {
                    uint64_t p, q;
                    secondaries(h, T.cap, p, q);
                    bidx = stage == 0 ? p : q;
                    stage++, lo = 0;
                }
                else
                {
                    in_ov = true;
                    r = ov_insert<VB>(T, key, sk, slot);
                    need = false;
                }
            }
        }
    }
    return r;
}

// Lookup policy (include/Streaming_Kmer_Hash_Table.hpp:2548-2574): the
// primary; absent unless it is full and its filter marks the minimizer; the
// secondaries; the overflow table iff both are full.
template <int VB>
__device__ __forceinline__ Slot<VB> find_key(const TableView& T, uint64_t key, uint64_t h, uint8_t c)
{
    const uint64_t sk = ~key;
    uint8_t* bb = bucket_at<VB>(T, bucket_of(h, T.cap));
    bool full;
    const Slot<VB> s = probe_find<VB>(bb, sk, c, full);
    if(s.k || !full || !may_spill(bb, h))
        return s;

    uint64_t p, q;
    secondaries(h, T.cap, p, q);
    bool full_p, full_q;
    const Slot<VB> sp = probe_find<VB>(bucket_at<VB>(T, p), sk, c, full_p);
    if(sp.k)
        return sp;
    const Slot<VB> sq = probe_find<VB>(bucket_at<VB>(T, q), sk, c, full_q);
    if(sq.k)
        return sq;
    if(full_p && full_q)
        return ov_find<VB>(T, key, sk);
    return Slot<VB>(nullptr, 0);
}

// ---------------------------------------------------------------------------
// Values.  An empty slot's value is 0; every operation here is one atomic on
// the value's word (plus a rare fix-up), written so that the winner of a
// key's insertion and concurrent updaters of it commute.  Operands are at
// most vmax, the width's largest value.

// v + x saturating at vmax (capi/Kmer_Table.hpp:38): the sum, and vmax
// written by whoever crosses it.  Every add whose sum passes vmax (one that
// wraps the word included) sees old > vmax - x, and its write of vmax comes
// after its add, so once any sum has crossed, the word's last operation is a
// write of vmax; adds that landed between a crossing and its write vanish
// into the maximum, as they must.
template <typename T>
__device__ __forceinline__ void sat_add(T* v, T x, T vmax)
{
    if(at_add(v, x) > vmax - x)
        at_xchg(v, vmax);
}

// v + x modulo 2^(8 * width): whoever carries out of the width clears it.
template <typename T>
__device__ __forceinline__ void wrap_add(T* v, T x, T vmax)
{
    const T old = at_add(v, x);
    if(vmax != ~T(0) && (old & vmax) + x > vmax)
        at_and(v, vmax);
}

// Applies `op` to the value `v` with operand `x`; `winner` iff this call just
// inserted the key, with the value 0 in place.
template <uint32_t OP, typename T>
__device__ __forceinline__ void value_op(T* v, uint64_t x64, uint64_t vmax64, bool winner)
{
    const T x = static_cast<T>(x64), vmax = static_cast<T>(vmax64);
    if constexpr(OP == OP_KEEP)
    {
        if(winner)
            at_store(v, x);
    }
    else if constexpr(OP == OP_ASSIGN)
        at_store(v, x);
    else if constexpr(OP == OP_ADD)
        sat_add(v, x, vmax);
    else if constexpr(OP == OP_OR)
        at_or(v, x);
    else if constexpr(OP == OP_MIN)
    {
        if(winner)
            at_store(v, x);
        else
            at_min(v, x);
    }
    else if constexpr(OP == OP_MAX)
        at_max(v, x);
    else
        wrap_add(v, x, vmax);
}

// ---------------------------------------------------------------------------
// The describer.  Positions go in tiles of TILE (one per block and turn);
// wave w of a block describes positions 64w .. 64w + 63 of the tile on its
// own, with no block barrier, so a wave waiting on its probes holds back no
// other wave.  Staging: 16 lanes pack 8 characters each (the wave's
// 64 + K + W - 2 bases, rounded up to 128) into 2-bit codes and non-ACGT
// masks in the wave's LDS words; every lane then builds its 32-base window
// from two words, hashes its l-mer, and takes the minimum of 10 consecutive
// l-mer hashes from LDS (include/Kmer_Batch.hpp:191-250, the same quantities
// per position).

struct Desc
{
    uint64_t key;
    uint64_t h;
    uint32_t valid;
    uint32_t fwd;
    uint8_t c;
};

constexpr int WAVE_WORDS = 4;           // 32-base words a wave stages
constexpr int PACK_LANES = 8 * WAVE_WORDS / 2;   // lanes packing 8 characters each

struct WaveLds
{
    uint64_t pk[WAVE_WORDS];            // 2-bit codes, first base most significant
    uint32_t bad[WAVE_WORDS];           // bit j: character j of the word is not ACGT
    uint64_t lh[64 + W];
};

__device__ __forceinline__ uint64_t window_at(const WaveLds& s, int u)
{
    const int wi = u >> 5, sh = (u & 31) * 2;
    const uint64_t a = s.pk[wi];
    return sh ? (a << sh) | (s.pk[wi + 1] >> (64 - sh)) : a;
}

// Starts loading the 8 characters a packing lane stages for the tile at p0:
// bytes p0 + 64w + 8i .. + 7 for lane i of wave w, within the TILE_WORDS * 32
// readable bytes that `text + p0` (16-byte aligned) must have.
__device__ __forceinline__ void load_wave_tile(const uint8_t* text, uint64_t p0, uint2& w)
{
    const uint32_t lane = __lane_id();
    if(lane < PACK_LANES)
        w = *reinterpret_cast<const uint2*>(text + p0 + 64 * (threadIdx.x >> 6) + 8 * lane);
}

// Describes position p0 + threadIdx.x of the tile at p0 from the wave's
// loaded characters `w`, then starts loading those of the tile at `p_next`
// into `w` (if `has_next`).  Positions at or past `pos_end` are not
// described; a k-mer must also end within the first `text_len` bytes.
__device__ __forceinline__ Desc describe_wave(const uint8_t* text, uint64_t p0, uint64_t pos_end, uint64_t text_len, WaveLds& s,
                                              uint2& w, uint64_t p_next, bool has_next)
{
    const int t = static_cast<int>(__lane_id());
    if(t < PACK_LANES)
    {
        // Lane t holds characters 8t .. 8t + 7: bits 16 (3 - t % 4) of word
        // t / 4 (its first character most significant) and byte t % 4 of its mask.
        uint32_t code, bad;
        pack8(static_cast<uint64_t>(w.x) | (static_cast<uint64_t>(w.y) << 32), code, bad);
        reinterpret_cast<uint16_t*>(s.pk)[(t & ~3) + 3 - (t & 3)] = static_cast<uint16_t>(code);
        reinterpret_cast<uint8_t*>(s.bad)[t] = static_cast<uint8_t>(bad);
    }
    wave_sync();
    if(has_next)
        load_wave_tile(text, p_next, w);

    Desc d;
    uint64_t kf, kr, lh;
    window_parts(window_at(s, t), kf, kr, lh);
    s.lh[t] = lh;
    if(t < W - 1)
    {
        uint64_t f2, r2, lh2;
        window_parts(window_at(s, 64 + t), f2, r2, lh2);
        s.lh[64 + t] = lh2;
    }
    const int wi = t >> 5;
    const uint64_t b64 = static_cast<uint64_t>(s.bad[wi]) | (static_cast<uint64_t>(s.bad[wi + 1]) << 32);
    const uint64_t p = p0 + threadIdx.x;
    d.valid = (((b64 >> (t & 31)) & 0x7FFFFFFFull) == 0) && (p < pos_end) && (p + K <= text_len);
    wave_sync();

    uint64_t h = s.lh[t];
#pragma unroll
    for(int o = 1; o < W; ++o)
    {
        const uint64_t y = s.lh[t + o];
        h = y < h ? y : h;
    }
    d.h = h;
    d.key = kf < kr ? kf : kr;
    d.fwd = kf <= kr;
    d.c = checksum(d.key);
    // The next tile's codes overwrite the LDS words only after every lane has read them.
    wave_sync();
    return d;
}

// Walks the positions [p_begin, pos_end) in tiles, block by block, each
// wave's words prefetched during its previous tile: calls f(tile, desc) on
// every lane, where tile t covers positions origin + t * TILE onwards (lane
// position origin + t * TILE + threadIdx.x) and origin is p_begin rounded
// down to 16 bytes (tiles load aligned words); positions before p_begin are
// not described.  `s` is the wave's own LDS.
// BT: positions per tile, the block's threads (64, 128 or 256: a tile's
// last wave reads up to BT + 64 bytes past the tile's start, within the 512
// readable bytes a device text has past its end).
template <int BT = TILE, typename F>
__device__ __forceinline__ void for_tiles(const uint8_t* text, uint64_t p_begin, uint64_t pos_end, uint64_t text_len, WaveLds& s, const F& f)
{
    static_assert(BT == 64 || BT == 128 || BT == 256, "tiles of 1, 2 or 4 waves");
    const uint64_t origin = p_begin & ~uint64_t(15);
    const uint64_t tiles = (pos_end - origin + BT - 1) / BT;
    uint2 w{};
    if(blockIdx.x < tiles)
        load_wave_tile(text, origin + blockIdx.x * uint64_t(BT), w);
    for(uint64_t tile = blockIdx.x; tile < tiles; tile += gridDim.x)
    {
        const uint64_t next = tile + gridDim.x;
        Desc d = describe_wave(text, origin + tile * BT, pos_end, text_len, s, w, origin + next * BT, next < tiles);
        d.valid &= (origin + tile * BT + threadIdx.x >= p_begin);
        f(tile, d);
    }
}

// Sums `v` over the block into `*dst` with one atomic.
__device__ __forceinline__ void block_add(uint64_t v, uint64_t* dst, uint64_t* red)
{
    for(int o = 32; o > 0; o >>= 1)
        v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if(lane == 0)
        red[wave] = v;
    __syncthreads();
    if(threadIdx.x == 0)
    {
        uint64_t s = 0;
        for(int i = 0; i < static_cast<int>(blockDim.x >> 6); ++i)
            s += red[i];
        if(s)
            add64(dst, s);
    }
    __syncthreads();
}

// Per-launch counters.
struct Counters
{
    uint64_t inserted;      // keys new to the table
    uint64_t ov_inserted;   // of which into the overflow table
    uint64_t present;       // positions found (updates, queries)
    uint64_t failed;        // positions that found no room (must stay 0)
    uint64_t positions;     // k-mer positions seen (queries)
};

// In device memory a set of counters is CNT_STRIPES copies of Counters, one
// per 128-byte line: block b adds its sums into copy b % CNT_STRIPES, so the
// blocks that end a launch together do not queue on one word (one atomic per
// block on one word measured 22.8 us per launch of 1,792 blocks, spread over
// 64 lines 2.4 us, as with none), and whoever reads the set adds the copies
// up (k_publish, hand_off).
constexpr uint32_t CNT_STRIPES = 64;
constexpr uint32_t CNT_STRIPE_WORDS = 16;
constexpr uint64_t CNT_SET_WORDS = uint64_t(CNT_STRIPES) * CNT_STRIPE_WORDS;
constexpr uint32_t CNT_WORDS = sizeof(Counters) / 8;
static_assert(CNT_WORDS <= CNT_STRIPE_WORDS, "a stripe holds the counters");

// This block's copy of counter set `c`.
__device__ __forceinline__ Counters* stripe_of(Counters* c)
{
    return reinterpret_cast<Counters*>(reinterpret_cast<uint64_t*>(c) + (blockIdx.x % CNT_STRIPES) * CNT_STRIPE_WORDS);
}

#endif

}
