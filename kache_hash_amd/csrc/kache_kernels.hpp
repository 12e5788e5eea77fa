// Kernels of the MI355X k-mer table.  All are grid-stride ("persistent"):
// the host launches about as many blocks as the device keeps resident, and
// each block walks tiles until the range is done.
#pragma once

#include "kache_device.hpp"

// Occupancy of the streaming kernels: minimum waves per SIMD the compiler
// must fit, which caps their registers.  The insert kernel's settle path
// needs more registers than the lookups: at 8 waves its common path reloads
// spilled values inside the loop; 7 waves (72 registers, a few spilled
// outside the common path) measured fastest, then 6 (80, none spilled).
#ifndef KAMD_STREAM_WAVES
#define KAMD_STREAM_WAVES 8
#endif
#ifndef KAMD_INSERT_WAVES
#define KAMD_INSERT_WAVES 7
#endif
#define KAMD_STREAM_BOUNDS __launch_bounds__(256, KAMD_STREAM_WAVES)
#define KAMD_INSERT_BOUNDS __launch_bounds__(256, KAMD_INSERT_WAVES)

namespace kamd
{

// Sum over the wave.
__device__ __forceinline__ uint64_t wave_sum(uint64_t v)
{
    for(int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, 64);
    return v;
}

// ---------------------------------------------------------------------------
// Inserts / upserts (include/Streaming_Kmer_Hash_Table.hpp:2217-2514): every
// valid position of [pos_begin, pos_end) of `text`.  `x` is the operand, or
// the forward one with ORIENT (x_rev for k-mers occurring reverse-complemented).
//
// Two paths per position.  The common one (a key already in its primary
// bucket) reads the checksum line and the slots whose checksums match, then
// applies the value atomic.  Every other position (a new key, or one placed
// in a secondary bucket or the overflow table) is queued in the wave's LDS
// ring and settled 64 at a time by insert_key_wave with the whole wave busy,
// instead of holding a wave of mostly finished lanes through its claim rounds.  Operations within a
// launch commute, so the order in which positions settle changes nothing.
// A stale checksum or key line read by the common path only sends a
// position to the queue; the queue's path reads slots coherently.
constexpr uint32_t DEFER_RING = 128;   // > 63 left over + 64 new

struct DeferRing
{
    uint64_t key[DEFER_RING];
    uint64_t h[DEFER_RING];
    uint32_t cf[DEFER_RING];   // checksum | forward << 8
};

// One wave (all lanes): adds up counter set `src` (lane l reads copy l) into
// `dst` (host memory: CNT_WORDS words) and zeroes the set.
__device__ __forceinline__ void publish_set(uint64_t* src, uint64_t* dst)
{
    static_assert(CNT_STRIPES == 64, "one lane per copy");
    const uint32_t lane = __lane_id();
    uint64_t* mine = src + uint64_t(lane) * CNT_STRIPE_WORDS;
    for(uint32_t i = 0; i < CNT_WORDS; ++i)
    {
        const uint64_t v = wave_sum(mine[i]);
        mine[i] = 0;
        if(lane == 0)
            __hip_atomic_store(dst + i, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// The counters of the previous launch of a pipeline, handed to the host by
// the next launch: set `src` (complete, since launches on a stream run in
// order) is added up into host memory and cleared, then `flag` is set to
// `seq`.
struct Handoff
{
    uint64_t* src;
    uint64_t* dst;
    uint64_t* flag;
    uint64_t seq;
};

__device__ __forceinline__ void hand_off(const Handoff& h)
{
    if(h.src && blockIdx.x == 0 && threadIdx.x < 64)
    {
        publish_set(h.src, h.dst);
        if(threadIdx.x == 0)
            __hip_atomic_store(h.flag, h.seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// WAVES: the waves per SIMD its registers must allow; BT: the block's
// threads, i.e. positions per tile (the launch shape the tuner chooses).
template <int VB, uint32_t OP, bool ORIENT, int WAVES = KAMD_INSERT_WAVES, int BT = TILE>
__global__ __launch_bounds__(BT, WAVES) void k_insert(TableView T, const uint8_t* __restrict__ text, uint64_t text_len, uint64_t pos_begin, uint64_t pos_end, uint64_t x, uint64_t x_rev, Counters* cnt, Handoff prev)
{
    hand_off(prev);
    __shared__ WaveLds wave_lds[BT / 64];
    __shared__ uint64_t red[BT / 64];
    __shared__ DeferRing rings[BT / 64];
    WaveLds& s = wave_lds[threadIdx.x >> 6];
    DeferRing& q = rings[threadIdx.x >> 6];
    const uint32_t lane = __lane_id();
    uint32_t head = 0, queued = 0;   // the wave's ring: entries [head, head + queued) mod DEFER_RING
    uint64_t ins = 0, ins_ov = 0, failed = 0;     // the wave's counts

    // Settles the ring entry of each lane with `active` (all lanes take part).
    auto settle = [&](bool active, uint32_t i)
    {
        const uint64_t key = active ? q.key[i] : 0, h = active ? q.h[i] : 0;
        const uint32_t cf = active ? q.cf[i] : 0;
        Slot<VB> slot;
        bool in_ov;
        const Res r = insert_key_wave<VB>(T, active, key, h, static_cast<uint8_t>(cf), slot, in_ov);
        // The wave's counts (uniform: scalar registers).
        ins += __popcll(__ballot(active && r == RES_NEW));
        ins_ov += __popcll(__ballot(active && r == RES_NEW && in_ov));
        failed += __popcll(__ballot(active && r == RES_FULL));
        if(active && r != RES_FULL)
            if constexpr(VB != 0)
                value_op<OP>(slot.v(), (ORIENT && !(cf >> 8)) ? x_rev : x, T.vmax, !ORIENT && r == RES_NEW);
    };

    for_tiles<BT>(text, pos_begin, pos_end, text_len, s, [&](uint64_t, const Desc& d)
    {
        Slot<VB> slot(nullptr, 0);
        if(d.valid)
        {
            bool full;
            slot = probe_find<VB>(bucket_at<VB>(T, bucket_of(d.h, T.cap)), ~d.key, d.c, full);
            if constexpr(VB != 0)
                if(slot.k)
                    value_op<OP>(slot.v(), (ORIENT && !d.fwd) ? x_rev : x, T.vmax, false);
        }
        const bool defer = d.valid && !slot.k;
        const uint64_t m = __ballot(defer);
        if(defer)
        {
            const uint32_t i = (head + queued + lanes_below(m)) % DEFER_RING;
            q.key[i] = d.key, q.h[i] = d.h, q.cf[i] = d.c | (d.fwd << 8);
        }
        queued += __popcll(m);
        if(queued >= 64)
        {
            wave_sync();
            settle(true, (head + lane) % DEFER_RING);
            head = (head + 64) % DEFER_RING, queued -= 64;
        }
    });
    if(queued)
    {
        wave_sync();
        settle(lane < queued, (head + lane) % DEFER_RING);
    }
    block_add(lane == 0 ? ins : 0, &stripe_of(cnt)->inserted, red);
    block_add(lane == 0 ? ins_ov : 0, &stripe_of(cnt)->ov_inserted, red);
    block_add(lane == 0 ? failed : 0, &stripe_of(cnt)->failed, red);
}

// Present-only updates (include/Streaming_Kmer_Hash_Table.hpp:2648-2691).
template <int VB, uint32_t OP>
__global__ KAMD_STREAM_BOUNDS void k_update(TableView T, const uint8_t* __restrict__ text, uint64_t text_len, uint64_t pos_begin, uint64_t pos_end, uint64_t x, Counters* cnt)
{
    __shared__ WaveLds wave_lds[TILE / 64];
    WaveLds& s = wave_lds[threadIdx.x >> 6];
    __shared__ uint64_t red[TILE / 64];
    uint64_t present = 0;
    for_tiles(text, pos_begin, pos_end, text_len, s, [&](uint64_t, const Desc& d)
    {
        if(d.valid)
            if(const Slot<VB> slot = find_key<VB>(T, d.key, d.h, d.c); slot.k)
            {
                present++;
                value_op<OP>(slot.v(), x, T.vmax, false);
            }
    });
    block_add(present, &stripe_of(cnt)->present, red);
}

// ---------------------------------------------------------------------------
// Queries.  The staged text holds segments: segment s starts at seg_start[s],
// belongs to sequence seg_q[s] and starts at its position seg_off[s];
// tile_seg[t] is the segment holding position t * TILE.

enum QueryMode : int { Q_TOTAL = 0, Q_COUNT = 1, Q_SUM = 2, Q_MASKED = 3, Q_LOOKUP = 4 };

struct QueryArgs
{
    const uint64_t* seg_start;
    const uint32_t* seg_q;
    const uint64_t* seg_off;
    const uint32_t* tile_seg;
    const uint64_t* out_off;    // per sequence: index of its first position in the lookup outputs
    uint64_t* per_seq;          // counts / sums / masked counts, per sequence
    uint8_t* found;
    uint8_t* values;
    uint8_t* fwd;
    uint32_t vbytes;
    uint64_t mask_fwd, mask_rev;
};

// The segment holding position p, from the tile's bracket.
__device__ __forceinline__ uint32_t segment_of(const QueryArgs& a, uint64_t tile, uint64_t p)
{
    uint32_t lo = a.tile_seg[tile], hi = a.tile_seg[tile + 1];
    while(lo < hi)
    {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if(a.seg_start[mid] <= p)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

template <int VB, int MODE>
__global__ KAMD_STREAM_BOUNDS void k_query(TableView T, const uint8_t* __restrict__ text, uint64_t pos_begin, uint64_t pos_end, QueryArgs a, Counters* cnt)
{
    __shared__ WaveLds wave_lds[TILE / 64];
    WaveLds& s = wave_lds[threadIdx.x >> 6];
    __shared__ uint64_t red[TILE / 64];
    uint64_t total = 0, positions = 0;
    for_tiles(text, pos_begin, pos_end, pos_end, s, [&](uint64_t tile, const Desc& d)
    {
        const uint64_t p = (pos_begin & ~uint64_t(15)) + tile * TILE + threadIdx.x;
        uint64_t v = 0;
        bool hit = false;
        if(d.valid)
            if(const Slot<VB> slot = find_key<VB>(T, d.key, d.h, d.c); slot.k)
            {
                hit = true;
                if constexpr(VB != 0)
                    v = *slot.v();
            }

        if constexpr(MODE == Q_TOTAL)
            total += hit, positions += d.valid;
        else if constexpr(MODE == Q_LOOKUP)
        {
            if(d.valid)
            {
                const uint32_t sg = segment_of(a, tile, p);
                const uint64_t o = a.out_off[a.seg_q[sg]] + a.seg_off[sg] + (p - a.seg_start[sg]);
                a.found[o] = hit ? 1 : 0;
                if(a.fwd)
                    a.fwd[o] = static_cast<uint8_t>(d.fwd);
                if(a.values)
                {
                    if(a.vbytes == 1)
                        a.values[o] = static_cast<uint8_t>(v);
                    else if(a.vbytes == 2)
                        reinterpret_cast<uint16_t*>(a.values)[o] = static_cast<uint16_t>(v);
                    else if(a.vbytes == 4)
                        reinterpret_cast<uint32_t*>(a.values)[o] = static_cast<uint32_t>(v);
                    else
                        reinterpret_cast<uint64_t*>(a.values)[o] = v;
                }
            }
        }
        else
        {
            // Per-sequence sums: lanes of one sequence add once per wave.
            uint64_t add = 0;
            if(hit)
                add = MODE == Q_COUNT ? 1 : MODE == Q_SUM ? v : ((v & (d.fwd ? a.mask_fwd : a.mask_rev)) != 0);
            uint32_t q = 0xFFFFFFFFu;
            if(add)
                q = a.seg_q[segment_of(a, tile, p)];
            uint64_t pending = __ballot(add != 0);
            while(pending)
            {
                const int leader = __builtin_ctzll(pending);
                const uint32_t q0 = __shfl(q, leader, 64);
                const bool mine = (add != 0) && (q == q0);
                const uint64_t sum = wave_sum(mine ? add : 0);
                if((threadIdx.x & 63) == leader)
                    add64(a.per_seq + q0, sum);
                pending &= ~__ballot(mine);
            }
        }
    });
    if constexpr(MODE == Q_TOTAL)
    {
        block_add(total, &stripe_of(cnt)->present, red);
        block_add(positions, &stripe_of(cnt)->positions, red);
    }
}

// The number of batch bounds bounds[1..n] at or below p (p the same on
// every lane): the batch of position p, or n past the last one.  One load
// per lane per 64 bounds.
__device__ __forceinline__ uint32_t bounds_at_or_below(const uint64_t* bounds, uint32_t n, uint64_t p)
{
    uint32_t c = 0;
    for(uint32_t base = 1; base <= n; base += 64)
    {
        const uint32_t i = base + __lane_id();
        c += __popcll(__ballot(i <= n && bounds[i] <= p));
    }
    return c;
}

// Lookups of the batches [bounds[i], bounds[i + 1]) of a device text in one
// launch (kache_amd_count_batches): k-mer positions and hits per batch, a
// k-mer counting only if it ends within its batch, as a launch per batch
// counts it.  This kernel counts the wave tiles (64 positions from p0) with
// no batch bound in (p0, p0 + 93], whose k-mers all lie in one batch; the
// others, the edge tiles (at most two per bound), are k_count_edges'.  A
// wave's tiles go up the text, so it leaves a batch for good: it sums its
// positions of the batch it is in, lane by lane in registers, and stores the
// sums into its own pair of part (n x waves pairs of u32 hits, positions,
// zeroed before) when it leaves.  k_batch_sums adds part and edge up.
constexpr uint64_t EDGE_REACH = 63 + K - 1;     // a bound within p0 + 1 .. p0 + EDGE_REACH makes an edge tile

template <int VB>
__global__ KAMD_STREAM_BOUNDS void k_count_batches(TableView T, const uint8_t* __restrict__ text, const uint64_t* __restrict__ bounds, uint32_t n,
                                                   uint32_t* __restrict__ part)
{
    __shared__ WaveLds wave_lds[TILE / 64];
    WaveLds& s = wave_lds[threadIdx.x >> 6];
    const uint32_t lane = __lane_id();
    const uint64_t waves = uint64_t(gridDim.x) * (TILE / 64);
    const uint64_t wave = uint64_t(blockIdx.x) * (TILE / 64) + (threadIdx.x >> 6);
    const uint64_t origin = bounds[0] & ~uint64_t(15);
    uint32_t cur = n;               // the batch summed in registers (n: none)
    uint64_t ce = 0;                // its end (the wave's tiles only go up: past its start)
    uint32_t lh = 0, lp = 0;        // this lane's hits and positions in it
    const auto flush = [&]
    {
        const uint32_t h = static_cast<uint32_t>(wave_sum(lh)), q = static_cast<uint32_t>(wave_sum(lp));
        if(cur < n && lane == 0)
        {
            uint32_t* dst = part + 2 * (uint64_t(cur) * waves + wave);
            dst[0] = h, dst[1] = q;
        }
        lh = 0, lp = 0;
    };
    for_tiles(text, bounds[0], bounds[n], bounds[n], s, [&](uint64_t tile, const Desc& d)
    {
        const uint64_t p0 = origin + tile * TILE + (threadIdx.x & ~63u);
        if(p0 + EDGE_REACH >= ce)
        {
            const uint32_t b0 = bounds_at_or_below(bounds, n, p0);
            if(b0 >= n || p0 + EDGE_REACH >= bounds[b0 + 1])
                return;         // an edge tile (or past the last batch)
            if(b0 != cur)
            {
                flush();
                cur = b0, ce = bounds[b0 + 1];
            }
        }
        lp += d.valid;
        if(d.valid && find_key<VB>(T, d.key, d.h, d.c).k)
            lh++;
    });
    flush();
}

// The edge tiles of k_count_batches: wave unit tiles[i] (positions origin +
// 64 tiles[i] onwards, origin = bounds[0] rounded down to 16), one wave per
// tile at a time; each lane counts its position in its own batch, a k-mer
// only if it ends within that batch, into edge (n pairs of u64 hits,
// positions) by wave-aggregated atomics.
template <int VB>
__global__ __launch_bounds__(64) void k_count_edges(TableView T, const uint8_t* __restrict__ text, const uint64_t* __restrict__ bounds, uint32_t n,
                                                    const uint64_t* __restrict__ tiles, uint64_t n_tiles, uint64_t* __restrict__ edge)
{
    __shared__ WaveLds s;
    const uint32_t lane = __lane_id();
    const uint64_t origin = bounds[0] & ~uint64_t(15);
    for(uint64_t i = blockIdx.x; i < n_tiles; i += gridDim.x)
    {
        const uint64_t p0 = origin + tiles[i] * 64, p = p0 + lane;
        uint2 w{};
        load_wave_tile(text, p0, w);
        const Desc d = describe_wave(text, p0, bounds[n], bounds[n], s, w, 0, false);
        uint32_t bl = bounds_at_or_below(bounds, n, p0);
        while(bl < n && bounds[bl + 1] <= p)
            bl++;
        const bool v = d.valid && p >= bounds[0] && bl < n && p + K <= bounds[bl + 1];
        const bool hit = v && find_key<VB>(T, d.key, d.h, d.c).k != nullptr;
        uint64_t pending = __ballot(v);
        while(pending)
        {
            const uint32_t b = __shfl(bl, __builtin_ctzll(pending), 64);
            const bool mine = v && bl == b;
            const uint64_t sh = wave_sum(mine && hit), sp = wave_sum(mine);
            if(lane == 0)
                add64(edge + 2 * b, sh), add64(edge + 2 * b + 1, sp);
            pending &= ~__ballot(mine);
        }
    }
}

// Sums of k_count_batches: block b adds batch b's pairs of part and edge
// into out[2b] (hits) and out[2b + 1] (positions).
__global__ __launch_bounds__(256) void k_batch_sums(const uint32_t* __restrict__ part, uint64_t waves, const uint64_t* __restrict__ edge, uint64_t* out)
{
    __shared__ uint64_t red[2][4];
    const uint64_t b = blockIdx.x;
    uint64_t h = 0, q = 0;
    for(uint64_t w = threadIdx.x; w < waves; w += blockDim.x)
        h += part[2 * (b * waves + w)], q += part[2 * (b * waves + w) + 1];
    h = wave_sum(h), q = wave_sum(q);
    if((threadIdx.x & 63) == 0)
        red[0][threadIdx.x >> 6] = h, red[1][threadIdx.x >> 6] = q;
    __syncthreads();
    if(threadIdx.x == 0)
    {
        out[2 * b] = edge[2 * b] + red[0][0] + red[0][1] + red[0][2] + red[0][3];
        out[2 * b + 1] = edge[2 * b + 1] + red[1][0] + red[1][1] + red[1][2] + red[1][3];
    }
}

// ---------------------------------------------------------------------------
// Resize: doubling (include/Streaming_Kmer_Hash_Table.hpp:1473-1681).  Old
// bucket b's primaries move to new buckets 2b and 2b + 1, in their slot
// order; the keys away from their primary bucket are listed and inserted
// again afterwards, with the old overflow table's.
//
// In place (the reference also grows its store in place and splits from the
// top, include/utility.hpp:87-116): the store's address range is mapped to
// twice its size first, and the new buckets are written into the same range.
// Phases run from the top: phase [lo, hi) reads old buckets lo .. hi - 1 and
// writes new buckets 2 lo .. 2 hi - 1, all at or above hi (2 lo >= hi): in
// the new half, or old buckets that an earlier phase has already read.  A
// phase never writes its own sources, nor those of a later phase.  The lowest
// K buckets are copied aside before the first phase and split last, from the
// copy, so that there are log2(cap / K) + 1 phases rather than log2(cap) + 1.
// Out of place: one phase [0, cap) from the old store into a new one.
//
// The keys away from their primary bucket cannot wait in their old slots
// (in place those are overwritten), so the split lists them with their values
// in the away list: `regions` regions of `region` entries, region r shared by the
// waves whose index is r modulo `regions`.  A wave's step takes entries for
// its away keys from its region's cursor by one returning atomic, and writes
// them during the next step, so that the atomic's round trip overlaps a
// step's work instead of stalling it (waiting for it on every step made the
// split 60% slower).  A phase that finds its region full sets `fail`
// (1 + the phase), and the phases after it return at once; the host then
// grows the list and runs that phase again from the cursors it had before it
// (kept by k_away_snap).

constexpr uint32_t AWAY_REGIONS = 256;
constexpr uint32_t AWAY_CURSOR_STRIDE = 8;     // words between cursors: one 64-byte line each

struct AwayList
{
    uint64_t* keys;       // stored keys (the complement of the key)
    uint64_t* vals;       // their value words (maps)
    uint64_t region;      // entries per region
    uint64_t* cursor;     // regions cursors, AWAY_CURSOR_STRIDE words apart
    uint32_t* fail;       // 0, or 1 + the first phase that found its region full
    uint32_t regions;
};

template <int VB>
__global__ __launch_bounds__(TILE) void k_split(const uint8_t* src, TableView N, uint64_t old_cap, uint64_t lo, uint64_t hi, AwayList L, uint32_t phase)
{
    if(__hip_atomic_load(L.fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)
        return;
    constexpr uint32_t BB = Geo<VB>::bucket_bytes, IMG = 2 * BB;
    __shared__ __align__(16) uint8_t images[TILE / 64][2][IMG];
    const uint32_t lane = threadIdx.x & 63, j = lane & 31;
    const uint64_t half_mask = (lane < 32) ? 0xFFFFFFFFull : 0xFFFFFFFF00000000ull;
    const uint64_t below = (uint64_t(1) << lane) - 1;
    const uint64_t pairs = (hi - lo + 1) / 2;
    const uint64_t stride = (uint64_t(gridDim.x) * blockDim.x) >> 6;
    uint64_t w = (blockIdx.x * uint64_t(blockDim.x) + threadIdx.x) >> 6;
    const uint32_t r = static_cast<uint32_t>(w % L.regions);
    uint64_t* cursor = L.cursor + uint64_t(r) * AWAY_CURSOR_STRIDE;
    uint64_t* rkeys = L.keys + r * L.region;
    uint64_t* rvals = VB != 0 ? L.vals + r * L.region : nullptr;
    // The last step's away keys, waiting for their entries: p_at (lane 0)
    // is the cursor before them, p_n (wave-uniform) their number.
    uint64_t p_sk = 0, p_val = 0, p_at = 0;
    uint32_t p_n = 0, p_rank = 0;
    bool p_aw = false;
    // Writes them; false once the region is full (the wave then stops: the
    // phase runs again).
    const auto flush = [&]
    {
        if(p_n == 0)
            return true;
        const uint64_t at = readlane64(p_at, 0);
        if(at + p_n > L.region)
        {
            if(lane == 0)
            {
                uint32_t none = 0;
                __hip_atomic_compare_exchange_strong(L.fail, &none, phase + 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            return false;
        }
        if(p_aw)
        {
            const uint64_t i = at + p_rank;
            rkeys[i] = p_sk;
            if constexpr(VB != 0)
                rvals[i] = p_val;
        }
        return true;
    };
    // Slot j of old bucket lo + 2w + lane / 32, read a pair ahead of its use.
    struct Old
    {
        uint64_t sk, val;
        uint32_t c;
    };
    const auto read = [&](uint64_t w)
    {
        Old o{0, 0, 0};
        const uint64_t b = lo + 2 * w + (lane >> 5);
        if(w < pairs && b < hi)
        {
            uint8_t* ob = const_cast<uint8_t*>(src) + b * BB;
            const Slot<VB> os = slot_ref<VB>(ob, j);
            o.c = ob[j], o.sk = os.k[0];
            if constexpr(VB != 0)
                o.val = *os.v();
        }
        return o;
    };
    Old next = read(w);
    for(; w < pairs; w += stride)
    {
        const Old cur = next;
        next = read(w + stride);
        const uint64_t b = lo + 2 * w + (lane >> 5);
        bool lo_half = false, hi_half = false, aw = false;
        const uint64_t sk = cur.sk, val = cur.val;
        uint8_t c = static_cast<uint8_t>(cur.c);
        if(c && !is_key(sk))
            c = 0;      // a tombstone: dropped
        if(c)
        {
            const uint64_t h = minimizer_hash(~sk);
            if(bucket_of(h, old_cap) == b)
            {
                hi_half = (bucket_of(h, N.cap) == 2 * b + 1);
                lo_half = !hi_half;
            }
            else
                aw = true;
        }
        const uint64_t m_lo = __ballot(lo_half) & half_mask, m_hi = __ballot(hi_half) & half_mask, m_away = __ballot(aw);
        uint8_t* img = images[threadIdx.x >> 6][lane >> 5];
        wave_sync();    // the previous pair's image has been read out
        for(uint32_t k = j; k < IMG / 16; k += 32)
            reinterpret_cast<uint4*>(img)[k] = make_uint4(0, 0, 0, 0);
        wave_sync();
        if(lo_half || hi_half)
        {
            const uint32_t rank = __popcll((lo_half ? m_lo : m_hi) & below);
            uint8_t* nb = img + (lo_half ? 0 : BB);
            const Slot<VB> ns = slot_ref<VB>(nb, rank);
            ns.k[0] = sk;
            if constexpr(VB != 0)
                *ns.v() = static_cast<typename Geo<VB>::V>(val);
            nb[rank] = c;
        }
        wave_sync();
        if(b < hi)
        {
            uint4* dst = reinterpret_cast<uint4*>(bucket_at<VB>(N, 2 * b));
            for(uint32_t k = j; k < IMG / 16; k += 32)
                dst[k] = reinterpret_cast<const uint4*>(img)[k];
        }
        // The previous step's away keys, now that their entries are known;
        // then this step's entries are taken.
        if(!flush())
            return;
        p_n = __popcll(m_away);
        if(p_n)
        {
            if(lane == 0)
                p_at = __hip_atomic_fetch_add(cursor, uint64_t(p_n), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            p_aw = aw, p_rank = __popcll(m_away & below), p_sk = sk, p_val = val;
        }
    }
    flush();
}

// The away list's cursors before a phase: `regions` words to `snap`.
__global__ void k_away_snap(const uint64_t* cursor, uint32_t regions, uint64_t* snap)
{
    for(uint32_t r = threadIdx.x; r < regions; r += blockDim.x)
        snap[r] = cursor[uint64_t(r) * AWAY_CURSOR_STRIDE];
}

// Moves the away list into one with larger regions: region r's first
// before[r] entries (its cursor before the phase that found it full), then
// resets the cursors to those counts and clears the fail word, so that the
// phase can run again on the new list.
__global__ void k_away_regrow(AwayList from, AwayList to, const uint64_t* before)
{
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for(uint32_t r = 0; r < from.regions; ++r)
    {
        const uint64_t n = before[r];
        for(uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += stride)
        {
            to.keys[r * to.region + i] = from.keys[r * from.region + i];
            if(from.vals)
                to.vals[r * to.region + i] = from.vals[r * from.region + i];
        }
    }
    if(blockIdx.x == 0)
    {
        for(uint32_t r = threadIdx.x; r < from.regions; r += blockDim.x)
            to.cursor[uint64_t(r) * AWAY_CURSOR_STRIDE] = before[r];
        if(threadIdx.x == 0)
            *to.fail = 0;
    }
}

// The keys of the away list and of the old overflow table into the doubled
// table, with the insertion policy and the reinsert's claims (insert_new_wave).
// With `clear` (in place), each key inserted is cleared from its list entry
// or old overflow slot, so that a run which finds the new overflow table full
// can be repeated, after that table has grown, for the keys left; without it
// (out of place) the old overflow table is only read.
template <int VB>
__global__ __launch_bounds__(TILE) void k_reinsert(AwayList L, uint64_t* old_ov, uint64_t old_ov_slots, TableView N, Counters* cnt, bool clear)
{
    __shared__ uint64_t red[TILE / 64];
    uint64_t ov_n = 0, failed = 0;
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t per_region = (L.region + 63) / 64;
    const uint64_t list_chunks = uint64_t(L.regions) * per_region;
    const uint64_t chunks = list_chunks + (old_ov_slots + 63) / 64;
    const uint64_t waves = (uint64_t(gridDim.x) * blockDim.x) >> 6;
    for(uint64_t ch = (blockIdx.x * uint64_t(blockDim.x) + threadIdx.x) >> 6; ch < chunks; ch += waves)
    {
        uint64_t* at = nullptr;
        uint64_t sk = 0, val = 0;
        if(ch < list_chunks)
        {
            const uint64_t r = ch / per_region, i = (ch % per_region) * 64 + lane;
            const uint64_t used = L.cursor[r * AWAY_CURSOR_STRIDE];
            if(i < used && i < L.region)
            {
                at = L.keys + r * L.region + i;
                sk = *at;
                if constexpr(VB != 0)
                    val = L.vals[r * L.region + i];
            }
        }
        else
        {
            const uint64_t i = (ch - list_chunks) * 64 + lane;
            if(i < old_ov_slots)
            {
                const Slot<VB> os = ov_ref<VB>(old_ov, i);
                at = os.k;
                sk = *at;
                if constexpr(VB != 0)
                    if(sk)
                        val = *os.v();
            }
        }
        const bool active = sk != 0;
        Slot<VB> slot;
        bool in_ov;
        const uint64_t key = ~sk;
        const Res res = insert_new_wave<VB>(N, active, key, active ? minimizer_hash(key) : 0, active ? checksum(key) : 0, slot, in_ov);
        if(active)
        {
            if(res == RES_FULL)
                failed++;
            else
            {
                ov_n += in_ov;
                if constexpr(VB != 0)
                    *slot.v() = static_cast<typename Geo<VB>::V>(val);
                if(clear)
                    *at = 0;
            }
        }
    }
    block_add(ov_n, &cnt->ov_inserted, red);
    block_add(failed, &cnt->failed, red);
}

// Overflow growth: the old overflow table's keys into a new one
// (include/Streaming_Kmer_Hash_Table.hpp:1412-1443).
template <int VB>
__global__ __launch_bounds__(TILE) void k_ov_rehash(const uint64_t* old_ov, uint64_t old_slots, TableView N, Counters* cnt)
{
    __shared__ uint64_t red[TILE / 64];
    uint64_t failed = 0;
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for(uint64_t g = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; g < old_slots; g += stride)
    {
        const Slot<VB> os = ov_ref<VB>(const_cast<uint64_t*>(old_ov), g);
        const uint64_t sk = os.k[0];
        if(!sk)
            continue;
        Slot<VB> slot;
        if(ov_insert<VB>(N, ~sk, sk, slot) == RES_FULL)
            failed++;
        else if constexpr(VB != 0)
            at_store(slot.v(), *os.v());
    }
    block_add(failed, &cnt->failed, red);
}

// ---------------------------------------------------------------------------
// Scans: iteration shares, reductions and placement statistics.

// Copies the keys (2-bit encodings) and values of buckets [b0, b1) and
// overflow slots [o0, o1) to out_keys / out_vals, at cursor positions.
template <int VB>
__global__ __launch_bounds__(TILE) void k_collect(TableView T, uint64_t b0, uint64_t b1, uint64_t o0, uint64_t o1, uint64_t* out_keys, uint64_t* out_vals, uint64_t* cursor)
{
    const uint64_t main_n = (b1 - b0) * B, all = main_n + (o1 - o0);
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    const uint64_t base_g = blockIdx.x * uint64_t(blockDim.x) + (threadIdx.x & ~63u);
    for(uint64_t g0 = base_g; g0 < all; g0 += stride)
    {
        const uint64_t g = g0 + (threadIdx.x & 63);
        uint64_t sk = 0, val = 0;
        if(g < main_n)
        {
            uint8_t* bk = bucket_at<VB>(T, b0 + g / B);
            if(bk[g % B])
            {
                const Slot<VB> s = slot_ref<VB>(bk, g % B);
                sk = s.k[0];
                if(!is_key(sk))
                    sk = 0;
                else if constexpr(VB != 0)
                    val = *s.v();
            }
        }
        else if(g < all)
        {
            const Slot<VB> s = ov_ref<VB>(T.ov, o0 + g - main_n);
            sk = s.k[0];
            if constexpr(VB != 0)
                if(sk)
                    val = *s.v();
        }
        const uint64_t m = __ballot(sk != 0);
        if(!m)
            continue;
        const uint32_t lane = threadIdx.x & 63;
        uint64_t base = 0;
        if(lane == static_cast<uint32_t>(__builtin_ctzll(m)))
            base = add64(cursor, __popcll(m));
        base = __shfl(base, __builtin_ctzll(m), 64);
        if(sk)
        {
            const uint64_t i = base + __popcll(m & ((uint64_t(1) << lane) - 1));
            out_keys[i] = ~sk;
            if constexpr(VB != 0)
                out_vals[i] = val;
        }
    }
}

// capi/Kmer_Table.hpp:203-230: KEYS (v & m nonzero; every key of a set),
// POPCOUNT, SUM of v & m.
template <int VB>
__global__ __launch_bounds__(TILE) void k_reduce(TableView T, uint32_t what, uint64_t mask, uint64_t* out)
{
    __shared__ uint64_t red[TILE / 64];
    const uint64_t main_n = T.cap * B, all = main_n + T.ov_mask + 1;
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    uint64_t acc = 0;
    for(uint64_t g = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; g < all; g += stride)
    {
        Slot<VB> s;
        if(g < main_n)
        {
            uint8_t* bk = bucket_at<VB>(T, g / B);
            if(!bk[g % B])
                continue;
            s = slot_ref<VB>(bk, g % B);
            if(!is_key(s.k[0]))
                continue;
        }
        else
        {
            s = ov_ref<VB>(T.ov, g - main_n);
            if(!s.k[0])
                continue;
        }
        if constexpr(VB == 0)
        {
            acc++;
            continue;
        }
        const uint64_t v = *s.v() & mask;
        acc += what == 0 ? (v != 0) : what == 1 ? static_cast<uint64_t>(__popcll(v)) : v;
    }
    block_add(acc, out, red);
}

// Histogram of a map's values: out[v] += number of keys with value v, for v
// in [1, max] (counter/Kmer_Set_Counter.cpp:97-120).  LDS bins per block
// when they fit (dynamic shared memory of (max + 1) words), else global
// atomics.
template <int VB, bool LDS>
__global__ __launch_bounds__(TILE) void k_histogram(TableView T, uint64_t max, uint64_t* out)
{
    extern __shared__ uint32_t bins[];
    if constexpr(LDS)
    {
        for(uint64_t i = threadIdx.x; i <= max; i += blockDim.x)
            bins[i] = 0;
        __syncthreads();
    }
    const uint64_t main_n = T.cap * B, all = main_n + T.ov_mask + 1;
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for(uint64_t g = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; g < all; g += stride)
    {
        Slot<VB> s;
        if(g < main_n)
        {
            uint8_t* bk = bucket_at<VB>(T, g / B);
            if(!bk[g % B])
                continue;
            s = slot_ref<VB>(bk, g % B);
            if(!is_key(s.k[0]))
                continue;
        }
        else
        {
            s = ov_ref<VB>(T.ov, g - main_n);
            if(!s.k[0])
                continue;
        }
        const uint64_t v = *s.v();
        if(v < 1 || v > max)
            continue;
        if constexpr(LDS)
            atomicAdd(&bins[v], 1u);
        else
            add64(out + v, 1);
    }
    if constexpr(LDS)
    {
        __syncthreads();
        for(uint64_t i = threadIdx.x; i <= max; i += blockDim.x)
            if(bins[i])
                add64(out + i, bins[i]);
    }
}

// Keys of the main table at their primary bucket, and elsewhere
// (include/Streaming_Kmer_Hash_Table.hpp:1109-1120, 2728-2752).
template <int VB>
__global__ __launch_bounds__(TILE) void k_placement(TableView T, uint64_t* out)
{
    __shared__ uint64_t red[TILE / 64];
    const uint64_t all = T.cap * B, stride = uint64_t(gridDim.x) * blockDim.x;
    uint64_t prim = 0, sec = 0;
    for(uint64_t g = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; g < all; g += stride)
    {
        uint8_t* bk = bucket_at<VB>(T, g / B);
        if(!bk[g % B] || !is_key(slot_at<VB>(bk, g % B)[0]))
            continue;
        const uint64_t h = minimizer_hash(~slot_at<VB>(bk, g % B)[0]);
        if(bucket_of(h, T.cap) == g / B)
            prim++;
        else
            sec++;
    }
    block_add(prim, out, red);
    block_add(sec, out + 1, red);
}

__global__ void k_fill(uint8_t* p, uint64_t n, uint8_t v)
{
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for(uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += stride)
        p[i] = v;
}

// Zeroes n u64 words (16-byte aligned buffers, n even or odd).
__global__ void k_zero64(uint64_t* p, uint64_t n)
{
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for(uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += stride)
        p[i] = 0;
}

// Adds up `sets` consecutive counter sets from `src` (device) into as many
// Counters at `dst` (host memory) and zeroes them: a launch's counters read
// back and cleared for the next launch by one dispatch of one wave.
__global__ __launch_bounds__(64) void k_publish(uint64_t* src, uint64_t* dst, uint64_t sets)
{
    for(uint64_t k = 0; k < sets; ++k)
        publish_set(src + k * CNT_SET_WORDS, dst + k * CNT_WORDS);
}

// Describes the valid positions of [pos_begin, pos_end) into arrays (for
// tests: the device describer checked position by position).
__global__ __launch_bounds__(TILE) void k_describe(const uint8_t* __restrict__ text, uint64_t text_len, uint64_t pos_begin, uint64_t pos_end, uint64_t* key, uint64_t* h, uint8_t* c, uint8_t* fwd, uint8_t* valid)
{
    __shared__ WaveLds wave_lds[TILE / 64];
    WaveLds& s = wave_lds[threadIdx.x >> 6];
    for_tiles(text, pos_begin, pos_end, text_len, s, [&](uint64_t tile, const Desc& d)
    {
        // Tiles start at pos_begin rounded down to 16: positions below
        // pos_begin are described but have no output slot.
        const uint64_t p = (pos_begin & ~uint64_t(15)) + tile * TILE + threadIdx.x;
        if(p >= pos_begin && p < pos_end)
        {
            const uint64_t i = p - pos_begin;
            valid[i] = static_cast<uint8_t>(d.valid);
            key[i] = d.valid ? d.key : 0;
            h[i] = d.valid ? d.h : 0;
            c[i] = d.valid ? d.c : 0;
            fwd[i] = d.valid ? static_cast<uint8_t>(d.fwd) : 0;
        }
    });
}

// The hash and placement functions of the table, evaluated on the device for
// tests: per input x, HASH_COLUMNS words: wang_u64(x), checksum(x),
// minimizer_hash(x & (2^62 - 1)), xxh3_8(x, 0), xxh3_8(x, 2^63),
// bucket_of(x, cap), filter_pos(x), pick_bucket(x, cap) and the secondaries
// p, q of x as a minimizer hash in a table of cap buckets.
constexpr uint32_t HASH_COLUMNS = 10;

__global__ __launch_bounds__(TILE) void k_hashes(const uint64_t* __restrict__ x, uint64_t n, uint64_t cap, uint64_t* __restrict__ out)
{
    for(uint64_t i = blockIdx.x * uint64_t(TILE) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * TILE)
    {
        const uint64_t v = x[i];
        uint64_t* o = out + i * HASH_COLUMNS;
        uint64_t p, q;
        secondaries(v, cap, p, q);
        o[0] = wang_u64(v);
        o[1] = checksum(v);
        o[2] = minimizer_hash(v & KMER_MASK);
        o[3] = xxh3_8(v, 0);
        o[4] = xxh3_8(v, uint64_t(1) << 63);
        o[5] = bucket_of(v, cap);
        o[6] = filter_pos(v);
        o[7] = pick_bucket(v, cap);
        o[8] = p;
        o[9] = q;
    }
}

// Segment tables of a per-sequence count over a device text, made on the
// device: sequence i is segment i (seg_q[i] = i) and its counter starts at 0;
// tile t's bracket tile_seg[t], for t in [0, tiles], is the last segment that
// starts at or before position t * TILE (0 if none does), as segment_of reads
// it.  The host builds the same tables for the per-sequence calls that stage
// them.
__global__ __launch_bounds__(TILE) void k_seg_setup(const uint64_t* __restrict__ seg_start, uint64_t ns, uint64_t tiles,
                                                   uint32_t* __restrict__ seg_q, uint32_t* __restrict__ tile_seg,
                                                   uint64_t* __restrict__ per_seq)
{
    const uint64_t stride = uint64_t(gridDim.x) * TILE;
    for(uint64_t i = blockIdx.x * uint64_t(TILE) + threadIdx.x; i < ns; i += stride)
        seg_q[i] = static_cast<uint32_t>(i), per_seq[i] = 0;
    for(uint64_t t = blockIdx.x * uint64_t(TILE) + threadIdx.x; t <= tiles; t += stride)
    {
        const uint64_t p = t * TILE;
        uint64_t lo = 0, hi = ns - 1;
        while(lo < hi)
        {
            const uint64_t mid = (lo + hi + 1) >> 1;
            if(seg_start[mid] <= p)
                lo = mid;
            else
                hi = mid - 1;
        }
        tile_seg[t] = static_cast<uint32_t>(lo);
    }
}

}
