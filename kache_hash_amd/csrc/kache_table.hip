// Host side of the MI355X k-mer table: the C ABI of include/kache_capi.h and
// the device extensions of include/kache_amd.h.
//
// A table owns its HBM bucket store, its overflow table, a HIP stream and
// staging buffers.  Each call stages its sequences (pinned host -> HBM),
// launches the fused describe-and-probe kernels, and reads back results.
// Growth follows the reference's schedule exactly (per-user flush steps,
// include/Streaming_Kmer_Hash_Table.hpp:2244-2286), so capacities match it.

#include "kache_capi.h"
#include "kache_amd.h"
#include "kache_kernels.hpp"

#include <algorithm>
#include <cerrno>
#include <cstddef>
#include <cstdio>
#include <memory>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <future>
#include <map>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <chrono>
#include <vector>
#include <condition_variable>
#include <functional>
#include <thread>

using namespace kamd;

namespace
{

thread_local std::string last_error;

struct hip_failure : std::runtime_error
{
    using std::runtime_error::runtime_error;
};

void check(hipError_t e, const char* what)
{
    if(e == hipSuccess)
        return;
    if(e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
    {
        last_error = std::string(what) + ": out of device memory";
        (void)hipGetLastError();
        throw std::bad_alloc();
    }
    throw hip_failure(std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP(call) check((call), #call)

// Runs `f`, returning its status, or that of the exception it throws
// (capi/kache_capi.cpp:14-29).
template <typename F>
int guard(const F& f) noexcept
{
    try
    {
        return f();
    }
    catch(const std::bad_alloc&)
    {
        if(last_error.empty())
            last_error = "out of memory";
        return KACHE_ENOMEM;
    }
    catch(const std::exception& e)
    {
        last_error = e.what();
        return KACHE_EINTERNAL;
    }
    catch(...)
    {
        last_error = "unknown error";
        return KACHE_EINTERNAL;
    }
}

constexpr uint64_t MAX_CAP = (uint64_t(1) << 32) - 1;   // include/Streaming_Kmer_Hash_Table.hpp:84
constexpr uint64_t OV_MIN = 1024;                       // :67
constexpr uint64_t LAUNCH_POS = uint64_t(1) << 24;      // positions per insert launch
constexpr uint64_t STAGE_BYTES = uint64_t(1) << 28;     // host sequence bytes staged at once
constexpr uint64_t PIPE_BYTES = uint64_t(1) << 24;      // chunk of the double-buffered host pipeline
constexpr uint64_t COLLECT_BUCKETS = uint64_t(1) << 18; // buckets per iteration copy-out

uint64_t ceil_pow2(uint64_t x)
{
    uint64_t p = 1;
    while(p < x)
        p <<= 1;
    return p;
}

// include/Streaming_Kmer_Hash_Table.hpp:1003-1011.
uint64_t cap_for(uint64_t n, double lf, bool pow2)
{
    uint64_t cap = std::max<uint64_t>(1, static_cast<uint64_t>(std::ceil(static_cast<double>(n) / lf / B)));
    while(static_cast<uint64_t>(static_cast<double>(cap * B) * lf) <= n)
        cap++;
    return pow2 ? ceil_pow2(cap) : cap;
}

// :799-804.
bool size_fits(uint64_t n, double lf, bool pow2)
{
    return static_cast<double>(n) / lf / B <= static_cast<double>(MAX_CAP) && cap_for(n, lf, pow2) <= MAX_CAP;
}

// :383.
uint64_t step_for(uint64_t th) { return std::clamp<uint64_t>(th / (8 * KACHE_MAX_USERS), 1, 16384); }

// Slots of an overflow table with room for n keys at load 0.8
// (include/Concurrent_Hash_Table.hpp:161).
uint64_t ov_slots_for(uint64_t n) { return ceil_pow2(static_cast<uint64_t>(std::ceil(static_cast<double>(n) / 0.8))); }

template <typename T>
T* dev_alloc(uint64_t n)
{
    void* p = nullptr;
    if(n == 0)
        n = 1;
    HIP(hipMalloc(&p, n * sizeof(T)));
    return static_cast<T*>(p);
}

int device_cus()
{
    static int cus = [](){
        int dev = 0, n = 0;
        if(hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            return 256;
        return n > 0 ? n : 256;
    }();
    return cus;
}

// A growable device buffer.
struct DevBuf
{
    void* p = nullptr;
    uint64_t bytes = 0;
    void reserve(uint64_t n)
    {
        if(n <= bytes)
            return;
        if(p)
            HIP(hipFree(p)), p = nullptr, bytes = 0;
        n = std::max<uint64_t>(n, 4096);
        HIP(hipMalloc(&p, n));
        bytes = n;
    }
    ~DevBuf() { if(p) (void)hipFree(p); }
    template <typename T> T* as() const { return static_cast<T*>(p); }
};

// A growable pinned host buffer.
struct HostBuf
{
    void* p = nullptr;
    uint64_t bytes = 0;
    void reserve(uint64_t n)
    {
        if(n <= bytes)
            return;
        if(p)
            HIP(hipHostFree(p)), p = nullptr, bytes = 0;
        n = std::max<uint64_t>(n, 4096);
        HIP(hipHostMalloc(&p, n, hipHostMallocDefault));
        bytes = n;
    }
    ~HostBuf() { if(p) (void)hipHostFree(p); }
    template <typename T> T* as() const { return static_cast<T*>(p); }
};

// Device arrays in reserved address ranges backed by physical chunks (HIP
// virtual memory).  A range is mapped from its start; it can be mapped
// further, in place, up to its reservation, so a bucket store doubles without
// a second copy beside it.  Physical memory is taken in chunks of at most
// CHUNK bytes (a few mappings per table: each mapping costs the host ~50 us).
// Ranges are also what the array cache hands out for large arrays: random
// 64-bit atomics measured 20.0 G lines/s on such memory against 17.3 G/s on
// hipMalloc memory of the same size (random reads equal, 48 G lines/s).
class VmSpace
{
public:
    static constexpr uint64_t GRAIN = uint64_t(2) << 20;
    static constexpr uint64_t CHUNK = uint64_t(1) << 30;

private:
    struct Range
    {
        int dev = 0;
        uint64_t reserved = 0, mapped = 0;
        std::vector<std::pair<hipMemGenericAllocationHandle_t, uint64_t>> chunks;
    };
    std::mutex mu;
    std::map<const void*, Range> ranges;
    std::map<int, int> support;

public:
    // A physical chunk not yet mapped.
    struct Chunk
    {
        hipMemGenericAllocationHandle_t h;
        uint64_t n;
    };

private:
    static hipMemAllocationProp prop(int dev)
    {
        hipMemAllocationProp p = {};
        p.type = hipMemAllocationTypePinned;
        p.location.type = hipMemLocationTypeDevice;
        p.location.id = dev;
        return p;
    }

    // Maps [r.mapped, bytes) of the range at `base`, then grants the device
    // access to the whole mapped range in one call (access granted chunk by
    // chunk, or to the new part alone, is refused with "invalid argument"
    // once the range holds chunks of different sizes).  On failure the range
    // is left as it was and the error thrown.
    static void map_to(void* base, Range& r, uint64_t bytes, std::vector<Chunk>* ready = nullptr)
    {
        const hipMemAllocationProp pr = prop(r.dev);
        hipMemAccessDesc acc = {};
        acc.location = pr.location;
        acc.flags = hipMemAccessFlagsProtReadWrite;
        const size_t first = r.chunks.size();
        const uint64_t from = r.mapped;
        try
        {
            while(r.mapped < bytes)
            {
                const uint64_t n = std::min(CHUNK, bytes - r.mapped);
                hipMemGenericAllocationHandle_t h;
                if(ready && !ready->empty() && ready->front().n == n)
                    h = ready->front().h, ready->erase(ready->begin());
                else
                    check(hipMemCreate(&h, n, &pr, 0), "hipMemCreate");
                const hipError_t e = hipMemMap(static_cast<char*>(base) + r.mapped, n, 0, h, 0);
                if(e != hipSuccess)
                {
                    (void)hipMemRelease(h);
                    check(e, "hipMemMap");
                }
                r.chunks.push_back({h, n});
                r.mapped += n;
            }
            check(hipMemSetAccess(base, r.mapped, &acc, 1), "hipMemSetAccess");
        }
        catch(...)
        {
            uint64_t off = from;
            for(size_t i = first; i < r.chunks.size(); ++i)
            {
                (void)hipMemUnmap(static_cast<char*>(base) + off, r.chunks[i].second);
                (void)hipMemRelease(r.chunks[i].first);
                off += r.chunks[i].second;
            }
            r.chunks.resize(first);
            r.mapped = from;
            throw;
        }
    }

    static void unmap_all(void* base, Range& r)
    {
        uint64_t off = 0;
        for(const auto& [h, n] : r.chunks)
        {
            (void)hipMemUnmap(static_cast<char*>(base) + off, n);
            (void)hipMemRelease(h);
            off += n;
        }
        r.chunks.clear();
        r.mapped = 0;
    }

public:
    static VmSpace& get()
    {
        static VmSpace* s = new VmSpace;     // lives as long as the process
        return *s;
    }

    static uint64_t rounded(uint64_t b) { return (std::max<uint64_t>(b, 1) + GRAIN - 1) / GRAIN * GRAIN; }

    bool supported(int dev)
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = support.find(dev);
        if(it == support.end())
        {
            int v = 0;
            if(hipDeviceGetAttribute(&v, hipDeviceAttributeVirtualMemoryManagementSupported, dev) != hipSuccess)
                v = 0, (void)hipGetLastError();
            it = support.emplace(dev, v).first;
        }
        return it->second != 0;
    }

    // A range of `reserve` bytes (at least `bytes`) on the current device,
    // which must be `dev`, with its first `bytes` mapped.
    void* make(int dev, uint64_t bytes, uint64_t reserve)
    {
        bytes = rounded(bytes), reserve = std::max(rounded(reserve), bytes);
        void* base = nullptr;
        if(hipMemAddressReserve(&base, reserve, GRAIN, nullptr, 0) != hipSuccess)
        {
            // No room for the growth reservation: a range of the array alone.
            (void)hipGetLastError();
            reserve = bytes;
            check(hipMemAddressReserve(&base, reserve, GRAIN, nullptr, 0), "hipMemAddressReserve");
        }
        Range r;
        r.dev = dev, r.reserved = reserve;
        try
        {
            map_to(base, r, bytes);
        }
        catch(...)
        {
            unmap_all(base, r);
            (void)hipMemAddressFree(base, reserve);
            throw;
        }
        std::lock_guard<std::mutex> lk(mu);
        ranges[base] = std::move(r);
        return base;
    }

    bool owns(const void* p)
    {
        std::lock_guard<std::mutex> lk(mu);
        return ranges.count(p) != 0;
    }

    uint64_t reserved(const void* p)
    {
        std::lock_guard<std::mutex> lk(mu);
        const auto it = ranges.find(p);
        return it == ranges.end() ? 0 : it->second.reserved;
    }

    uint64_t mapped(const void* p)
    {
        std::lock_guard<std::mutex> lk(mu);
        const auto it = ranges.find(p);
        return it == ranges.end() ? 0 : it->second.mapped;
    }

    // Physical chunks for mapping a range from `from` to `to` bytes, in the
    // sizes map_to takes them: creating them is what costs (the driver clears
    // memory it hands out), so a helper thread makes them ahead of a doubling.
    static std::vector<Chunk> make_chunks(int dev, uint64_t from, uint64_t to)
    {
        const hipMemAllocationProp pr = prop(dev);
        std::vector<Chunk> out;
        try
        {
            for(uint64_t at = from; at < to;)
            {
                const uint64_t n = std::min(CHUNK, to - at);
                hipMemGenericAllocationHandle_t h;
                check(hipMemCreate(&h, n, &pr, 0), "hipMemCreate");
                out.push_back({h, n});
                at += n;
            }
        }
        catch(...)
        {
            release_chunks(out);
            throw;
        }
        return out;
    }

    static void release_chunks(std::vector<Chunk>& c)
    {
        for(const Chunk& k : c)
            (void)hipMemRelease(k.h);
        c.clear();
    }

    // Maps the range at `p` up to `bytes` (rounded; at most its reservation),
    // with chunks from `ready` where they fit (those used are taken out).
    void grow(void* p, uint64_t bytes, std::vector<Chunk>* ready = nullptr)
    {
        std::unique_lock<std::mutex> lk(mu);
        const auto it = ranges.find(p);
        if(it == ranges.end() || rounded(bytes) > it->second.reserved)
            throw hip_failure("growing an array past its address reservation");
        Range r = it->second;
        lk.unlock();
        // Only the range's owner maps it: nothing else changes it meanwhile.
        // A failure leaves the range as it was.
        map_to(p, r, rounded(bytes), ready);
        lk.lock();
        ranges[p] = std::move(r);
    }

    // Unmaps the chunks of the range at `p` that lie wholly past its first
    // `bytes`, returning their memory to the device; returns the bytes left
    // mapped.
    uint64_t shrink(void* p, uint64_t bytes)
    {
        std::lock_guard<std::mutex> lk(mu);
        const auto it = ranges.find(p);
        if(it == ranges.end())
            return 0;
        Range& r = it->second;
        while(!r.chunks.empty() && r.mapped - r.chunks.back().second >= bytes)
        {
            const auto [h, n] = r.chunks.back();
            (void)hipMemUnmap(static_cast<char*>(p) + r.mapped - n, n);
            (void)hipMemRelease(h);
            r.mapped -= n;
            r.chunks.pop_back();
        }
        return r.mapped;
    }

    void release(void* p)
    {
        Range r;
        {
            std::lock_guard<std::mutex> lk(mu);
            const auto it = ranges.find(p);
            if(it == ranges.end())
                return;
            r = std::move(it->second);
            ranges.erase(it);
        }
        unmap_all(p, r);
        (void)hipMemAddressFree(p, r.reserved);
    }
};

#ifndef KAMD_VMM_MIN_BYTES
#define KAMD_VMM_MIN_BYTES (uint64_t(2) << 20)
#endif

// Frees an array of the array cache (a range or a hipMalloc allocation).
void dev_free(void* p)
{
    if(!p)
        return;
    if(VmSpace::get().owns(p))
        VmSpace::get().release(p);
    else
        (void)hipFree(p);
}

// Large device arrays (bucket stores, overflow tables, doubling away lists) are
// kept when a table frees them and handed out again for an array of the same
// size on the same device, so a new table of the same shape, and the same
// growth steps, reuse them.  The driver wipes device memory it gets back
// before it hands it out again; at tens of gigabytes per table that wipe, not
// the table's own work, would bound a program that makes and drops tables.
// The cache holds at most half of a device's memory; an allocation that finds
// the device full releases the device's cached arrays first, and
// kache_amd_trim_cache releases them on request.  Arrays are handed back only
// after the work that used them has ended (their table's stream synchronized).
class ArrayCache
{
    std::mutex mu;
    std::multimap<std::pair<int, uint64_t>, void*> kept;
    std::map<int, uint64_t> kept_bytes, limits;
    static constexpr uint64_t GRAIN = uint64_t(2) << 20;

    static uint64_t rounded(uint64_t b) { return (std::max<uint64_t>(b, 1) + GRAIN - 1) / GRAIN * GRAIN; }

    uint64_t limit(int dev)
    {
        auto it = limits.find(dev);
        if(it != limits.end())
            return it->second;
        size_t total = 0;
        const uint64_t l = hipDeviceTotalMem(&total, dev) == hipSuccess ? total / 2 : 0;
        limits[dev] = l;
        return l;
    }

public:
    static ArrayCache& get()
    {
        static ArrayCache* c = new ArrayCache;     // lives as long as the process: no frees after the runtime's teardown
        return *c;
    }

    // An array of at least `bytes` on the current device, which must be `dev`.
    void* take(int dev, uint64_t bytes)
    {
        const uint64_t r = rounded(bytes);
        {
            std::lock_guard<std::mutex> lk(mu);
            const auto it = kept.find({dev, r});
            if(it != kept.end())
            {
                void* p = it->second;
                kept.erase(it);
                kept_bytes[dev] -= r;
                return p;
            }
        }
        if(r >= KAMD_VMM_MIN_BYTES && VmSpace::get().supported(dev))
        {
            try
            {
                return VmSpace::get().make(dev, r, r);
            }
            catch(const std::bad_alloc&)
            {
                if(trim(dev) == 0)
                    throw;
            }
            return VmSpace::get().make(dev, r, r);
        }
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, r);
        if((e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) && trim(dev) > 0)
        {
            (void)hipGetLastError();
            e = hipMalloc(&p, r);
        }
        check(e, "hipMalloc");
        return p;
    }

    // A bucket store that can double in place: an address range reserving
    // `reserve` bytes with at least `bytes` mapped.  A kept range whose
    // mapping covers `bytes` (and is at most 16 times it) and whose
    // reservation lets it double three times is handed out again, the
    // smallest first: the next table of a program that makes and drops tables
    // then grows into memory that is already mapped.
    void* take_range(int dev, uint64_t bytes, uint64_t reserve)
    {
        const uint64_t r = rounded(bytes);
        {
            std::lock_guard<std::mutex> lk(mu);
            for(auto it = kept.lower_bound({dev, r}); it != kept.end() && it->first.first == dev && it->first.second <= 16 * r; ++it)
                if(VmSpace::get().reserved(it->second) >= 8 * r)
                {
                    void* p = it->second;
                    kept_bytes[dev] -= it->first.second;
                    kept.erase(it);
                    return p;
                }
        }
        try
        {
            return VmSpace::get().make(dev, r, reserve);
        }
        catch(const std::bad_alloc&)
        {
            if(trim(dev) == 0)
                throw;
        }
        return VmSpace::get().make(dev, r, reserve);
    }

    // Keeps an array that take() or take_range() gave for `bytes` on device
    // `dev` (a range: what is mapped of it), if the cache has room for it;
    // returns whether it did.
    bool keep(int dev, void* p, uint64_t bytes)
    {
        const uint64_t mapped = VmSpace::get().mapped(p);
        const uint64_t r = rounded(mapped ? mapped : bytes);
        std::lock_guard<std::mutex> lk(mu);
        if(kept_bytes[dev] + r > limit(dev))
            return false;
        kept.insert({{dev, r}, p});
        kept_bytes[dev] += r;
        return true;
    }

    // Keeps the array, or frees it when the cache is full.
    void give(int dev, void* p, uint64_t bytes)
    {
        if(p != nullptr && !keep(dev, p, bytes))
            dev_free(p);
    }

    // Frees the device's cached arrays; returns their bytes.
    uint64_t trim(int dev)
    {
        std::vector<void*> out;
        uint64_t n = 0;
        {
            std::lock_guard<std::mutex> lk(mu);
            for(auto it = kept.begin(); it != kept.end();)
                if(it->first.first == dev)
                    out.push_back(it->second), n += it->first.second, it = kept.erase(it);
                else
                    ++it;
            kept_bytes[dev] = 0;
        }
        for(void* p : out)
            dev_free(p);
        return n;
    }

    uint64_t bytes_kept(int dev)
    {
        std::lock_guard<std::mutex> lk(mu);
        return kept_bytes[dev];
    }
};

int current_device()
{
    int dev = 0;
    HIP(hipGetDevice(&dev));
    return dev;
}

// Host threads that pack sequences into the pinned staging buffers: one
// packing thread copies at a fraction of the link's speed, so large stages
// are copied by several.  One job at a time (callers queue on `run_mu`);
// the calling thread works too.
class CopyPool
{
    std::vector<std::thread> workers;
    std::mutex mu, run_mu;
    std::condition_variable cv_job, cv_done;
    const std::function<void(size_t)>* job = nullptr;
    size_t n_tasks = 0, next = 0, finished = 0;
    uint64_t gen = 0;
    bool stop = false;

    // Runs tasks of the current job until none is left; returns with `lk` held.
    void drain(std::unique_lock<std::mutex>& lk)
    {
        while(next < n_tasks)
        {
            const size_t i = next++;
            const auto* f = job;
            lk.unlock();
            (*f)(i);
            lk.lock();
            if(++finished == n_tasks)
                cv_done.notify_all();
        }
    }

    explicit CopyPool(unsigned n) : threads(n + 1)
    {
        for(unsigned i = 0; i < n; ++i)
            workers.emplace_back([this]
            {
                std::unique_lock<std::mutex> lk(mu);
                uint64_t seen = 0;
                while(true)
                {
                    cv_job.wait(lk, [&]{ return stop || gen != seen; });
                    if(stop)
                        return;
                    seen = gen;
                    drain(lk);
                }
            });
    }

public:
    static constexpr unsigned THREADS = 8;
    const unsigned threads;     // the workers and the calling thread

    ~CopyPool()
    {
        {
            std::lock_guard<std::mutex> lk(mu);
            stop = true;
        }
        cv_job.notify_all();
        for(auto& w : workers)
            w.join();
    }

    static CopyPool& get()
    {
        static CopyPool pool(THREADS - 1);
        return pool;
    }

    // A second, smaller pool for work done ahead (the serving loop's stager
    // packing the next wave), so that it never holds up the packing of a
    // call in progress: jobs of one pool run one at a time.
    static CopyPool& aside()
    {
        static CopyPool pool(3);
        return pool;
    }

    void run(size_t n, const std::function<void(size_t)>& f)
    {
        if(n <= 1)
        {
            for(size_t i = 0; i < n; ++i)
                f(i);
            return;
        }
        std::lock_guard<std::mutex> one(run_mu);
        std::unique_lock<std::mutex> lk(mu);
        job = &f, n_tasks = n, next = 0, finished = 0, gen++;
        cv_job.notify_all();
        drain(lk);
        cv_done.wait(lk, [&]{ return finished == n_tasks; });
        job = nullptr;
    }
};

// Whether p lies in page-locked host memory (hipHostMalloc, as
// kache_amd_host_alloc gives), which the copy engines reach directly.
static bool pinned_host(const void* p)
{
    hipPointerAttribute_t a{};
    if(hipPointerGetAttributes(&a, p) != hipSuccess)
    {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeHost;
}

// Copies n bytes, in parallel when large.
void par_copy(char* dst, const char* src, uint64_t n, CopyPool& pool = CopyPool::get())
{
    constexpr uint64_t PIECE = uint64_t(2) << 20;
    if(n < 2 * PIECE)
    {
        std::memcpy(dst, src, n);
        return;
    }
    const uint64_t parts = std::min<uint64_t>(pool.threads, n / PIECE);
    const uint64_t each = (n + parts - 1) / parts;
    pool.run(parts, [&](size_t i)
    {
        const uint64_t a = i * each, b = std::min(n, a + each);
        if(a < b)
            std::memcpy(dst + a, src + a, b - a);
    });
}

// Sequences staged for one device pass: segments of the caller's sequences,
// each followed by a separator byte (include/Streaming_Kmer_Hash_Table.hpp:1697-1709
// splits at non-ACGT characters, so the separator splits too).
struct Stage
{
    uint64_t text_len = 0;
    std::vector<uint64_t> seg_start;    // n_seg + 1
    std::vector<uint32_t> seg_q;
    std::vector<uint64_t> seg_off;
};

// Cursor over the caller's sequences: sequence q from its character `off`.
struct SeqCursor
{
    size_t q = 0;
    uint64_t off = 0;
};

}


// ---------------------------------------------------------------------------
// Scratch memory of the allocation-free per-sequence count
// (kache_amd_count_seqs_device_static): per-sequence counters, segment starts
// and owners, and tile brackets for at most STATIC_MAX_SEQS sequences over at
// most STATIC_MAX_TEXT bytes of device text.  One copy per device (about
// 36 MiB), held until the process ends.  Host code never uses the array's
// name as an address: each device's copy is found once, at the first table's
// set-up on that device, by hipGetSymbolAddress, and that address is what the
// kernels and copies take.

constexpr uint64_t STATIC_MAX_SEQS = uint64_t(1) << 20;
constexpr uint64_t STATIC_MAX_TEXT = uint64_t(1) << 30;
constexpr uint64_t SCR_PER_SEQ = 0;
constexpr uint64_t SCR_SEG_START = SCR_PER_SEQ + STATIC_MAX_SEQS * 8;
constexpr uint64_t SCR_SEG_Q = SCR_SEG_START + (STATIC_MAX_SEQS + 1) * 8;
constexpr uint64_t SCR_TILE_SEG = SCR_SEG_Q + STATIC_MAX_SEQS * 4;
constexpr uint64_t SCRATCH_BYTES = (SCR_TILE_SEG + (STATIC_MAX_TEXT / TILE + 2) * 4 + 255) / 256 * 256;

static __device__ __align__(256) uint8_t scratch[SCRATCH_BYTES];

namespace
{

struct DeviceScratch
{
    uint8_t* ptr = nullptr;     // this device's copy of `scratch`
    std::mutex calls;           // serializes the calls that use it
};

std::mutex scratch_mu;
std::map<int, std::unique_ptr<DeviceScratch>> scratch_of_device;

// The scratch of device `dev`, which must be the current device: its
// address is resolved on the first call for the device and kept.
DeviceScratch& device_scratch(int dev)
{
    std::lock_guard<std::mutex> lk(scratch_mu);
    auto& d = scratch_of_device[dev];
    if(!d)
    {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if(e != hipSuccess || p == nullptr)
        {
            (void)hipGetLastError();
            throw hip_failure(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        }
        auto n = std::make_unique<DeviceScratch>();
        n->ptr = static_cast<uint8_t*>(p);
        d = std::move(n);
    }
    return *d;
}

}

// ---------------------------------------------------------------------------
// The table.

struct kache_table
{
    const uint32_t value_bytes;
    std::mutex reg;                         // serializes registrations (capi/kache_table.hpp:22-23)
    std::atomic<uint32_t> user_c{0};

    mutable std::mutex mu;                  // serializes device work
    hipStream_t stream = nullptr;

    const double lf;
    const bool map;
    const uint32_t vw;                      // value bytes per slot: 0 (a set), 4 (widths 1-4), 8 (widths 5-8)
    const int device;                       // the device current at creation: the table's arrays live there
    DeviceScratch* dscratch = nullptr;      // that device's scratch (allocation-free calls)

    uint64_t cap = 0;                       // buckets
    uint8_t* buckets = nullptr;
    uint64_t* ov = nullptr;
    uint64_t ov_slots = 0;
    uint64_t ov_count = 0;
    uint64_t size = 0;                      // keys
    uint64_t resizes = 0;

    // The reference's growth schedule (include/Streaming_Kmer_Hash_Table.hpp:224-237).
    uint64_t approx_sz = 0;
    uint64_t resize_th = 0;
    uint64_t step = 1;
    uint64_t checkpoint[KACHE_MAX_USERS] = {};

    // Staging and scratch.
    mutable HostBuf h_text, h_misc;
    HostBuf h_away;                         // an in-place doubling's cursors and fail word
    // Double-buffered staging of host sequences: chunk i + 1 is packed and
    // copied (on copy_stream) while chunk i's kernels run.
    mutable HostBuf h_pipe[2];
    mutable DevBuf d_pipe[2];
    mutable hipStream_t copy_stream = nullptr;
    mutable hipEvent_t ev_copied[2] = {}, ev_freed[2] = {};
    // Double-buffered download of large outputs to the caller's memory.
    mutable HostBuf h_dl[2];
    mutable hipEvent_t ev_dl[2] = {};
    mutable DevBuf d_text, d_misc, d_out, d_out2, d_out3;
    mutable DevBuf d_bcnt, d_goff;     // per-batch counters of count_batches, sequence offsets of lookup_groups
    mutable Counters* d_cnt = nullptr;
    mutable Counters* h_cnt = nullptr;
    uint32_t blocks_per_cu = 0;
    mutable std::map<const void*, int> resident;   // blocks per CU at full occupancy, per kernel

    // Profiling.
    bool profiling = false;
    mutable std::map<std::string, std::pair<double, uint64_t>> prof;
    mutable std::vector<hipEvent_t> ev_pool;
    mutable std::vector<std::pair<const char*, size_t>> ev_pending;
    mutable size_t ev_used = 0;

    kache_table(const kache_options& o)
        : value_bytes(o.value_bytes), lf(o.load_factor), map(o.value_bytes > 0), vw(value_word(o.value_bytes)), device(current_device())
    {
        HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        HIP(hipStreamCreateWithFlags(&copy_stream, hipStreamNonBlocking));
        for(int i = 0; i < 2; ++i)
        {
            HIP(hipEventCreateWithFlags(&ev_copied[i], hipEventDisableTiming));
            HIP(hipEventCreateWithFlags(&ev_freed[i], hipEventDisableTiming));
            HIP(hipEventCreateWithFlags(&ev_dl[i], hipEventDisableTiming));
        }

        dscratch = &device_scratch(device);
        d_cnt = reinterpret_cast<Counters*>(dev_alloc<uint64_t>(CNT_SET_WORDS));
        HIP(hipHostMalloc(reinterpret_cast<void**>(&h_cnt), sizeof(Counters), hipHostMallocDefault));
        zero_counters();
        try
        {
            cap = cap_for(o.expected_keys, lf, o.pow2_capacity != 0);
            resize_th = static_cast<uint64_t>(static_cast<double>(cap * B) * lf);
            step = step_for(resize_th);
            buckets = alloc_buckets(cap);
            ov_slots = ov_slots_for(std::max<uint64_t>(OV_MIN, cap));
            ov = alloc_ov(ov_slots);
            HIP(hipStreamSynchronize(stream));
        }
        catch(...)
        {
            release();
            throw;
        }
    }

    ~kache_table() { release(); }

    void release()
    {
        wait_frees();
        if(prepared.valid())
            { Prepared p = prepared.get(); free_prepared(p); }
        if(stream)
            (void)hipStreamSynchronize(stream);
        free_graphs();
        if(buckets)
            ArrayCache::get().give(device, buckets, cap * bucket_bytes()), buckets = nullptr;
        if(ov)
            ArrayCache::get().give(device, ov, ov_bytes(ov_slots)), ov = nullptr;
        if(d_cnt)
            (void)hipFree(d_cnt), d_cnt = nullptr;
        if(h_cnt)
            (void)hipHostFree(h_cnt), h_cnt = nullptr;
        for(auto e : ev_pool)
            (void)hipEventDestroy(e);
        if(d_ring)
            (void)hipFree(d_ring), d_ring = nullptr;
        if(h_ring)
            (void)hipHostFree(h_ring), h_ring = nullptr;
        if(h_seq)
            (void)hipHostFree(h_seq), h_seq = nullptr;
        for(auto& e : ring_ev)
            if(e)
                (void)hipEventDestroy(e), e = nullptr;
        ev_pool.clear();
        if(copy_stream)
            (void)hipStreamSynchronize(copy_stream), (void)hipStreamDestroy(copy_stream), copy_stream = nullptr;
        for(int i = 0; i < 2; ++i)
        {
            if(ev_copied[i])
                (void)hipEventDestroy(ev_copied[i]), ev_copied[i] = nullptr;
            if(ev_freed[i])
                (void)hipEventDestroy(ev_freed[i]), ev_freed[i] = nullptr;
            if(ev_dl[i])
                (void)hipEventDestroy(ev_dl[i]), ev_dl[i] = nullptr;
        }
        if(stream)
            (void)hipStreamDestroy(stream), stream = nullptr;
    }

    static uint32_t bucket_bytes_of(uint32_t vw) { return vw == 0 ? Geo<0>::bucket_bytes : vw == 4 ? Geo<4>::bucket_bytes : Geo<8>::bucket_bytes; }
    uint32_t bucket_bytes() const { return bucket_bytes_of(vw); }
    uint64_t vmax() const { return value_bytes == 8 ? ~uint64_t(0) : value_bytes == 0 ? 0 : (uint64_t(1) << (8 * value_bytes)) - 1; }

    // Overflow slots come in blocks of 32 (at least OV_MIN of them).
    uint64_t ov_bytes(uint64_t slots) const { return slots * (8 + vw); }

    // Calls f with the table's value bytes per slot as a compile-time
    // constant (std::integral_constant<int, 0, 4 or 8>): the kernels'
    // layout parameter.
    template <typename F>
    decltype(auto) by_layout(F&& f) const
    {
        switch(vw)
        {
            case 0: return f(std::integral_constant<int, 0>{});
            case 4: return f(std::integral_constant<int, 4>{});
            default: return f(std::integral_constant<int, 8>{});
        }
    }

    // Bucket and overflow arrays come from the array cache and go back to it.
    // A store that can double in place comes as a range with room to grow.
    uint8_t* take_buckets(uint64_t n) const
    {
        const uint64_t bytes = n * bucket_bytes();
        if(growable(bytes))
            return static_cast<uint8_t*>(ArrayCache::get().take_range(device, bytes, growth_reserve(bytes)));
        return static_cast<uint8_t*>(ArrayCache::get().take(device, bytes));
    }

    uint8_t* alloc_buckets(uint64_t n) const
    {
        uint8_t* p = take_buckets(n);
        zero(p, n * bucket_bytes());
        return p;
    }

    uint64_t* alloc_ov(uint64_t slots) const
    {
        uint64_t* p = static_cast<uint64_t*>(ArrayCache::get().take(device, ov_bytes(slots)));
        zero(p, ov_bytes(slots));
        return p;
    }

    void zero(void* p, uint64_t bytes) const { HIP(hipMemsetAsync(p, 0, bytes, stream)); }

    TableView view() const { return view_of(buckets, cap, ov, ov_slots); }

    TableView view_of(uint8_t* b, uint64_t c, uint64_t* o, uint64_t os) const
    {
        TableView v;
        v.buckets = b, v.cap = c, v.ov = o, v.ov_mask = os - 1, v.vmax = vmax(), v.width = value_bytes;
        return v;
    }

    // Blocks of a persistent launch: as many as stay resident, capped by the work.
    template <typename Kern>
    uint32_t grid_for(Kern kern, uint64_t work_blocks) const
    {
        return grid_for(reinterpret_cast<const void*>(kern), work_blocks);
    }

    uint32_t grid_for(const void* kern, uint64_t work_blocks, uint32_t threads = TILE) const
    {
        int per_cu = static_cast<int>(blocks_per_cu);
        if(per_cu == 0)
        {
            const auto it = resident.find(kern);
            if(it != resident.end())
                per_cu = it->second;
            else
            {
                if(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, static_cast<int>(threads), 0) != hipSuccess || per_cu < 1)
                    per_cu = 4;
                resident[kern] = per_cu;
            }
        }
        const uint64_t g = static_cast<uint64_t>(per_cu) * device_cus();
        return static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>(g, work_blocks)));
    }

    // Launches `launch()` and, when profiling, brackets it with events from a
    // pool; their times are summed at the next synchronization of the stream.
    template <typename F>
    void timed(const char* name, const F& launch) const
    {
        if(!profiling)
        {
            launch();
            HIP(hipGetLastError());
            return;
        }
        if(ev_used + 2 > ev_pool.size())
            for(int i = 0; i < 2; ++i)
            {
                hipEvent_t e;
                HIP(hipEventCreate(&e));
                ev_pool.push_back(e);
            }
        hipEvent_t a = ev_pool[ev_used], b = ev_pool[ev_used + 1];
        HIP(hipEventRecord(a, stream));
        launch();
        HIP(hipGetLastError());
        HIP(hipEventRecord(b, stream));
        ev_pending.push_back({name, ev_used});
        ev_used += 2;
    }

    // Synchronizes the stream, then sums the times of the launches profiled
    // since the last synchronization.
    void sync() const
    {
        HIP(hipStreamSynchronize(stream));
        for(const auto& [name, i] : ev_pending)
        {
            float ms = 0;
            HIP(hipEventElapsedTime(&ms, ev_pool[i], ev_pool[i + 1]));
            auto& e = prof[name];
            e.first += ms, e.second++;
        }
        ev_pending.clear();
        ev_used = 0;
    }

    void zero_counters() const
    {
        hipLaunchKernelGGL(k_zero64, dim3(4), dim3(256), 0, stream, reinterpret_cast<uint64_t*>(d_cnt), CNT_SET_WORDS);
        HIP(hipGetLastError());
    }

    // Enqueues the read-back of `n` counter sets of the launches so far into
    // `out` (host memory, n Counters) and their zeroing for the next launches.
    void publish_counters(Counters* dev, Counters* out, uint64_t n = 1) const
    {
        hipLaunchKernelGGL(k_publish, dim3(1), dim3(64), 0, stream, reinterpret_cast<uint64_t*>(dev), reinterpret_cast<uint64_t*>(out), n);
        HIP(hipGetLastError());
    }

    // Reads the counters of the launches since the last read, and zeroes them
    // again for the next (they are zero whenever no launch is in flight).
    Counters read_counters() const
    {
        publish_counters(d_cnt, h_cnt);
        sync();
        return *h_cnt;
    }

    // --- Overflow growth (include/Streaming_Kmer_Hash_Table.hpp:1412-1443) ---

    void rehash_ov(uint64_t new_slots)
    {
        check_budget(ov_bytes(new_slots), "growing the overflow table");
        uint64_t* nov = alloc_ov(new_slots);
        try
        {
            const TableView N = view_of(buckets, cap, nov, new_slots);
            const uint64_t blocks = (ov_slots + TILE - 1) / TILE;
            timed("ov_rehash", [&]{
                by_layout([&](auto L){
                    const auto kern = k_ov_rehash<decltype(L)::value>;
                    hipLaunchKernelGGL(kern, dim3(grid_for(kern, blocks)), dim3(TILE), 0, stream, ov, ov_slots, N, d_cnt);
                });
            });
            const Counters c = read_counters();
            if(c.failed)
                throw hip_failure("overflow rehash found no room");
        }
        catch(...)
        {
            // The old overflow table is untouched: the table stays as it was.
            (void)hipStreamSynchronize(stream);
            ArrayCache::get().give(device, nov, ov_bytes(new_slots));
            throw;
        }
        ArrayCache::get().give(device, ov, ov_bytes(ov_slots));
        ov = nov, ov_slots = new_slots;
    }

    // --- Device-memory budget ---

    // Most bytes of HBM the table may hold at once, its transient arrays
    // during growth included (0: no limit but the device's own).  One
    // exception: an away list that must grow part way through an in-place
    // doubling grows beyond it, since that doubling cannot stop there.
    uint64_t hbm_budget = 0;

    // Set when an in-place doubling failed after its first phase had
    // overwritten old buckets: the table's keys are then partly split, and
    // every later call that reads or changes them is refused with `damage`
    // (written before the flag is set).
    std::atomic<bool> damaged{false};
    std::string damage;

    uint64_t device_bytes() const { return cap * bucket_bytes() + ov_bytes(ov_slots); }

    // --- Growth mode: doublings in place or out of place ---

    uint32_t growth_mode = KACHE_AMD_GROW_AUTO;
    uint64_t away_min = 0;          // entries of an in-place doubling's away list, at least
    uint64_t in_place_doublings = 0;   // doublings done in place
    uint64_t away_regrows = 0;      // times such a doubling's away list grew
    uint64_t last_away = 0;         // keys the last one listed

    static constexpr uint64_t AWAY_CTL_WORDS = AWAY_REGIONS * AWAY_CURSOR_STRIDE + 8 + 16 * AWAY_REGIONS;   // cursors, fail, snapshots
    static constexpr size_t MAX_PHASES = 16;

    // Whether a store of `bytes` is taken as an address range that can
    // double in place.
    bool growable(uint64_t bytes) const
    {
        return growth_mode != KACHE_AMD_GROW_OUT_OF_PLACE && VmSpace::get().supported(device) &&
               (bytes >= KAMD_VMM_MIN_BYTES || growth_mode == KACHE_AMD_GROW_IN_PLACE);
    }

    // Address space reserved for a growable store of `bytes`: room for four
    // doublings, within twice the device's memory.
    uint64_t growth_reserve(uint64_t bytes) const
    {
        size_t total = 0;
        const uint64_t top = hipDeviceTotalMem(&total, device) == hipSuccess ? 2 * uint64_t(total) : uint64_t(1) << 40;
        return std::max(VmSpace::rounded(bytes), std::min(16 * VmSpace::rounded(bytes), top));
    }

    // HBM an out-of-place doubling needs beside the current arrays: the
    // doubled store, its overflow table and the away list.
    uint64_t out_of_place_bytes() const { return 2 * cap * bucket_bytes() + ov_bytes(next_ov_slots()) + list_bytes(); }

    // Whether the next doubling runs in place.  It can when the store is a
    // range whose reservation holds twice its buckets.  Automatically it
    // does only when an out-of-place doubling would not fit: beyond the
    // table's budget, or leaving less than 1/16 of the device's memory free
    // (out of place measured 2-3% faster on the batched workload: one split
    // launch instead of eight phases, no mapping).
    bool in_place_next() const
    {
        if(growth_mode == KACHE_AMD_GROW_OUT_OF_PLACE || VmSpace::get().reserved(buckets) < VmSpace::rounded(2 * cap * bucket_bytes()))
            return false;
        if(growth_mode == KACHE_AMD_GROW_IN_PLACE)
            return true;
        const uint64_t out = out_of_place_bytes();
        if(hbm_budget != 0 && device_bytes() + out > hbm_budget)
            return true;
        size_t free_b = 0, total_b = 0;
        if(hipMemGetInfo(&free_b, &total_b) != hipSuccess)
            return true;
        return out + total_b / 16 > free_b + ArrayCache::get().bytes_kept(device);
    }

    // Share of the keys the last in-place doubling listed (keys away from
    // their primary bucket); the first guess is 0.3 (a table at load 0.8 of
    // 150-bp reads holds ~20% of its keys in secondary buckets).
    double away_share = 0.3;

    uint64_t away_entries() const
    {
        if(away_min)
            return std::max<uint64_t>(away_min, AWAY_REGIONS);
        const double share = std::min(1.0, away_share * 1.25 + 0.01);
        return std::max<uint64_t>(uint64_t(AWAY_REGIONS) * 256, static_cast<uint64_t>(static_cast<double>(size) * share));
    }

    // A doubling's away list and its control words.
    uint64_t list_bytes() const { return away_entries() * (vw ? 16 : 8) + AWAY_CTL_WORDS * 8; }

    // Transient arrays of an in-place doubling: the copy of the lowest
    // buckets, the away list and its control words.
    uint64_t in_place_extra() const { return std::max<uint64_t>(1, cap >> 6) * bucket_bytes() + list_bytes(); }

    // HBM the next doubling needs beside the current arrays.  In place: the
    // store's second half, the new overflow table and the transient arrays.
    // Out of place: the doubled store, its overflow table and the away list.
    uint64_t next_doubling_bytes() const
    {
        if(in_place_next())
            return cap * bucket_bytes() + ov_bytes(next_ov_slots()) + in_place_extra();
        return out_of_place_bytes();
    }

    // Throws bad_alloc (KACHE_ENOMEM) before growth that would need `more`
    // bytes beside the table's current arrays, beyond the budget, or that
    // needs `fresh` bytes of memory not yet mapped for it (`more` if not
    // given) beyond the device's free memory; nothing has changed then.
    void check_budget(uint64_t more, const char* what, uint64_t fresh = ~uint64_t(0)) const
    {
        if(fresh == ~uint64_t(0))
            fresh = more;
        const uint64_t need = device_bytes() + more;
        if(hbm_budget != 0 && need > hbm_budget)
        {
            last_error = std::string(what) + " needs " + std::to_string(need) + " bytes of HBM, beyond the table's budget of " +
                         std::to_string(hbm_budget);
            throw std::bad_alloc();
        }
        size_t free_b = 0, total_b = 0;
        if(hipMemGetInfo(&free_b, &total_b) == hipSuccess && fresh > free_b + ArrayCache::get().bytes_kept(device))
        {
            last_error = std::string(what) + " needs " + std::to_string(fresh) + " more bytes of HBM, " + std::to_string(free_b) +
                         " free";
            throw std::bad_alloc();
        }
    }

    // Grows the overflow table until it has room for `incoming` more keys
    // without passing the reference's load (0.75, :66) by more than them.
    void ov_room(uint64_t incoming)
    {
        uint64_t slots = ov_slots;
        while(ov_count + incoming + 1 > slots || static_cast<double>(ov_count) >= 0.75 * static_cast<double>(slots))
            slots *= 2;
        if(slots != ov_slots)
            rehash_ov(slots);
    }

    // --- Resize: doubling (include/Streaming_Kmer_Hash_Table.hpp:1473-1681) ---

    // The arrays a doubling replaced go back to the array cache; those it
    // does not keep are freed by a helper thread, so that the launches after
    // it need not wait for the unmapping.  Those frees are waited for before
    // the next doubling allocates and when the table is released, so at most
    // one set is pending.
    std::vector<std::thread> freeing;

    void free_later(const std::vector<std::pair<void*, uint64_t>>& arrays)
    {
        std::vector<void*> rest;
        for(const auto& [p, bytes] : arrays)
            if(p && !ArrayCache::get().keep(device, p, bytes))
                rest.push_back(p);
        if(rest.empty())
            return;
        freeing.emplace_back([dev = device, rest]{
            (void)hipSetDevice(dev);
            for(void* p : rest)
                dev_free(p);
        });
    }

    void wait_frees()
    {
        for(auto& t : freeing)
            t.join();
        freeing.clear();
    }

    // The arrays of the next doubling, taken ahead by a helper thread once
    // one is near: mapping gigabytes takes the host a while, which would
    // otherwise stall the launches around the doubling.
    struct Prepared
    {
        uint64_t cap = 0, ov_slots = 0, bucket_b = 0, ov_b = 0;
        bool in_place = false;      // in place: the overflow table and the store's new chunks; else the new store too
        void* buckets = nullptr;
        void* ov = nullptr;
        std::vector<VmSpace::Chunk> chunks;     // in place: physical memory for the store's second half
    };
    std::future<Prepared> prepared;
    bool prepared_in_place = false;     // the mode the prepared arrays are for

    uint64_t next_ov_slots() const { return std::max<uint64_t>(ov_slots_for(std::max<uint64_t>(OV_MIN, 2 * cap)), ov_slots); }

    void prepare_resize()
    {
        if(prepared.valid() || 2 * cap > MAX_CAP)
            return;
        wait_frees();
        Prepared p;
        p.cap = 2 * cap, p.ov_slots = next_ov_slots(), p.in_place = in_place_next();
        prepared_in_place = p.in_place;
        p.ov_b = ov_bytes(p.ov_slots);
        if(!p.in_place)
            p.bucket_b = p.cap * bucket_bytes();
        if(hbm_budget != 0 && device_bytes() + next_doubling_bytes() > hbm_budget)
            return;     // the doubling would be refused: nothing to prepare
        const bool store_range = !p.in_place && growable(p.bucket_b);
        const uint64_t reserve = store_range ? growth_reserve(p.bucket_b) : 0;
        const uint64_t mapped = p.in_place ? VmSpace::get().mapped(buckets) : 0;
        const uint64_t target = VmSpace::rounded(p.cap * bucket_bytes());
        prepared = std::async(std::launch::async, [p, reserve, store_range, mapped, target, dev = device]() mutable
        {
            (void)hipSetDevice(dev);
            try
            {
                p.ov = ArrayCache::get().take(dev, p.ov_b);
                if(p.in_place && mapped < target)
                    p.chunks = VmSpace::make_chunks(dev, mapped, target);
                if(!p.in_place)
                    p.buckets = store_range ? ArrayCache::get().take_range(dev, p.bucket_b, reserve) : ArrayCache::get().take(dev, p.bucket_b);
            }
            catch(...)
            {
            }
            return p;
        });
    }

    void free_prepared(Prepared& p) const
    {
        VmSpace::release_chunks(p.chunks);
        ArrayCache::get().give(device, p.buckets, p.bucket_b);
        ArrayCache::get().give(device, p.ov, p.ov_b);
    }

    // The prepared arrays, if they are those of the doubling about to run.
    Prepared take_prepared(uint64_t ncap, uint64_t nov_slots, bool in_place_now)
    {
        if(!prepared.valid())
            return {};
        Prepared p = prepared.get();
        if(p.cap != ncap || p.ov_slots != nov_slots || p.in_place != in_place_now || !p.ov || (!p.in_place && !p.buckets))
        {
            free_prepared(p);
            return {};
        }
        return p;
    }

    void resize()
    {
        if(2 * cap > MAX_CAP)
            throw hip_failure("growing the table past the most buckets supported");
        wait_frees();
        // Automatically, the mode chosen when the arrays were prepared (the
        // prepared arrays themselves take free memory the choice looks at).
        const bool prep = growth_mode == KACHE_AMD_GROW_AUTO && prepared.valid();
        resize_split(prep ? prepared_in_place : in_place_next());
        resizes++;
        resize_th = static_cast<uint64_t>(static_cast<double>(cap * B) * lf);
        step = step_for(resize_th);
    }

    // Doubling (include/Streaming_Kmer_Hash_Table.hpp:1473-1681) by
    // k_split and k_reinsert.
    // In place: the store's range is mapped to twice its buckets, the lowest
    // K = cap / 64 buckets are copied aside, the phases split the store from
    // the top and the copy last.  Peak memory: the doubled store, the two
    // overflow tables, the copy and the away list.  Once the first phase has
    // run the old buckets are being overwritten: a device error from there on
    // leaves the table unusable.
    // Out of place: one phase [0, cap) from the old store into a new one of
    // twice its size; the old arrays are only read, so any failure leaves the
    // table as it was and the new arrays go.
    // Either way the split lists the keys away from their primary bucket, with
    // their values; k_reinsert inserts those and the old overflow
    // table's keys.  A full away list region, or a new overflow table with no
    // room left for the reinsert, is handled by growing the array and running
    // that step again.
    void resize_split(bool in_place)
    {
        const uint32_t bb = bucket_bytes();
        const uint64_t ncap = 2 * cap;
        uint64_t nov_slots = next_ov_slots();
        const uint64_t K = in_place ? std::max<uint64_t>(1, cap >> 6) : 0;
        std::vector<std::pair<uint64_t, uint64_t>> phases;
        if(in_place)
        {
            for(uint64_t hi = cap; hi > K;)
            {
                const uint64_t lo = std::max(K, (hi + 1) / 2);
                phases.push_back({lo, hi});
                hi = lo;
            }
            phases.push_back({0, K});
        }
        else
            phases.push_back({0, cap});
        if(phases.size() > MAX_PHASES)
            throw hip_failure("an in-place doubling with too many phases");
        const uint64_t mapped = in_place ? VmSpace::get().mapped(buckets) : 0;
        const uint64_t extra = in_place ? in_place_extra() : list_bytes();
        Prepared pre = take_prepared(ncap, nov_slots, in_place);
        // Memory the helper thread took already (the second half's chunks or
        // the new store, the new overflow table) is not asked of the device
        // again.
        uint64_t have = (pre.ov ? ov_bytes(nov_slots) : 0) + (pre.buckets ? ncap * bb : 0);
        for(const auto& c : pre.chunks)
            have += c.n;
        const uint64_t store_more = in_place ? cap * bb : ncap * bb;
        const uint64_t store_fresh = in_place ? (2 * cap * bb > mapped ? 2 * cap * bb - mapped : 0) : ncap * bb;
        const uint64_t fresh = store_fresh + ov_bytes(nov_slots) + extra;

        uint8_t* dst = in_place ? buckets : static_cast<uint8_t*>(pre.buckets);   // the doubled store
        uint64_t* nov = static_cast<uint64_t*>(pre.ov);
        uint8_t* copy = nullptr;
        uint64_t* ctl = nullptr;
        uint64_t entries = away_entries();
        uint64_t* lkeys = nullptr;
        uint64_t* lvals = nullptr;
        const auto give_new = [&]
        {
            (void)hipStreamSynchronize(stream);
            VmSpace::release_chunks(pre.chunks);
            if(!in_place)
                ArrayCache::get().give(device, dst, ncap * bb), dst = nullptr;
            ArrayCache::get().give(device, nov, ov_bytes(nov_slots)), nov = nullptr;
            ArrayCache::get().give(device, copy, K * bb), copy = nullptr;
            ArrayCache::get().give(device, ctl, AWAY_CTL_WORDS * 8), ctl = nullptr;
            ArrayCache::get().give(device, lkeys, entries * 8), lkeys = nullptr;
            ArrayCache::get().give(device, lvals, entries * 8), lvals = nullptr;
        };
        try
        {
            check_budget(store_more + ov_bytes(nov_slots) + extra, "doubling the table", fresh > have ? fresh - have : 0);
            if(in_place)
            {
                VmSpace::get().grow(buckets, ncap * bb, &pre.chunks);
                VmSpace::release_chunks(pre.chunks);     // any not used
            }
            else if(!dst)
                dst = take_buckets(ncap);
            if(nov)
                zero(nov, ov_bytes(nov_slots));
            else
                nov = alloc_ov(nov_slots);
            if(in_place)
                copy = static_cast<uint8_t*>(ArrayCache::get().take(device, K * bb));
            ctl = static_cast<uint64_t*>(ArrayCache::get().take(device, AWAY_CTL_WORDS * 8));
            lkeys = static_cast<uint64_t*>(ArrayCache::get().take(device, entries * 8));
            if(vw)
                lvals = static_cast<uint64_t*>(ArrayCache::get().take(device, entries * 8));
        }
        catch(...)
        {
            // Nothing of the table has changed (in place, its range may stay
            // mapped further).
            give_new();
            throw;
        }

        try
        {
            // Every phase runs on the first (largest) phase's grid, so that
            // the waves of each away list region are the same in every phase.
            uint32_t grid = 0;
            by_layout([&](auto V){
                const auto kern = k_split<decltype(V)::value>;
                const uint64_t pairs = (phases[0].second - phases[0].first + 1) / 2;
                grid = grid_for(kern, (pairs * 64 + TILE - 1) / TILE);
            });
            const uint32_t regions = static_cast<uint32_t>(std::min<uint64_t>(AWAY_REGIONS, uint64_t(grid) * (TILE / 64)));
            AwayList L{lkeys, lvals, entries / regions, ctl, reinterpret_cast<uint32_t*>(ctl + AWAY_REGIONS * AWAY_CURSOR_STRIDE), regions};
            uint64_t* snaps = ctl + AWAY_REGIONS * AWAY_CURSOR_STRIDE + 8;
            const TableView N = view_of(dst, ncap, nov, nov_slots);
            const auto run_phases = [&](size_t from)
            {
                for(size_t k = from; k < phases.size(); ++k)
                {
                    hipLaunchKernelGGL(k_away_snap, dim3(1), dim3(256), 0, stream, static_cast<const uint64_t*>(L.cursor), regions,
                                       snaps + k * AWAY_REGIONS);
                    const uint8_t* src = (in_place && k + 1 == phases.size()) ? copy : buckets;
                    timed("split", [&]{
                        by_layout([&](auto V){
                            hipLaunchKernelGGL(k_split<decltype(V)::value>, dim3(grid), dim3(TILE), 0, stream, src, N, cap,
                                               phases[k].first, phases[k].second, L, static_cast<uint32_t>(k));
                        });
                    });
                }
            };
            zero(ctl, AWAY_CTL_WORDS * 8);
            if(in_place)
                HIP(hipMemcpyAsync(copy, buckets, K * bb, hipMemcpyDeviceToDevice, stream));
            run_phases(0);
            constexpr uint64_t CUR_WORDS = AWAY_REGIONS * AWAY_CURSOR_STRIDE;
            h_away.reserve((CUR_WORDS + 1) * 8);
            uint64_t listed = 0;
            for(;;)
            {
                // The cursors and the fail word, in one copy.
                HIP(hipMemcpyAsync(h_away.p, ctl, (CUR_WORDS + 1) * 8, hipMemcpyDeviceToHost, stream));
                sync();
                const uint64_t* hc = h_away.as<uint64_t>();
                const uint64_t fail = hc[CUR_WORDS] & 0xFFFFFFFFull;
                if(fail == 0)
                {
                    for(uint32_t r = 0; r < regions; ++r)
                        listed += hc[uint64_t(r) * AWAY_CURSOR_STRIDE];
                    break;
                }
                // Phase fail - 1 found its region full: grow the list, keep
                // each region's entries from before that phase, and run it
                // again (its sources are untouched).
                const size_t k = fail - 1;
                const uint64_t region2 = 2 * L.region, entries2 = region2 * regions;
                if(std::getenv("KACHE_AMD_TEST_FAIL_AWAY_REGROW") != nullptr)
                {
                    last_error = "growing a doubling's away list: refused (KACHE_AMD_TEST_FAIL_AWAY_REGROW is set)";
                    throw std::bad_alloc();
                }
                AwayList L2 = L;
                L2.keys = static_cast<uint64_t*>(ArrayCache::get().take(device, entries2 * 8));
                L2.vals = nullptr;
                L2.region = region2;
                if(vw)
                {
                    try
                    {
                        L2.vals = static_cast<uint64_t*>(ArrayCache::get().take(device, entries2 * 8));
                    }
                    catch(...)
                    {
                        ArrayCache::get().give(device, L2.keys, entries2 * 8);
                        throw;
                    }
                }
                hipLaunchKernelGGL(k_away_regrow, dim3(std::min<uint64_t>(1024, (L.region + 255) / 256 + 1)), dim3(256), 0, stream, L, L2,
                                   static_cast<const uint64_t*>(snaps + k * AWAY_REGIONS));
                HIP(hipGetLastError());
                sync();
                ArrayCache::get().give(device, lkeys, entries * 8);
                ArrayCache::get().give(device, lvals, entries * 8);
                lkeys = L2.keys, lvals = L2.vals, entries = entries2;
                L = L2;
                away_regrows++;
                run_phases(k);
            }

            // The listed keys and the old overflow table's, into the doubled table.
            uint64_t ov_n = 0;
            for(int attempt = 0;; ++attempt)
            {
                const TableView NV = view_of(dst, ncap, nov, nov_slots);
                const uint64_t chunks = uint64_t(regions) * ((L.region + 63) / 64) + (ov_slots + 63) / 64;
                timed("reinsert", [&]{
                    by_layout([&](auto V){
                        const auto kern = k_reinsert<decltype(V)::value>;
                        hipLaunchKernelGGL(kern, dim3(grid_for(kern, (chunks * 64 + TILE - 1) / TILE)), dim3(TILE), 0, stream, L, ov, ov_slots,
                                           NV, d_cnt, in_place);
                    });
                });
                const Counters c = read_counters();
                ov_n += c.ov_inserted;
                if(!c.failed)
                    break;
                // Out of place the table is left as it was; in place the keys
                // left are tried again.
                if(!in_place || attempt == 8)
                    throw hip_failure("a doubling found no room in the overflow table");
                // The keys left stay listed: a larger new overflow table, then again.
                uint64_t* nov2 = alloc_ov(2 * nov_slots);
                const TableView N2 = view_of(dst, ncap, nov2, 2 * nov_slots);
                by_layout([&](auto V){
                    const auto kern = k_ov_rehash<decltype(V)::value>;
                    hipLaunchKernelGGL(kern, dim3(grid_for(kern, (nov_slots + TILE - 1) / TILE)), dim3(TILE), 0, stream, nov, nov_slots, N2, d_cnt);
                });
                HIP(hipGetLastError());
                if(read_counters().failed)
                {
                    ArrayCache::get().give(device, nov2, ov_bytes(2 * nov_slots));
                    throw hip_failure("a doubling found no room in the overflow table");
                }
                ArrayCache::get().give(device, nov, ov_bytes(nov_slots));
                nov = nov2, nov_slots *= 2;
            }

            std::vector<std::pair<void*, uint64_t>> old = {{ov, ov_bytes(ov_slots)}, {copy, K * bb}, {ctl, AWAY_CTL_WORDS * 8},
                                                           {lkeys, entries * 8}, {lvals, entries * 8}};
            if(!in_place)
                old.push_back({buckets, cap * bb});
            free_later(old);
            last_away = listed;
            if(size > 0)
                away_share = static_cast<double>(listed) / static_cast<double>(size);
            buckets = dst, cap = ncap, ov = nov, ov_slots = nov_slots, ov_count = ov_n;
            if(in_place)
                in_place_doublings++;
        }
        catch(...)
        {
            // Out of place the old arrays were only read: the table is as it
            // was, and the new arrays go.  In place the first phase has
            // overwritten old buckets: the table is refused from now on, and
            // the transient arrays go (the store stays, for kache_free).
            if(in_place)
            {
                std::string why = last_error;
                try
                {
                    throw;
                }
                catch(const std::bad_alloc&)
                {
                }
                catch(const std::exception& e)
                {
                    why = e.what();
                }
                catch(...)
                {
                }
                damage = "the table is unusable: an in-place doubling failed part way (" + (why.empty() ? std::string("out of memory") : why) + ")";
                damaged.store(true, std::memory_order_release);
            }
            give_new();
            throw;
        }
    }

    // The growth schedule of one insert call by one user: the keys it inserts
    // are flushed into the table's size every `step` keys, and a flush that
    // reaches the threshold doubles the table
    // (include/Streaming_Kmer_Hash_Table.hpp:2246-2286, 2509-2513).
    struct CallSched
    {
        uint64_t pending, flushed = 0, ins_c = 0, due;
    };

    CallSched sched_begin(uint32_t user) const
    {
        CallSched s;
        s.pending = checkpoint[user];
        s.due = s.pending < step ? step - s.pending : 1;
        return s;
    }

    // Accounts for `n` more keys inserted; returns the number of doublings due.
    uint32_t sched_add(CallSched& s, uint64_t n)
    {
        uint32_t doublings = 0;
        const uint64_t end = s.ins_c + n;
        while(s.due <= end)
        {
            s.ins_c = s.due;
            approx_sz += s.pending + (s.ins_c - s.flushed);
            s.pending = 0, s.flushed = s.ins_c;
            s.due = s.ins_c + step;
            if(approx_sz >= resize_th)
            {
                doublings++;
                // The threshold and step of the doubled table govern later flushes.
                const uint64_t c2 = cap << doublings;
                resize_th = static_cast<uint64_t>(static_cast<double>(c2 * B) * lf);
                step = step_for(resize_th);
            }
        }
        s.ins_c = end;
        return doublings;
    }

    void sched_end(const CallSched& s, uint32_t user) { checkpoint[user] = s.pending + (s.ins_c - s.flushed); }

    void grow(uint32_t doublings)
    {
        // sched_add already moved the thresholds; resize() recomputes them for
        // the final capacity, which is the same.
        const uint64_t th = resize_th, st = step;
        for(uint32_t i = 0; i < doublings; ++i)
        {
            try
            {
                resize();
            }
            catch(...)
            {
                // The table keeps the capacity it has: the threshold and step
                // are that capacity's again, and the flushed size stays at or
                // above the threshold, so the next flush tries the doubling
                // again (include/Streaming_Kmer_Hash_Table.hpp:2266-2286).
                resize_th = static_cast<uint64_t>(static_cast<double>(cap * B) * lf);
                step = step_for(resize_th);
                throw;
            }
        }
        resize_th = th, step = st;
    }

    // --- Staging of host sequences ---

    // Packs the caller's sequences from `cur` into the pinned staging buffer,
    // up to about STAGE_BYTES, splitting a sequence too long for one stage
    // into pieces that overlap by k - 1 characters.  Returns false when done.
    bool stage_next(const char* const* seqs, const size_t* lens, size_t n, SeqCursor& cur, Stage& st) const
    {
        return stage_into(seqs, lens, n, cur, st, h_text, STAGE_BYTES);
    }

    bool stage_into(const char* const* seqs, const size_t* lens, size_t n, SeqCursor& cur, Stage& st, HostBuf& hb, uint64_t limit) const
    {
        st.text_len = 0, st.seg_start.clear(), st.seg_q.clear(), st.seg_off.clear();
        while(cur.q < n && (lens[cur.q] < static_cast<size_t>(K) || cur.off + K > lens[cur.q]))
            cur.q++, cur.off = 0;
        if(cur.q >= n)
            return false;

        // Size the stage.
        uint64_t need = 0;
        for(size_t q = cur.q; q < n && need < limit; ++q)
            need += (q == cur.q ? lens[q] - cur.off : lens[q]) + 1;
        need = std::min<uint64_t>(need, limit + K);
        hb.reserve(need + TEXT_PAD);
        char* out = hb.as<char>();

        while(cur.q < n)
        {
            const size_t q = cur.q;
            if(lens[q] < static_cast<size_t>(K) || cur.off + K > lens[q])
            {
                cur.q++, cur.off = 0;
                continue;
            }
            const uint64_t room = need - st.text_len;
            const uint64_t rest = lens[q] - cur.off;
            if(rest + 1 <= room)
            {
                st.seg_start.push_back(st.text_len), st.seg_q.push_back(static_cast<uint32_t>(q)), st.seg_off.push_back(cur.off);
                st.text_len += rest + 1;
                cur.q++, cur.off = 0;
                continue;
            }
            if(room >= static_cast<uint64_t>(K) + 1 && (st.seg_start.empty() || room > (1u << 20)))
            {
                // A piece of the sequence; the next piece starts k - 1 characters before its end.
                const uint64_t take = room - 1;
                st.seg_start.push_back(st.text_len), st.seg_q.push_back(static_cast<uint32_t>(q)), st.seg_off.push_back(cur.off);
                st.text_len += take + 1;
                cur.off += take - (K - 1);
            }
            break;
        }
        st.seg_start.push_back(st.text_len);
        // The copies: segment i fills [seg_start[i], seg_start[i + 1] - 1), then '\n'.
        const size_t ns = st.seg_q.size();
        const auto copy = [&](size_t a, size_t b)
        {
            for(size_t i = a; i < b; ++i)
            {
                const uint64_t len = st.seg_start[i + 1] - st.seg_start[i] - 1;
                std::memcpy(out + st.seg_start[i], seqs[st.seg_q[i]] + st.seg_off[i], len);
                out[st.seg_start[i] + len] = '\n';
            }
        };
        if(st.text_len < (uint64_t(4) << 20) || ns < 2)
        {
            if(ns == 1 && st.text_len >= (uint64_t(4) << 20))
            {
                par_copy(out, seqs[st.seg_q[0]] + st.seg_off[0], st.text_len - 1);
                out[st.text_len - 1] = '\n';
            }
            else
                copy(0, ns);
        }
        else
        {
            // Ranges of segments of about equal bytes, one per task.
            const size_t parts = CopyPool::THREADS;
            std::vector<size_t> cut(parts + 1, ns);
            cut[0] = 0;
            for(size_t p = 1, i = 0; p < parts; ++p)
            {
                const uint64_t target = st.text_len * p / parts;
                while(i < ns && st.seg_start[i] < target)
                    ++i;
                cut[p] = i;
            }
            CopyPool::get().run(parts, [&](size_t p){ copy(cut[p], std::max(cut[p], cut[p + 1])); });
        }
        std::memset(out + st.text_len, '\n', TEXT_PAD);
        return true;
    }

    // Copies the staged text to the device.
    const uint8_t* upload(const Stage& st) const
    {
        d_text.reserve(st.text_len + TEXT_PAD);
        HIP(hipMemcpyAsync(d_text.p, h_text.p, st.text_len + TEXT_PAD, hipMemcpyHostToDevice, stream));
        return d_text.as<uint8_t>();
    }

    // --- Inserts ---

    // The insert kernel of an operation, at `WV` waves per SIMD and `BT`
    // threads per block.
    template <int VB, int WV, int BT>
    static const void* insert_kernel_of(uint32_t op, bool oriented)
    {
        const auto f = [](auto kern) { return reinterpret_cast<const void*>(kern); };
        if(oriented)
            switch(op)
            {
            case OP_ASSIGN: return f(k_insert<VB, OP_ASSIGN, true, WV, BT>);
            case OP_ADD:    return f(k_insert<VB, OP_ADD, true, WV, BT>);
            case OP_OR:     return f(k_insert<VB, OP_OR, true, WV, BT>);
            default:        return f(k_insert<VB, OP_MAX, true, WV, BT>);
            }
        switch(op)
        {
        case OP_KEEP:   return f(k_insert<VB, OP_KEEP, false, WV, BT>);
        case OP_ASSIGN: return f(k_insert<VB, OP_ASSIGN, false, WV, BT>);
        case OP_ADD:    return f(k_insert<VB, OP_ADD, false, WV, BT>);
        case OP_OR:     return f(k_insert<VB, OP_OR, false, WV, BT>);
        case OP_MIN:    return f(k_insert<VB, OP_MIN, false, WV, BT>);
        case OP_MAX:    return f(k_insert<VB, OP_MAX, false, WV, BT>);
        default:        return f(k_insert<VB, OP_ADD_WRAP, false, WV, BT>);
        }
    }

    template <int WV, int BT>
    const void* insert_kernel_at(uint32_t op, bool oriented) const
    {
        if(!map)
            return reinterpret_cast<const void*>(k_insert<0, OP_KEEP, false, WV, BT>);
        return vw == 4 ? insert_kernel_of<4, WV, BT>(op, oriented) : insert_kernel_of<8, WV, BT>(op, oriented);
    }

    template <int BT>
    const void* insert_kernel_bt(uint32_t op, bool oriented, uint32_t waves) const
    {
        switch(waves)
        {
        case 6:  return insert_kernel_at<6, BT>(op, oriented);
        case 8:  return insert_kernel_at<8, BT>(op, oriented);
        default: return insert_kernel_at<KAMD_INSERT_WAVES, BT>(op, oriented);
        }
    }

    const void* insert_kernel(uint32_t op, bool oriented, uint32_t waves = 0, uint32_t threads = 0) const
    {
        waves = waves ? waves : insert_waves;
        switch(threads ? threads : insert_threads)
        {
        case 64:  return insert_kernel_bt<64>(op, oriented, waves);
        case 128: return insert_kernel_bt<128>(op, oriented, waves);
        default:  return insert_kernel_bt<TILE>(op, oriented, waves);
        }
    }

    // The arguments of one insert launch, in k_insert's order.  bind() once
    // the object is in place.
    struct InsertArgs
    {
        TableView T;
        const uint8_t* text;
        uint64_t text_len, b, e, x, xr;
        Counters* cnt;
        Handoff prev;
        void* ptrs[9];
        void bind()
        {
            void* p[9] = {&T, &text, &text_len, &b, &e, &x, &xr, &cnt, &prev};
            std::copy(p, p + 9, ptrs);
        }
    };

    void insert_args(InsertArgs& a, const uint8_t* text, uint64_t text_len, uint64_t b, uint64_t e, uint64_t x, uint64_t xr, Counters* cnt,
                     bool oriented, const Handoff& prev = {}) const
    {
        a.T = view(), a.text = text, a.text_len = text_len, a.b = b, a.e = e;
        a.x = map ? x : 0, a.xr = (map && oriented) ? xr : 0, a.cnt = cnt ? cnt : d_cnt, a.prev = prev;
        a.bind();
    }

    uint32_t insert_grid(const void* kern, uint64_t b, uint64_t e) const
    {
        return grid_for(kern, (e - (b & ~uint64_t(15)) + insert_threads - 1) / insert_threads, insert_threads);
    }

    void dispatch_insert(uint32_t op, bool oriented, const uint8_t* text, uint64_t text_len, uint64_t b, uint64_t e, uint64_t x, uint64_t xr,
                         Counters* cnt = nullptr, const Handoff& prev = {}) const
    {
        const void* kern = insert_kernel(op, oriented);
        InsertArgs a;
        insert_args(a, text, text_len, b, e, x, xr, cnt, oriented, prev);
        HIP(hipLaunchKernel(kern, dim3(insert_grid(kern, b, e)), dim3(insert_threads), a.ptrs, 0, stream));
    }

    // --- Launch shape and its online tuner ---

    // Positions per insert launch are 2^launch_log2; blocks per CU as
    // blocks_per_cu; with use_graph, a launch and its counter read-back run
    // as one replay of a captured hipGraph whose kernel node takes the new
    // arguments.
    uint32_t launch_log2 = 24;
    bool use_graph = false;
    uint32_t insert_waves = KAMD_INSERT_WAVES;   // occupancy shape of the insert kernel (6, 7 or 8 waves per SIMD)
    uint32_t insert_threads = TILE;              // its tile: threads per block = positions per tile (64, 128, 256)

    // The tuner: a coordinate descent over the shape's dimensions (waves per
    // SIMD, blocks per CU, positions per launch, graph replay).  A stage
    // varies one dimension of the best shape so far; its launches alternate
    // that shape (the reference) and each candidate, `tune_rounds` times, and
    // a candidate's score is the median ratio of its time per position to
    // the mean of the reference launches either side of it.  The ratio
    // cancels what changes during tuning besides the shape (the table's
    // fill, the share of new keys), which differs more between consecutive
    // launches than the shapes do.  A candidate replaces the reference if its
    // score is below 0.99.  Launches during tuning are real work.
    struct Trial
    {
        kache_amd_tuning cfg;
        std::vector<double> ns_per_pos, rel;
    };
    std::vector<Trial> trials;          // every shape tried; the first of each stage is its reference
    uint32_t tune_rounds = 0;           // > 0 while tuning
    uint32_t tune_stage = 0;
    std::vector<int> tune_plan;         // the stage's launches: trial indices
    size_t tune_next = 0;               // next launch of the plan
    std::vector<std::pair<int, double>> tune_seen;   // the stage's launches so far: (trial, ns per position)
    kache_amd_tuning tune_best{};

    // One graph per insert kernel: [kernel, counters to host and zeroed].
    struct InsertGraph
    {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        hipGraphNode_t kernel = nullptr;
    };
    mutable std::map<const void*, InsertGraph> graphs;

    void free_graphs() const
    {
        for(auto& [k, g] : graphs)
        {
            if(g.exec)
                (void)hipGraphExecDestroy(g.exec);
            if(g.graph)
                (void)hipGraphDestroy(g.graph);
        }
        graphs.clear();
    }

    // Launches one insert and its counter read-back as a graph replay; the
    // counters are then in h_cnt once the stream is synchronized.
    void graph_insert(uint32_t op, bool oriented, const uint8_t* text, uint64_t text_len, uint64_t b, uint64_t e, uint64_t x, uint64_t xr) const
    {
        const void* kern = insert_kernel(op, oriented);
        InsertArgs a;
        insert_args(a, text, text_len, b, e, x, xr, nullptr, oriented);
        const uint32_t grid = insert_grid(kern, b, e);
        auto& g = graphs[kern];
        if(!g.exec)
        {
            HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed));
            hipError_t le = hipLaunchKernel(kern, dim3(grid), dim3(insert_threads), a.ptrs, 0, stream);
            if(le == hipSuccess)
            {
                hipLaunchKernelGGL(k_publish, dim3(1), dim3(64), 0, stream, reinterpret_cast<uint64_t*>(d_cnt), reinterpret_cast<uint64_t*>(h_cnt),
                                   uint64_t(1));
                le = hipGetLastError();
            }
            hipGraph_t graph = nullptr;
            const hipError_t ce = hipStreamEndCapture(stream, &graph);
            HIP(le);
            HIP(ce);
            g.graph = graph;
            HIP(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
            size_t n = 0;
            HIP(hipGraphGetNodes(g.graph, nullptr, &n));
            std::vector<hipGraphNode_t> nodes(n);
            HIP(hipGraphGetNodes(g.graph, nodes.data(), &n));
            for(auto node : nodes)
            {
                hipGraphNodeType ty;
                HIP(hipGraphNodeGetType(node, &ty));
                if(ty != hipGraphNodeTypeKernel)
                    continue;
                hipKernelNodeParams kp{};
                HIP(hipGraphKernelNodeGetParams(node, &kp));
                if(kp.func == kern)
                    g.kernel = node;
            }
            if(!g.kernel)
                throw hip_failure("the captured insert graph has no insert kernel node");
        }
        hipKernelNodeParams kp{};
        kp.func = const_cast<void*>(kern);
        kp.gridDim = dim3(grid);
        kp.blockDim = dim3(insert_threads);
        kp.sharedMemBytes = 0;
        kp.kernelParams = a.ptrs;
        kp.extra = nullptr;
        HIP(hipGraphExecKernelNodeSetParams(g.exec, g.kernel, &kp));
        HIP(hipGraphLaunch(g.exec, stream));
    }

    // One insert launch of [b, e) and its counters, by the current launch
    // shape; while tuning, the shape is the next candidate's and the launch
    // is timed.
    Counters insert_launch(uint32_t op, bool oriented, const uint8_t* text, uint64_t text_len, uint64_t b, uint64_t e, uint64_t x, uint64_t xr)
    {
        Trial* trial = nullptr;
        int trial_i = -1;
        if(tune_rounds)
        {
            trial_i = tune_plan[tune_next];
            trial = &trials[trial_i];
            apply_shape(trial->cfg);
        }
        const auto t0 = std::chrono::steady_clock::now();
        Counters c;
        if(use_graph)
        {
            timed("insert", [&]{ graph_insert(op, oriented, text, text_len, b, e, x, xr); });
            sync();
            c = *h_cnt;
        }
        else
        {
            timed("insert", [&]{ dispatch_insert(op, oriented, text, text_len, b, e, x, xr); });
            c = read_counters();
        }
        if(trial)
        {
            const double ns = std::chrono::duration<double, std::nano>(std::chrono::steady_clock::now() - t0).count() / double(e - b);
            trial->ns_per_pos.push_back(ns);
            tune_seen.push_back({trial_i, ns});
            if(++tune_next == tune_plan.size())
                end_stage();
        }
        return c;
    }

    void apply_shape(const kache_amd_tuning& cfg)
    {
        blocks_per_cu = cfg.blocks_per_cu, use_graph = cfg.graph != 0, launch_log2 = cfg.launch_log2;
        insert_waves = cfg.waves_per_simd ? cfg.waves_per_simd : KAMD_INSERT_WAVES;
        insert_threads = cfg.threads_per_block ? cfg.threads_per_block : TILE;
    }

    // The shape the next launch takes: the tuner's while it runs.
    uint32_t next_launch_log2() const { return tune_rounds ? trials[tune_plan[tune_next]].cfg.launch_log2 : launch_log2; }

    void start_tuning(uint32_t rounds)
    {
        trials.clear(), tune_seen.clear(), tune_plan.clear();
        tune_next = 0, tune_stage = 0;
        tune_rounds = rounds;
        if(!rounds)
            return;
        tune_best = kache_amd_tuning{blocks_per_cu, launch_log2, use_graph ? 1u : 0u, 0, insert_waves, insert_threads};
        begin_stage();
    }

    // Plans the launches of the current stage (the reference, then each
    // candidate after a reference, `tune_rounds` times, then a last reference).
    void begin_stage()
    {
        std::vector<kache_amd_tuning> cands;
        kache_amd_tuning c = tune_best;
        switch(tune_stage)
        {
        case 0:
            for(uint32_t wv : {6u, 7u, 8u})
                if(wv != tune_best.waves_per_simd)
                    c.waves_per_simd = wv, cands.push_back(c);
            break;
        case 1:
            // The tile: threads per block, i.e. waves of one block walking
            // the positions together (each wave describes and probes its 64
            // on its own); all blocks that fit stay resident.
            for(uint32_t bt : {64u, 128u, 256u})
                if(bt != tune_best.threads_per_block)
                    c.threads_per_block = bt, c.blocks_per_cu = 0, cands.push_back(c);
            break;
        case 2:
        {
            int resident = 0;
            const uint32_t bt = tune_best.threads_per_block ? tune_best.threads_per_block : TILE;
            if(hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, insert_kernel(OP_ADD, false, tune_best.waves_per_simd, bt),
                                                            static_cast<int>(bt), 0) != hipSuccess ||
               resident < 1)
                resident = 8;
            for(uint32_t b : {0u, static_cast<uint32_t>(resident * 3 / 4), static_cast<uint32_t>(resident / 2)})
                if(b != tune_best.blocks_per_cu && (b == 0 || b < static_cast<uint32_t>(resident)))
                    c.blocks_per_cu = b, cands.push_back(c);
            break;
        }
        case 3:
            for(uint32_t lg : {23u, 24u, 25u})
                if(lg != tune_best.launch_log2)
                    c.launch_log2 = lg, cands.push_back(c);
            break;
        case 4:
            c.graph = tune_best.graph ? 0u : 1u, cands.push_back(c);
            break;
        default:
            break;
        }
        if(cands.empty())
        {
            finish_tuning();
            return;
        }
        const int ref = static_cast<int>(trials.size());
        trials.push_back({tune_best, {}, {}});
        for(const auto& x : cands)
            trials.push_back({x, {}, {}});
        tune_plan.clear(), tune_seen.clear();
        tune_next = 0;
        for(uint32_t r = 0; r < tune_rounds; ++r)
            for(size_t i = 0; i < cands.size(); ++i)
                tune_plan.push_back(ref), tune_plan.push_back(ref + 1 + static_cast<int>(i));
        tune_plan.push_back(ref);
    }

    // Scores the stage's candidates against the reference launches beside
    // them, keeps the best shape, and starts the next stage.
    void end_stage()
    {
        const int ref = tune_plan.front();
        for(size_t k = 1; k + 1 < tune_seen.size(); ++k)
            if(tune_seen[k].first != ref)
                trials[tune_seen[k].first].rel.push_back(tune_seen[k].second / (0.5 * (tune_seen[k - 1].second + tune_seen[k + 1].second)));
        trials[ref].rel.assign(1, 1.0);
        double best = 0.99;
        for(size_t i = ref + 1; i < trials.size(); ++i)
        {
            std::vector<double> v = trials[i].rel;
            if(v.empty())
                continue;
            std::sort(v.begin(), v.end());
            const double med = v[v.size() / 2];
            if(med < best)
                best = med, tune_best = trials[i].cfg;
        }
        tune_stage++;
        begin_stage();
    }

    void finish_tuning()
    {
        tune_rounds = 0;
        tune_plan.clear(), tune_seen.clear();
        tune_best.tuning = 0;
        apply_shape(tune_best);
    }

    // Inserts the positions of a device text, in launches of at most
    // 2^launch_log2 positions, growing the overflow table ahead of each so it
    // cannot fill, and the table after each as the schedule says.
    uint64_t insert_text(CallSched& s, uint32_t op, bool oriented, const uint8_t* text, uint64_t text_len, uint64_t x, uint64_t xr,
                         uint64_t first = 0)
    {
        uint64_t inserted = 0;
        for(uint64_t b = first; b < text_len;)
        {
            const uint64_t e = std::min(text_len, b + (uint64_t(1) << next_launch_log2()));
            ov_room(e - b);
            if(size + 2 * (e - b) >= resize_th)
                prepare_resize();
            const Counters c = insert_launch(op, oriented, text, text_len, b, e, x, xr);
            if(c.failed)
                throw hip_failure("an insert found no room in the overflow table");
            inserted += c.inserted, size += c.inserted, ov_count += c.ov_inserted;
            grow(sched_add(s, c.inserted));
            b = e;
        }
        return inserted;
    }

    // Runs `work(text, text_len)` on each chunk of the caller's sequences, as
    // packed by stage_into in chunks of about PIPE_BYTES: while chunk i's
    // kernels run on `stream`, chunk i + 1 is packed into the other pinned
    // buffer and copied on copy_stream.  `work` enqueues its kernels on
    // `stream` and may synchronize it; chunk i + 1's copy never waits for it.
    template <typename Work>
    void host_pipeline(const char* const* seqs, const size_t* lens, size_t n, const Work& work) const
    {
        SeqCursor cur;
        pipeline([&](Stage& st, HostBuf& hb){ return stage_into(seqs, lens, n, cur, st, hb, PIPE_BYTES); }, work);
    }

    // The same over a host text: sequences separated by non-ACGT bytes, cut
    // into chunks of about PIPE_BYTES that end after such a byte (a run of
    // bases longer than a chunk is cut with k - 1 bases of overlap).
    template <typename Work>
    void text_pipeline(const char* text, uint64_t bytes, const Work& work) const
    {
        uint64_t off = 0;
        pipeline([&](Stage& st, HostBuf& hb)
        {
            st.text_len = 0;
            if(off >= bytes)
                return false;
            uint64_t end = std::min(bytes, off + PIPE_BYTES), next = end;
            if(end < bytes)
            {
                uint64_t e = end;
                while(e > off && is_base(static_cast<uint8_t>(text[e - 1])))
                    e--;
                if(e > off)
                    end = next = e;
                else
                    next = end - (K - 1);
            }
            hb.reserve(end - off + 1 + TEXT_PAD);
            char* out = hb.as<char>();
            par_copy(out, text + off, end - off);
            out[end - off] = '\n';
            std::memset(out + end - off + 1, '\n', TEXT_PAD);
            st.text_len = end - off + 1;
            off = next;
            return true;
        }, work);
    }

    template <typename Fill, typename Work>
    void pipeline(const Fill& fill, const Work& work) const
    {
        Stage st[2];
        const auto stage = [&](int slot)
        {
            // The pinned buffer is free once its last copy ended; the device
            // text once the kernels of two chunks ago are done.
            HIP(hipEventSynchronize(ev_copied[slot]));
            if(!fill(st[slot], h_pipe[slot]))
                return false;
            const uint64_t bytes = st[slot].text_len + TEXT_PAD;
            if(bytes > d_pipe[slot].bytes)
            {
                HIP(hipStreamSynchronize(stream));
                d_pipe[slot].reserve(bytes);
            }
            HIP(hipStreamWaitEvent(copy_stream, ev_freed[slot], 0));
            HIP(hipMemcpyAsync(d_pipe[slot].p, h_pipe[slot].p, bytes, hipMemcpyHostToDevice, copy_stream));
            HIP(hipEventRecord(ev_copied[slot], copy_stream));
            return true;
        };
        try
        {
            bool have = stage(0);
            for(int slot = 0; have; slot ^= 1)
            {
                HIP(hipStreamWaitEvent(stream, ev_copied[slot], 0));
                work(d_pipe[slot].as<uint8_t>(), st[slot].text_len, [&]{ return stage(slot ^ 1); }, have);
                HIP(hipEventRecord(ev_freed[slot], stream));
            }
        }
        catch(...)
        {
            (void)hipStreamSynchronize(copy_stream);
            throw;
        }
    }

    // Inserts host sequences (seqs, lens, n) or, with seqs null, a host text.
    uint64_t insert_host(uint32_t user, const char* const* seqs, const size_t* lens, size_t n, uint32_t op, bool oriented, uint64_t x, uint64_t xr,
                         const char* host_text = nullptr, uint64_t text_bytes = 0)
    {
        std::lock_guard<std::mutex> lock(mu);
        CallSched s = sched_begin(user);
        uint64_t inserted = 0;
        try
        {
            const auto work = [&](const uint8_t* text, uint64_t text_len, const auto& stage_next_chunk, bool& more)
            {
                // Launches of at most LAUNCH_POS positions; the next chunk is
                // staged while the first one runs, before its count is read.
                bool staged = false;
                for(uint64_t b = 0; b < text_len; b += LAUNCH_POS)
                {
                    const uint64_t e = std::min(text_len, b + LAUNCH_POS);
                    ov_room(e - b);
                    timed("insert", [&]{ dispatch_insert(op, oriented, text, text_len, b, e, x, xr); });
                    if(!staged)
                        more = stage_next_chunk(), staged = true;
                    const Counters c = read_counters();
                    if(c.failed)
                        throw hip_failure("an insert found no room in the overflow table");
                    inserted += c.inserted, size += c.inserted, ov_count += c.ov_inserted;
                    grow(sched_add(s, c.inserted));
                }
                if(!staged)
                    more = stage_next_chunk();
            };
            if(seqs)
                host_pipeline(seqs, lens, n, work);
            else
                text_pipeline(host_text, text_bytes, work);
        }
        catch(...)
        {
            sched_end(s, user);
            throw;
        }
        sched_end(s, user);
        return inserted;
    }

    uint64_t insert_device(uint32_t user, const uint8_t* text, uint64_t text_len, uint32_t op, bool oriented, uint64_t x, uint64_t xr)
    {
        std::lock_guard<std::mutex> lock(mu);
        CallSched s = sched_begin(user);
        uint64_t inserted = 0;
        try
        {
            inserted = insert_text(s, op, oriented, text, text_len, x, xr);
        }
        catch(...)
        {
            sched_end(s, user);
            throw;
        }
        sched_end(s, user);
        return inserted;
    }

    // --- Pipelined batches ---

    // Inserts the batches [bounds[i], bounds[i + 1]) of a device text, each
    // as one insert call of `user` (its count at out[i]), launching batch
    // i + 1 before batch i's count is read whenever the keys of all batches
    // in flight cannot reach the growth threshold nor fill the overflow
    // table: then no doubling can fall due among them, and the growth
    // schedule, applied batch by batch as their counts arrive, is exact.
    // Each batch counts into its own ring slot.  The next batch's kernel
    // hands a batch's counters to the host (copy, clear, then a sequence
    // flag the host polls); a copy kernel does it instead when the host
    // needs the count before it can launch the next batch.
    uint64_t insert_batches(uint32_t user, const uint8_t* text, const uint64_t* bounds, uint64_t n, uint32_t op, bool oriented,
                            uint64_t x, uint64_t xr, uint64_t* out)
    {
        std::lock_guard<std::mutex> lock(mu);
        ensure_ring();
        // How a batch's counters reach the host: ev >= 0, the ring event
        // recorded after their copy kernel; seq > 0, the flag value the next
        // batch's kernel sets; neither yet while both are -1 / 0.
        struct Flight { uint64_t batch; int slot; uint64_t positions; int ev; uint64_t seq; };
        std::deque<Flight> q;
        uint64_t in_flight = 0, total = 0, next_slot = 0;

        const auto publish = [&](Flight& f)
        {
            publish_counters(ring_set(f.slot), h_ring + f.slot);
            HIP(hipEventRecord(ring_ev[f.slot], stream));
            f.ev = f.slot;
        };

        const auto resolve = [&]()
        {
            if(q.front().ev < 0 && q.front().seq == 0)
                publish(q.front());
            const Flight f = q.front();
            q.pop_front();
            if(f.seq)
            {
                // The flag is set once the next batch's kernel starts; if the
                // stream has drained without it, the hand-off was lost.
                for(uint64_t spins = 1; __atomic_load_n(h_seq + f.slot, __ATOMIC_ACQUIRE) != f.seq; ++spins)
                {
                    __builtin_ia32_pause();
                    if(spins % 4096 == 0 && hipStreamQuery(stream) != hipErrorNotReady &&
                       __atomic_load_n(h_seq + f.slot, __ATOMIC_ACQUIRE) != f.seq)
                        throw hip_failure("a batch's counters were not handed to the host");
                }
            }
            else
                HIP(hipEventSynchronize(ring_ev[f.ev]));
            const Counters c = h_ring[f.slot];
            in_flight -= f.positions;
            if(c.failed)
                throw hip_failure("an insert found no room in the overflow table");
            size += c.inserted, ov_count += c.ov_inserted, total += c.inserted;
            if(out)
                out[f.batch] = c.inserted;
            CallSched s = sched_begin(user);
            const uint32_t d = sched_add(s, c.inserted);
            sched_end(s, user);
            if(d)
            {
                sync();
                grow(d);
            }
        };

        // Whether a launch of `pos` positions could go now: no doubling can
        // fall due among the launches in flight, and the overflow table has a
        // free slot for every position in flight.  (Its load is kept below
        // 0.75 by ov_room whenever the pipeline drains; the positions in
        // flight bound the keys it can gain, of which few spill that far.)
        const auto fits_now = [&](uint64_t pos)
        {
            return q.size() < RING && size + in_flight + pos < resize_th && ov_count + in_flight + pos + 1 <= ov_slots;
        };

        try
        {
            for(uint64_t i = 0; i < n; ++i)
            {
                const uint64_t b = bounds[i], e = bounds[i + 1];
                const uint64_t pos = e - b;
                if(pos > LAUNCH_POS)
                {
                    // A batch too large for one launch goes alone, as a call.
                    while(!q.empty())
                        resolve();
                    sync();
                    CallSched s = sched_begin(user);
                    const uint64_t c = insert_text(s, op, oriented, text + (b & ~uint64_t(15)), e - (b & ~uint64_t(15)), x, xr, b & 15);
                    sched_end(s, user);
                    total += c;
                    if(out)
                        out[i] = c;
                    continue;
                }
                while(!q.empty() && !fits_now(pos))
                    resolve();
                if(q.empty())
                {
                    sync();
                    ov_room(pos);
                }
                const int slot = static_cast<int>(next_slot++ % RING);
                Handoff prev{};
                if(!q.empty() && q.back().ev < 0 && q.back().seq == 0)
                {
                    Flight& p = q.back();
                    p.seq = ++ring_seq;
                    prev = {reinterpret_cast<uint64_t*>(ring_set(p.slot)), reinterpret_cast<uint64_t*>(h_ring + p.slot), h_seq + p.slot, p.seq};
                }
                timed("insert", [&]{ dispatch_insert(op, oriented, text, e, b, e, x, xr, ring_set(slot), prev); });
                q.push_back({i, slot, pos, -1, 0});
                in_flight += pos;
                const uint64_t next = i + 1 < n ? bounds[i + 2] - bounds[i + 1] : 0;
                if(size + in_flight + 2 * next >= resize_th)
                    prepare_resize();
                // The count is needed before the next launch: copy it now.
                if(i + 1 == n || next > LAUNCH_POS || !fits_now(next))
                    publish(q.back());
            }
            while(!q.empty())
                resolve();
            sync();
        }
        catch(...)
        {
            // Counts of launches that were in flight stay in their slots:
            // clear the ring for the next call (it is otherwise zero whenever
            // no launch is in flight).
            (void)hipStreamSynchronize(stream);
            hipLaunchKernelGGL(k_zero64, dim3(16), dim3(256), 0, stream, reinterpret_cast<uint64_t*>(d_ring), RING * CNT_SET_WORDS);
            (void)hipStreamSynchronize(stream);
            throw;
        }
        return total;
    }

    // Lookups of the batches of a device text: hits and k-mer positions of
    // batch i at hits[i], positions[i].  Reads commute, so up to
    // COUNT_GROUP batches run in one launch (k_count_batches), with the
    // same counts as a launch per batch: the ramp-up and tail of a launch
    // per batch cost ~19 us each (one launch over 31 batches of 9.9M
    // positions: 3.0 instead of 3.6 ms).
    static constexpr uint64_t COUNT_GROUP = 256;

    void count_batches(const uint8_t* text, const uint64_t* bounds, uint64_t n, uint64_t* hits, uint64_t* positions) const
    {
        std::lock_guard<std::mutex> lock(mu);
        uint32_t grid = 0;
        by_layout([&](auto L){
            const auto kern = k_count_batches<decltype(L)::value>;
            grid = grid_for(kern, ~uint64_t(0) >> 1);
        });
        const uint64_t waves = uint64_t(grid) * (TILE / 64);
        const uint64_t g = std::min(n, COUNT_GROUP);
        // Per group: its n + 1 bounds and its edge tiles (at most two per
        // bound), the per-wave pairs (u32), the edge sums and the results
        // (u64 pairs); results of all groups after them.
        const uint64_t bounds_w = COUNT_GROUP + 2, tiles_w = 2 * COUNT_GROUP, part_w = g * waves, edge_w = 2 * g;
        d_bcnt.reserve((bounds_w + tiles_w + part_w + edge_w + 2 * n) * 8);
        uint64_t* d_bounds = d_bcnt.as<uint64_t>();
        uint64_t* d_tiles = d_bounds + bounds_w;
        uint32_t* part = reinterpret_cast<uint32_t*>(d_tiles + tiles_w);
        uint64_t* edge = d_tiles + tiles_w + part_w;
        uint64_t* out = edge + edge_w;
        std::vector<uint64_t> staged(bounds_w + tiles_w);
        for(uint64_t g0 = 0; g0 < n; g0 += COUNT_GROUP)
        {
            const uint64_t m = std::min(COUNT_GROUP, n - g0);
            // Edge tiles: wave units (64 positions from origin + 64 u) with a
            // bound of bounds[g0 + 1 .. g0 + m] in (start, start + EDGE_REACH].
            const uint64_t origin = bounds[g0] & ~uint64_t(15);
            uint64_t n_edges = 0;
            for(uint64_t i = 1; i <= m; ++i)
            {
                const uint64_t B = bounds[g0 + i];
                if(B <= origin)
                    continue;
                const uint64_t first = B - origin > EDGE_REACH ? (B - origin - EDGE_REACH + 63) / 64 : 0, last = (B - origin - 1) / 64;
                for(uint64_t u = first; u <= last; ++u)
                    if(n_edges == 0 || staged[bounds_w + n_edges - 1] < u)
                        staged[bounds_w + n_edges++] = u;
            }
            std::copy(bounds + g0, bounds + g0 + m + 1, staged.begin());
            HIP(hipMemcpyAsync(d_bounds, staged.data(), (bounds_w + n_edges) * 8, hipMemcpyHostToDevice, stream));
            const uint64_t part_words = m * waves;      // u32 pairs: one word each
            hipLaunchKernelGGL(k_zero64, dim3(std::max<uint64_t>(1, std::min<uint64_t>((part_words + 255) / 256, 1024))), dim3(256), 0, stream,
                               reinterpret_cast<uint64_t*>(part), part_words);
            hipLaunchKernelGGL(k_zero64, dim3(1), dim3(256), 0, stream, edge, 2 * m);
            HIP(hipGetLastError());
            const uint64_t tiles = (bounds[g0 + m] - origin + TILE - 1) / TILE;
            const uint32_t gr = static_cast<uint32_t>(std::min<uint64_t>(grid, std::max<uint64_t>(tiles, 1)));
            const uint64_t w = uint64_t(gr) * (TILE / 64);
            timed("query", [&]{
                by_layout([&](auto L){
                    hipLaunchKernelGGL(k_count_batches<decltype(L)::value>, dim3(gr), dim3(TILE), 0, stream, view(), text,
                                       static_cast<const uint64_t*>(d_bounds), static_cast<uint32_t>(m), part);
                    if(n_edges)
                        hipLaunchKernelGGL(k_count_edges<decltype(L)::value>, dim3(static_cast<uint32_t>(std::min<uint64_t>(n_edges, 2048))), dim3(64), 0,
                                           stream, view(), text, static_cast<const uint64_t*>(d_bounds), static_cast<uint32_t>(m),
                                           static_cast<const uint64_t*>(d_tiles), n_edges, edge);
                });
            });
            hipLaunchKernelGGL(k_batch_sums, dim3(static_cast<uint32_t>(m)), dim3(256), 0, stream, static_cast<const uint32_t*>(part), w,
                               static_cast<const uint64_t*>(edge), out + 2 * g0);
            HIP(hipGetLastError());
        }
        std::vector<uint64_t> h(2 * n);
        HIP(hipMemcpyAsync(h.data(), out, 2 * n * 8, hipMemcpyDeviceToHost, stream));
        sync();
        for(uint64_t i = 0; i < n; ++i)
        {
            if(hits)
                hits[i] = h[2 * i];
            if(positions)
                positions[i] = h[2 * i + 1];
        }
    }

    static constexpr int RING = 16;         // counter slots of batches in flight
    mutable Counters* d_ring = nullptr;     // RING counter sets (CNT_SET_WORDS words each)
    Counters* ring_set(uint64_t slot) const { return reinterpret_cast<Counters*>(reinterpret_cast<uint64_t*>(d_ring) + slot * CNT_SET_WORDS); }
    mutable Counters* h_ring = nullptr;
    mutable uint64_t* h_seq = nullptr;      // per slot: the last hand-off's sequence number
    mutable uint64_t ring_seq = 0;
    mutable hipEvent_t ring_ev[RING] = {};

    void ensure_ring() const
    {
        if(d_ring)
            return;
        d_ring = reinterpret_cast<Counters*>(dev_alloc<uint64_t>(RING * CNT_SET_WORDS));
        HIP(hipHostMalloc(reinterpret_cast<void**>(&h_ring), RING * sizeof(Counters), hipHostMallocDefault));
        HIP(hipHostMalloc(reinterpret_cast<void**>(&h_seq), RING * sizeof(uint64_t), hipHostMallocDefault));
        std::memset(h_seq, 0, RING * sizeof(uint64_t));
        for(auto& e : ring_ev)
            HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        hipLaunchKernelGGL(k_zero64, dim3(16), dim3(256), 0, stream, reinterpret_cast<uint64_t*>(d_ring), RING * CNT_SET_WORDS);
        HIP(hipGetLastError());
    }

    // --- Updates ---

    template <uint32_t OP>
    void launch_update(const uint8_t* text, uint64_t text_len, uint64_t b, uint64_t e, uint64_t x) const
    {
        const uint64_t tiles = (e - b + TILE - 1) / TILE;
        const auto kern = vw == 4 ? k_update<4, OP> : k_update<8, OP>;
        hipLaunchKernelGGL(kern, dim3(grid_for(kern, tiles)), dim3(TILE), 0, stream, view(), text, text_len, b, e, x, d_cnt);
    }

    uint64_t update_text(uint32_t op, const uint8_t* text, uint64_t text_len, uint64_t x) const
    {
        launch_update_op(op, text, text_len, x);
        return read_counters().present;
    }

    void launch_update_op(uint32_t op, const uint8_t* text, uint64_t text_len, uint64_t x) const
    {
        timed("update", [&]{
            switch(op)
            {
            case OP_KEEP:   launch_update<OP_KEEP>(text, text_len, 0, text_len, x); break;
            case OP_ASSIGN: launch_update<OP_ASSIGN>(text, text_len, 0, text_len, x); break;
            case OP_ADD:    launch_update<OP_ADD>(text, text_len, 0, text_len, x); break;
            case OP_OR:     launch_update<OP_OR>(text, text_len, 0, text_len, x); break;
            case OP_MIN:    launch_update<OP_MIN>(text, text_len, 0, text_len, x); break;
            case OP_MAX:    launch_update<OP_MAX>(text, text_len, 0, text_len, x); break;
            default:        launch_update<OP_ADD_WRAP>(text, text_len, 0, text_len, x); break;
            }
        });
    }

    uint64_t update_host(const char* const* seqs, const size_t* lens, size_t n, uint32_t op, uint64_t x,
                         const char* host_text = nullptr, uint64_t text_bytes = 0)
    {
        std::lock_guard<std::mutex> lock(mu);
        // Counts add up in the device counters; one read at the end.
        const auto work = [&](const uint8_t* text, uint64_t text_len, const auto& stage_next_chunk, bool& more)
        {
            launch_update_op(op, text, text_len, x);
            more = stage_next_chunk();
        };
        if(seqs)
            host_pipeline(seqs, lens, n, work);
        else
            text_pipeline(host_text, text_bytes, work);
        return read_counters().present;
    }

    // --- Queries ---

    template <int MODE>
    void launch_query(const uint8_t* text, uint64_t text_len, const QueryArgs& a) const
    {
        const uint64_t tiles = (text_len + TILE - 1) / TILE;
        timed("query", [&]{
            by_layout([&](auto L){
                const auto kern = k_query<decltype(L)::value, MODE>;
                hipLaunchKernelGGL(kern, dim3(grid_for(kern, tiles)), dim3(TILE), 0, stream, view(), text, uint64_t(0), text_len, a, d_cnt);
            });
        });
    }

    // Uploads the segment tables of a stage and fills `a` with them.
    void upload_segments(const Stage& st, QueryArgs& a) const
    {
        const uint64_t ns = st.seg_q.size();
        const uint64_t tiles = (st.text_len + TILE - 1) / TILE;
        // Layout in one pinned block: seg_start (ns+1), seg_off (ns), seg_q (ns), tile_seg (tiles+1).
        const uint64_t bytes = (ns + 1) * 8 + ns * 8 + ns * 4 + (tiles + 2) * 4 + 64;
        h_misc.reserve(bytes);
        d_misc.reserve(bytes);
        uint8_t* h = h_misc.as<uint8_t>();
        uint64_t* h_start = reinterpret_cast<uint64_t*>(h);
        uint64_t* h_off = h_start + ns + 1;
        uint32_t* h_q = reinterpret_cast<uint32_t*>(h_off + ns);
        uint32_t* h_tile = h_q + ns;
        std::memcpy(h_start, st.seg_start.data(), (ns + 1) * 8);
        std::memcpy(h_off, st.seg_off.data(), ns * 8);
        std::memcpy(h_q, st.seg_q.data(), ns * 4);
        uint32_t sg = 0;
        for(uint64_t t = 0; t <= tiles; ++t)
        {
            const uint64_t p = t * TILE;
            while(sg + 1 < ns && st.seg_start[sg + 1] <= p)
                sg++;
            h_tile[t] = sg;
        }
        HIP(hipMemcpyAsync(d_misc.p, h, bytes, hipMemcpyHostToDevice, stream));
        uint8_t* d = d_misc.as<uint8_t>();
        a.seg_start = reinterpret_cast<const uint64_t*>(d);
        a.seg_off = a.seg_start + ns + 1;
        a.seg_q = reinterpret_cast<const uint32_t*>(a.seg_off + ns);
        a.tile_seg = a.seg_q + ns;
    }

    // Per-sequence counts, sums or masked counts (capi/Kmer_Table.hpp:134-161).
    template <int MODE>
    void per_sequence(const char* const* seqs, const size_t* lens, size_t n, uint64_t* out, uint64_t mf, uint64_t mr) const
    {
        std::lock_guard<std::mutex> lock(mu);
        std::fill(out, out + n, 0);
        if(n == 0)
            return;
        d_out.reserve(n * 8);
        HIP(hipMemsetAsync(d_out.p, 0, n * 8, stream));
        SeqCursor cur;
        Stage st;
        bool any = false;
        while(stage_next(seqs, lens, n, cur, st))
        {
            any = true;
            const uint8_t* text = upload(st);
            QueryArgs a{};
            upload_segments(st, a);
            a.per_seq = d_out.as<uint64_t>();
            a.mask_fwd = mf, a.mask_rev = mr;
            launch_query<MODE>(text, st.text_len, a);
            sync();  // the staging buffers are reused
        }
        if(any)
        {
            HIP(hipMemcpyAsync(out, d_out.p, n * 8, hipMemcpyDeviceToHost, stream));
            sync();
        }
    }

    // The segments of sequences already in a device text: sequence i spans
    // [off[i], off[i + 1] - 1) and its separator is at off[i + 1] - 1.
    static Stage stage_of_offsets(const uint64_t* off, size_t n)
    {
        Stage st;
        st.text_len = off[n];
        st.seg_start.assign(off, off + n + 1);
        st.seg_q.resize(n);
        st.seg_off.assign(n, 0);
        for(size_t i = 0; i < n; ++i)
            st.seg_q[i] = static_cast<uint32_t>(i);
        return st;
    }

    template <int MODE>
    void per_sequence_device(const uint8_t* text, const uint64_t* off, size_t n, uint64_t* out, uint64_t mf, uint64_t mr) const
    {
        std::lock_guard<std::mutex> lock(mu);
        std::fill(out, out + n, 0);
        if(n == 0)
            return;
        d_out.reserve(n * 8);
        HIP(hipMemsetAsync(d_out.p, 0, n * 8, stream));
        const Stage st = stage_of_offsets(off, n);
        QueryArgs a{};
        upload_segments(st, a);
        a.per_seq = d_out.as<uint64_t>();
        a.mask_fwd = mf, a.mask_rev = mr;
        launch_query<MODE>(text, st.text_len, a);
        HIP(hipMemcpyAsync(out, d_out.p, n * 8, hipMemcpyDeviceToHost, stream));
        sync();
    }

    // Where a group of consecutive sequences [first, next group's first)
    // puts its lookup outputs (values and fwd may be null).
    struct LookupDest
    {
        uint64_t first;
        uint8_t* found;
        void* values;
        uint8_t* fwd;
    };

    // capi/Kmer_Table.hpp:163-190.
    void lookup(const char* const* seqs, const size_t* lens, size_t n, uint8_t* found, void* values, uint8_t* fwd) const
    {
        const LookupDest d{0, found, values, fwd};
        lookup_groups(seqs, lens, n, &d, 1);
    }

    // One lookup pass over all sequences; each group's outputs, laid out as
    // a lookup of its own sequences alone, are copied to its destinations.
    // With `dev_text`, the sequences are already there, at offsets `dev_off`
    // (as per_sequence_device takes them), and seqs / lens are unused.
    // Copies `n` bytes of device memory to the caller's memory, in stream
    // order: a small copy, or one into page-locked memory (kache_amd_host_alloc),
    // which the copy engine writes directly, is only enqueued (the caller
    // synchronizes the stream before it reads); a large one into pageable
    // memory goes through two pinned chunks, the copy engine filling chunk
    // i + 1 while the packing threads copy chunk i out, several times the
    // rate of a copy into pageable memory, and has ended when this returns.
    void download(void* dst, const void* src, uint64_t n) const
    {
        constexpr uint64_t CHUNK = uint64_t(8) << 20;
        if(n < 2 * CHUNK || pinned_host(dst))
        {
            HIP(hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, stream));
            return;
        }
        h_dl[0].reserve(CHUNK), h_dl[1].reserve(CHUNK);
        const uint64_t chunks = (n + CHUNK - 1) / CHUNK;
        const auto issue = [&](uint64_t i)
        {
            const uint64_t a = i * CHUNK, len = std::min(CHUNK, n - a);
            HIP(hipMemcpyAsync(h_dl[i % 2].p, static_cast<const char*>(src) + a, len, hipMemcpyDeviceToHost, stream));
            HIP(hipEventRecord(ev_dl[i % 2], stream));
        };
        issue(0);
        for(uint64_t i = 0; i < chunks; ++i)
        {
            // Chunk i + 1's buffer held chunk i - 1, copied out in the last turn.
            if(i + 1 < chunks)
                issue(i + 1);
            HIP(hipEventSynchronize(ev_dl[i % 2]));
            const uint64_t a = i * CHUNK;
            par_copy(static_cast<char*>(dst) + a, h_dl[i % 2].as<char>(), std::min(CHUNK, n - a));
        }
    }

    void lookup_groups(const char* const* seqs, const size_t* lens, size_t n, const LookupDest* g, size_t ng,
                       const uint8_t* dev_text = nullptr, const uint64_t* dev_off = nullptr) const
    {
        std::lock_guard<std::mutex> lock(mu);
        std::vector<uint64_t> o(n + 1);
        uint64_t p = 0;
        for(size_t i = 0; i < n; ++i)
        {
            const uint64_t len = dev_text ? dev_off[i + 1] - dev_off[i] - 1 : lens[i];
            o[i] = p, p += (len >= static_cast<uint64_t>(K) ? len - K + 1 : 0);
        }
        o[n] = p;
        if(p == 0)
            return;
        bool values = false, fwd = false;
        for(size_t i = 0; i < ng; ++i)
            values |= g[i].values != nullptr, fwd |= g[i].fwd != nullptr;

        const uint32_t vb = value_bytes;
        d_out.reserve(p);                       // found
        d_out2.reserve(values ? p * vb : 1);    // values
        d_out3.reserve(fwd ? p : 1);            // fwd
        DevBuf& d_off = d_goff;
        d_off.reserve((n + 1) * 8);
        HIP(hipMemcpyAsync(d_off.p, o.data(), (n + 1) * 8, hipMemcpyHostToDevice, stream));
        fill(d_out.p, p, KACHE_NOT_KMER);
        if(values)
            fill(d_out2.p, p * vb, 0);
        if(fwd)
            fill(d_out3.p, p, 0);

        const auto launch = [&](const uint8_t* text, const Stage& st)
        {
            QueryArgs a{};
            upload_segments(st, a);
            a.out_off = d_off.as<uint64_t>();
            a.found = d_out.as<uint8_t>();
            a.values = values ? d_out2.as<uint8_t>() : nullptr;
            a.fwd = fwd ? d_out3.as<uint8_t>() : nullptr;
            a.vbytes = vb;
            launch_query<Q_LOOKUP>(text, st.text_len, a);
            sync();
        };
        if(dev_text)
            launch(dev_text, stage_of_offsets(dev_off, n));
        else
        {
            SeqCursor cur;
            Stage st;
            while(stage_next(seqs, lens, n, cur, st))
                launch(upload(st), st);
        }
        for(size_t i = 0; i < ng; ++i)
        {
            const uint64_t b = o[g[i].first], e = o[i + 1 < ng ? g[i + 1].first : n];
            if(e == b)
                continue;
            download(g[i].found, d_out.as<uint8_t>() + b, e - b);
            if(g[i].values)
                download(g[i].values, d_out2.as<uint8_t>() + b * vb, (e - b) * vb);
            if(g[i].fwd)
                download(g[i].fwd, d_out3.as<uint8_t>() + b, e - b);
        }
        sync();
    }

    void fill(void* p, uint64_t n, uint8_t v) const
    {
        const uint32_t g = static_cast<uint32_t>(std::min<uint64_t>((n + TILE - 1) / TILE, 4096));
        hipLaunchKernelGGL(k_fill, dim3(std::max<uint32_t>(g, 1)), dim3(TILE), 0, stream, static_cast<uint8_t*>(p), n, v);
        HIP(hipGetLastError());
    }

    // per_sequence_device<Q_COUNT> without allocating: the per-sequence
    // counters and segment tables live in the device's scratch, built there
    // by k_seg_setup from the offsets; the offsets are copied in and the
    // counts out by plain copies.  Calls that use one device's scratch are
    // serialized by its lock.
    void count_seqs_static(const uint8_t* text, const uint64_t* off, size_t n, uint64_t* out) const
    {
        std::lock_guard<std::mutex> lock(mu);
        std::lock_guard<std::mutex> one(dscratch->calls);
        std::fill(out, out + n, 0);
        if(n == 0)
            return;
        uint8_t* const scr = dscratch->ptr;
        uint64_t* const per_seq = reinterpret_cast<uint64_t*>(scr + SCR_PER_SEQ);
        uint64_t* const seg_start = reinterpret_cast<uint64_t*>(scr + SCR_SEG_START);
        uint32_t* const seg_q = reinterpret_cast<uint32_t*>(scr + SCR_SEG_Q);
        uint32_t* const tile_seg = reinterpret_cast<uint32_t*>(scr + SCR_TILE_SEG);
        const uint64_t text_len = off[n], tiles = (text_len + TILE - 1) / TILE;
        HIP(hipMemcpyAsync(seg_start, off, (n + 1) * 8, hipMemcpyHostToDevice, stream));
        const uint64_t setup_blocks = std::min<uint64_t>((std::max<uint64_t>(n, tiles + 1) + TILE - 1) / TILE, 4096);
        hipLaunchKernelGGL(k_seg_setup, dim3(setup_blocks), dim3(TILE), 0, stream, seg_start, uint64_t(n), tiles, seg_q, tile_seg, per_seq);
        HIP(hipGetLastError());
        QueryArgs a{};
        a.seg_start = seg_start, a.seg_q = seg_q, a.tile_seg = tile_seg, a.per_seq = per_seq;
        by_layout([&](auto L){
            const auto kern = k_query<decltype(L)::value, Q_COUNT>;
            hipLaunchKernelGGL(kern, dim3(grid_for(kern, tiles)), dim3(TILE), 0, stream, view(), text, uint64_t(0), text_len, a, d_cnt);
        });
        HIP(hipGetLastError());
        HIP(hipMemcpyAsync(out, per_seq, n * 8, hipMemcpyDeviceToHost, stream));
        HIP(hipStreamSynchronize(stream));
    }

    Counters count_device(const uint8_t* text, uint64_t text_len) const
    {
        std::lock_guard<std::mutex> lock(mu);
        QueryArgs a{};
        launch_query<Q_TOTAL>(text, text_len, a);
        return read_counters();
    }

    // --- Scans ---

    // capi/Kmer_Table.hpp:232-268 over include/Streaming_Kmer_Hash_Table.hpp:2703-2725:
    // share `shard` of `shards` is buckets [cap*w/W, cap*(w+1)/W) and its
    // share of the overflow slots, handed out in chunks of at most 4096 keys.
    void for_each(uint32_t shards, uint32_t shard, kache_keys_fn f, void* ctx) const
    {
        std::lock_guard<std::mutex> lock(mu);
        const uint64_t b0 = cap * shard / shards, b1 = cap * (shard + 1) / shards;
        const uint64_t range = (ov_slots + shards - 1) / shards;
        const uint64_t o0 = std::min<uint64_t>(shard * range, ov_slots), o1 = std::min<uint64_t>((shard + 1) * range, ov_slots);

        constexpr size_t chunk = 4096;
        std::vector<uint64_t> K_(chunk), Vw;
        std::vector<uint8_t> Vb(map ? chunk * value_bytes : 0);
        bool done = false;
        std::vector<uint64_t> hk, hv;

        const auto deliver = [&](uint64_t n)
        {
            for(uint64_t i = 0; i < n && !done; i += chunk)
            {
                const uint64_t c = std::min<uint64_t>(chunk, n - i);
                std::memcpy(K_.data(), hk.data() + i, c * 8);
                if(map)
                    for(uint64_t j = 0; j < c; ++j)
                        std::memcpy(Vb.data() + j * value_bytes, &hv[i + j], value_bytes);   // little-endian low bytes
                done = (f(ctx, K_.data(), map ? Vb.data() : nullptr, c) != 0);
            }
        };

        // Pieces of the share: bucket ranges, then the overflow range.
        struct Piece { uint64_t b0, b1, o0, o1; };
        std::vector<Piece> pieces;
        for(uint64_t b = b0; b < b1; b += COLLECT_BUCKETS)
            pieces.push_back({b, std::min(b1, b + COLLECT_BUCKETS), 0, 0});
        for(uint64_t o = o0; o < o1; o += COLLECT_BUCKETS * B)
            pieces.push_back({0, 0, o, std::min(o1, o + COLLECT_BUCKETS * B)});

        DevBuf dk, dv, dc;
        dk.reserve(COLLECT_BUCKETS * B * 8);
        if(map)
            dv.reserve(COLLECT_BUCKETS * B * 8);
        dc.reserve(8);
        for(const auto& pc : pieces)
        {
            if(done)
                break;
            HIP(hipMemsetAsync(dc.p, 0, 8, stream));
            const uint64_t work = (pc.b1 - pc.b0) * B + (pc.o1 - pc.o0);
            const uint64_t blocks = (work + TILE - 1) / TILE;
            timed("collect", [&]{
                by_layout([&](auto L){
                    const auto kern = k_collect<decltype(L)::value>;
                    hipLaunchKernelGGL(kern, dim3(grid_for(kern, blocks)), dim3(TILE), 0, stream, view(), pc.b0, pc.b1, pc.o0, pc.o1, dk.as<uint64_t>(),
                                       map ? dv.as<uint64_t>() : static_cast<uint64_t*>(nullptr), dc.as<uint64_t>());
                });
            });
            uint64_t n = 0;
            HIP(hipMemcpyAsync(&n, dc.p, 8, hipMemcpyDeviceToHost, stream));
            sync();
            if(n == 0)
                continue;
            hk.resize(n);
            HIP(hipMemcpyAsync(hk.data(), dk.p, n * 8, hipMemcpyDeviceToHost, stream));
            if(map)
            {
                hv.resize(n);
                HIP(hipMemcpyAsync(hv.data(), dv.p, n * 8, hipMemcpyDeviceToHost, stream));
            }
            sync();
            deliver(n);
        }
    }

    uint64_t reduce(uint32_t what, uint64_t mask) const
    {
        std::lock_guard<std::mutex> lock(mu);
        DevBuf out;
        out.reserve(8);
        HIP(hipMemsetAsync(out.p, 0, 8, stream));
        const uint64_t blocks = (cap * B + ov_slots + TILE - 1) / TILE;
        timed("reduce", [&]{
            by_layout([&](auto L){
                const auto kern = k_reduce<decltype(L)::value>;
                hipLaunchKernelGGL(kern, dim3(grid_for(kern, blocks)), dim3(TILE), 0, stream, view(), what, mask, out.as<uint64_t>());
            });
        });
        uint64_t r = 0;
        HIP(hipMemcpyAsync(&r, out.p, 8, hipMemcpyDeviceToHost, stream));
        sync();
        return r;
    }

    // Histogram of the values in [1, max] (counter/Kmer_Set_Counter.cpp:97-120):
    // out[c] = keys with value c, out[0] = 0.
    void histogram(uint64_t max, uint64_t* out) const
    {
        std::lock_guard<std::mutex> lock(mu);
        DevBuf d;
        d.reserve((max + 1) * 8);
        HIP(hipMemsetAsync(d.p, 0, (max + 1) * 8, stream));
        const uint64_t blocks = (cap * B + ov_slots + TILE - 1) / TILE;
        const uint64_t lds = (max + 1) * 4;
        timed("histogram", [&]{
            if(lds <= 48 * 1024)
            {
                const auto kern = vw == 4 ? k_histogram<4, true> : k_histogram<8, true>;
                int per_cu = 0;
                if(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, TILE, lds) != hipSuccess || per_cu < 1)
                    per_cu = 1;
                const uint64_t g = std::min<uint64_t>(blocks, uint64_t(per_cu) * device_cus());
                hipLaunchKernelGGL(kern, dim3(std::max<uint64_t>(g, 1)), dim3(TILE), lds, stream, view(), max, d.as<uint64_t>());
            }
            else
            {
                const auto kern = vw == 4 ? k_histogram<4, false> : k_histogram<8, false>;
                hipLaunchKernelGGL(kern, dim3(grid_for(kern, blocks)), dim3(TILE), 0, stream, view(), max, d.as<uint64_t>());
            }
        });
        HIP(hipMemcpyAsync(out, d.p, (max + 1) * 8, hipMemcpyDeviceToHost, stream));
        sync();
        out[0] = 0;
    }

    // --- Persistence: the device image with its host state ---

    struct FileHeader
    {
        char magic[8];
        uint32_t version;
        uint32_t value_bytes;
        double lf;
        uint64_t cap, ov_slots, ov_count, size, resizes, approx_sz, resize_th, step;
        uint32_t user_c, reserved;
        uint64_t checkpoint[KACHE_MAX_USERS];
        uint64_t data_bytes;
        uint64_t data_hash;
        uint64_t header_hash;       // of the bytes before it
    };
    static constexpr char MAGIC[8] = {'K', 'A', 'M', 'D', 'T', 'B', 'L', '1'};
    // Version 3: values of widths 1-4 in 4-byte words.
    static constexpr uint32_t FILE_VERSION = 3;
    static constexpr uint64_t IO_CHUNK = uint64_t(64) << 20;

    // A word-wise hash of file bytes (a multiply-rotate mix per 8 bytes).
    static uint64_t hash_words(uint64_t h, const void* p, uint64_t n)
    {
        const uint64_t* w = static_cast<const uint64_t*>(p);
        for(uint64_t i = 0; i < n / 8; ++i)
        {
            h ^= w[i];
            h *= 0x9E3779B97F4A7C15ull;
            h = (h << 29) | (h >> 35);
        }
        return h;
    }

    // The device regions of the image, in file order.
    std::vector<std::pair<uint8_t*, uint64_t>> regions() const
    {
        return {{buckets, cap * bucket_bytes()}, {reinterpret_cast<uint8_t*>(ov), ov_bytes(ov_slots)}};
    }

    void save(const char* path) const
    {
        std::lock_guard<std::mutex> lock(mu);
        sync();
        FileHeader h{};
        std::memcpy(h.magic, MAGIC, 8);
        h.version = FILE_VERSION, h.value_bytes = value_bytes, h.lf = lf;
        h.cap = cap, h.ov_slots = ov_slots, h.ov_count = ov_count, h.size = size, h.resizes = resizes;
        h.approx_sz = approx_sz, h.resize_th = resize_th, h.step = step;
        h.user_c = user_c.load();
        std::memcpy(h.checkpoint, checkpoint, sizeof(checkpoint));
        for(const auto& [p, n] : regions())
            h.data_bytes += n;
        FILE* f = std::fopen(path, "wb");
        if(!f)
            throw std::invalid_argument(std::string("cannot create ") + path + ": " + std::strerror(errno));
        std::unique_ptr<FILE, int (*)(FILE*)> guard_f(f, std::fclose);
        if(std::fwrite(&h, sizeof(h), 1, f) != 1)
            throw std::runtime_error("write failed");
        HostBuf buf;
        buf.reserve(IO_CHUNK);
        uint64_t dh = 0x6B616368655F616Dull;
        for(const auto& [p, n] : regions())
            for(uint64_t o = 0; o < n; o += IO_CHUNK)
            {
                const uint64_t c = std::min(IO_CHUNK, n - o);
                HIP(hipMemcpyAsync(buf.p, p + o, c, hipMemcpyDeviceToHost, stream));
                HIP(hipStreamSynchronize(stream));
                dh = hash_words(dh, buf.p, c);
                if(std::fwrite(buf.p, 1, c, f) != c)
                    throw std::runtime_error("write failed");
            }
        h.data_hash = dh;
        h.header_hash = hash_words(0x6B616368655F6864ull, &h, offsetof(FileHeader, header_hash));
        if(std::fseek(f, 0, SEEK_SET) != 0 || std::fwrite(&h, sizeof(h), 1, f) != 1)
            throw std::runtime_error("write failed");
        if(std::fflush(f) != 0)
            throw std::runtime_error("write failed");
    }

    // Reads and checks a file's header (no device work).
    static FileHeader read_header(FILE* f)
    {
        FileHeader h{};
        if(std::fread(&h, sizeof(h), 1, f) != 1 || std::memcmp(h.magic, MAGIC, 8) != 0 || h.version != FILE_VERSION)
            throw std::invalid_argument("not a table file of this version");
        if(h.header_hash != hash_words(0x6B616368655F6864ull, &h, offsetof(FileHeader, header_hash)))
            throw std::invalid_argument("the table file's header is corrupt");
        const uint32_t vb = h.value_bytes;
        const uint32_t w = value_word(vb);
        if(!(vb == 0 || vb == 1 || vb == 2 || vb == 4 || vb == 8) || !(h.lf > 0 && h.lf < 1) || h.cap == 0 || h.cap > MAX_CAP ||
           h.ov_slots < OV_MIN || (h.ov_slots & (h.ov_slots - 1)) || h.user_c > KACHE_MAX_USERS ||
           h.data_bytes != h.cap * bucket_bytes_of(w) + h.ov_slots * (8 + w))
            throw std::invalid_argument("the table file's header is inconsistent");
        return h;
    }

    // Replaces this (new, empty) table's store by the file's image and state.
    void load(FILE* f, const FileHeader& h)
    {
        std::lock_guard<std::mutex> lock(mu);
        sync();
        // Arrays prepared for the old table's next doubling are of no use now.
        wait_frees();
        if(prepared.valid())
            { Prepared p = prepared.get(); free_prepared(p); }
        uint8_t* nb = take_buckets(h.cap);
        uint64_t* nov = nullptr;
        try
        {
            nov = static_cast<uint64_t*>(ArrayCache::get().take(device, ov_bytes(h.ov_slots)));
        }
        catch(...)
        {
            ArrayCache::get().give(device, nb, h.cap * bucket_bytes());
            throw;
        }
        ArrayCache::get().give(device, buckets, cap * bucket_bytes());
        ArrayCache::get().give(device, ov, ov_bytes(ov_slots));
        buckets = nb, ov = nov, cap = h.cap, ov_slots = h.ov_slots;
        HostBuf buf;
        buf.reserve(IO_CHUNK);
        uint64_t dh = 0x6B616368655F616Dull;
        for(const auto& [p, n] : regions())
            for(uint64_t o = 0; o < n; o += IO_CHUNK)
            {
                const uint64_t c = std::min(IO_CHUNK, n - o);
                if(std::fread(buf.p, 1, c, f) != c)
                    throw std::invalid_argument("the table file is truncated");
                dh = hash_words(dh, buf.p, c);
                HIP(hipMemcpyAsync(p + o, buf.p, c, hipMemcpyHostToDevice, stream));
                HIP(hipStreamSynchronize(stream));
            }
        if(dh != h.data_hash)
            throw std::invalid_argument("the table file's data is corrupt");
        ov_count = h.ov_count, size = h.size, resizes = h.resizes;
        approx_sz = h.approx_sz, resize_th = h.resize_th, step = h.step;
        std::memcpy(checkpoint, h.checkpoint, sizeof(checkpoint));
        user_c.store(h.user_c);
    }

    // capi/Kmer_Table.hpp:270-279 with the placement of
    // include/Streaming_Kmer_Hash_Table.hpp:2728-2752.
    void stats(kache_stats& s) const
    {
        std::lock_guard<std::mutex> lock(mu);
        s = kache_stats{};
        s.capacity = cap * B;
        s.bytes = cap * bucket_bytes();
        s.keys = size;
        s.overflow_keys = ov_count;
        if(size == 0)
            return;
        DevBuf out;
        out.reserve(16);
        HIP(hipMemsetAsync(out.p, 0, 16, stream));
        const uint64_t blocks = (cap * B + TILE - 1) / TILE;
        by_layout([&](auto L){
            const auto kern = k_placement<decltype(L)::value>;
            hipLaunchKernelGGL(kern, dim3(grid_for(kern, blocks)), dim3(TILE), 0, stream, view(), out.as<uint64_t>());
        });
        HIP(hipGetLastError());
        uint64_t ps[2] = {0, 0};
        HIP(hipMemcpyAsync(ps, out.p, 16, hipMemcpyDeviceToHost, stream));
        sync();
        const double all = static_cast<double>(ps[0] + ps[1] + ov_count);
        s.primary_pct = ps[0] * 100.0 / all;
        s.secondary_pct = ps[1] * 100.0 / all;
        s.overflow_pct = ov_count * 100.0 / all;
    }
};


// ---------------------------------------------------------------------------
// The C ABI.

namespace
{

// capi/kache_capi.cpp:34-46.
bool valid(const char* const* seqs, const size_t* lens, size_t n)
{
    if(n == 0)
        return true;
    if(seqs == nullptr || lens == nullptr)
        return false;
    for(size_t i = 0; i < n; ++i)
        if(seqs[i] == nullptr && lens[i] > 0)
            return false;
    return true;
}

// capi/kache_capi.cpp:50-53.
bool fits(const kache_table* t, uint64_t x)
{
    return t->value_bytes > 0 && (t->value_bytes == 8 || x >> (8 * t->value_bytes) == 0);
}

// capi/kache_capi.cpp:57-60.
bool registered(const kache_table* t, uint32_t user) { return user < t->user_c.load(std::memory_order_acquire); }

// guard() for a call that reads or changes the keys of table `t`: a table
// that a failed in-place doubling left partly split refuses it.
template <typename F>
int guard(const kache_table* t, const F& f) noexcept
{
    if(t->damaged.load(std::memory_order_acquire))
    {
        last_error = t->damage;
        return KACHE_EINTERNAL;
    }
    return guard(f);
}

bool device_text_ok(const char* text, uint64_t bytes)
{
    return bytes == 0 || (text != nullptr && reinterpret_cast<uintptr_t>(text) % 16 == 0);
}

}

extern "C"
{

int kache_supports(uint32_t k, uint32_t value_bytes) noexcept
{
    return k == 31 && (value_bytes == 0 || value_bytes == 1 || value_bytes == 2 || value_bytes == 4 || value_bytes == 8);
}

kache_options kache_default_options(uint32_t k, uint32_t value_bytes) noexcept
{
    return kache_options{k, value_bytes, uint64_t(1) << 22, 0.8, 16, 0};
}

int kache_new(const kache_options* options, kache_table** out) noexcept
{
    if(out == nullptr)
        return KACHE_EINVAL;
    *out = nullptr;
    if(options == nullptr)
        return KACHE_EINVAL;
    const auto& o = *options;
    if(!kache_supports(o.k, o.value_bytes) || !(o.load_factor > 0 && o.load_factor < 1) || o.resize_workers < 1 ||
       o.pow2_capacity > 1 || !size_fits(o.expected_keys, o.load_factor, o.pow2_capacity != 0))
        return KACHE_EINVAL;

    return guard([&]()
    {
        int n = 0;
        if(hipGetDeviceCount(&n) != hipSuccess || n < 1)
        {
            (void)hipGetLastError();
            last_error = "no HIP device";
            return static_cast<int>(KACHE_EINTERNAL);
        }
        *out = new kache_table(o);
        return static_cast<int>(KACHE_OK);
    });
}

void kache_free(kache_table* t) noexcept
{
    delete t;
}

int kache_register(kache_table* t, uint32_t* user) noexcept
{
    if(t == nullptr || user == nullptr)
        return KACHE_EINVAL;
    return guard([&]()
    {
        const std::lock_guard<std::mutex> lock(t->reg);
        const auto u = t->user_c.load(std::memory_order_relaxed);
        if(u == KACHE_MAX_USERS)
            return static_cast<int>(KACHE_EUSERS);
        t->user_c.store(u + 1, std::memory_order_release);
        *user = u;
        return static_cast<int>(KACHE_OK);
    });
}

int kache_insert(kache_table* t, uint32_t user, const char* const* seqs, const size_t* lens, size_t n, uint32_t op, uint64_t x, uint64_t* inserted) noexcept
{
    if(t == nullptr || !registered(t, user) || !valid(seqs, lens, n) || (t->value_bytes > 0 && (op > KACHE_MAX || !fits(t, x))))
        return KACHE_EINVAL;
    return guard(t, [&]()
    {
        const uint64_t c = t->insert_host(user, seqs, lens, n, t->map ? op : 0, false, t->map ? x : 0, 0);
        if(inserted != nullptr)
            *inserted = c;
        return static_cast<int>(KACHE_OK);
    });
}

int kache_insert_oriented(kache_table* t, uint32_t user, const char* const* seqs, const size_t* lens, size_t n, uint32_t op, uint64_t x_fwd, uint64_t x_rev, uint64_t* inserted) noexcept
{
    const bool op_ok = (op == KACHE_ASSIGN || op == KACHE_ADD || op == KACHE_OR || op == KACHE_MAX);
    if(t == nullptr || !registered(t, user) || !valid(seqs, lens, n) || !op_ok || !fits(t, x_fwd) || !fits(t, x_rev))
        return KACHE_EINVAL;
    return guard(t, [&]()
    {
        const uint64_t c = t->insert_host(user, seqs, lens, n, op, true, x_fwd, x_rev);
        if(inserted != nullptr)
            *inserted = c;
        return static_cast<int>(KACHE_OK);
    });
}

int kache_update(kache_table* t, const char* const* seqs, const size_t* lens, size_t n, uint32_t op, uint64_t x, uint64_t* present) noexcept
{
    if(t == nullptr || !valid(seqs, lens, n) || op > KACHE_MAX || !fits(t, x))
        return KACHE_EINVAL;
    return guard(t, [&]()
    {
        const uint64_t c = t->update_host(seqs, lens, n, op, x);
        if(present != nullptr)
            *present = c;
        return static_cast<int>(KACHE_OK);
    });
}

int kache_count(const kache_table* t, const char* const* seqs, const size_t* lens, size_t n, uint64_t* counts) noexcept
{
    if(t == nullptr || !valid(seqs, lens, n) || (n > 0 && counts == nullptr))
        return KACHE_EINVAL;
    return guard(t, [&](){ t->per_sequence<Q_COUNT>(seqs, lens, n, counts, 0, 0); return static_cast<int>(KACHE_OK); });
}

int kache_count_masked(const kache_table* t, const char* const* seqs, const size_t* lens, size_t n, uint64_t mask_fwd, uint64_t mask_rev, uint64_t* counts) noexcept
{
    if(t == nullptr || t->value_bytes == 0 || !valid(seqs, lens, n) || (n > 0 && counts == nullptr))
        return KACHE_EINVAL;
    // Masks act on values of the table's width (capi/Kmer_Table.hpp:146).
    const uint64_t m = t->vmax();
    return guard(t, [&](){ t->per_sequence<Q_MASKED>(seqs, lens, n, counts, mask_fwd & m, mask_rev & m); return static_cast<int>(KACHE_OK); });
}

int kache_sum(const kache_table* t, const char* const* seqs, const size_t* lens, size_t n, uint64_t* sums) noexcept
{
    if(t == nullptr || t->value_bytes == 0 || !valid(seqs, lens, n) || (n > 0 && sums == nullptr))
        return KACHE_EINVAL;
    return guard(t, [&](){ t->per_sequence<Q_SUM>(seqs, lens, n, sums, 0, 0); return static_cast<int>(KACHE_OK); });
}

int kache_lookup(const kache_table* t, const char* const* seqs, const size_t* lens, size_t n, uint8_t* found, void* values, uint8_t* fwd) noexcept
{
    if(t == nullptr || !valid(seqs, lens, n) || (n > 0 && found == nullptr) || (t->value_bytes == 0 && values != nullptr))
        return KACHE_EINVAL;
    return guard(t, [&](){ t->lookup(seqs, lens, n, found, values, fwd); return static_cast<int>(KACHE_OK); });
}

namespace
{

// Offsets of sequences in a device text: nondecreasing by at least 1 (each
// sequence then its separator), within the text.
bool offsets_ok(const char* text, uint64_t text_bytes, const uint64_t* off, size_t n)
{
    if(n > 0 && (off == nullptr || !device_text_ok(text, text_bytes) || text == nullptr))
        return false;
    for(size_t i = 0; i < n; ++i)
        if(off[i + 1] < off[i] + 1)
            return false;
    return n == 0 || off[n] <= text_bytes;
}

}

int kache_amd_count_seqs_device(const kache_table* t, const char* text, uint64_t text_bytes, const uint64_t* offsets, size_t n, uint64_t* counts) noexcept
{
    if(t == nullptr || !offsets_ok(text, text_bytes, offsets, n) || (n > 0 && counts == nullptr))
        return KACHE_EINVAL;
    return guard(t, [&]()
    {
        t->per_sequence_device<Q_COUNT>(reinterpret_cast<const uint8_t*>(text), offsets, n, counts, 0, 0);
        return static_cast<int>(KACHE_OK);
    });
}

int kache_amd_count_seqs_device_static(const kache_table* t, const char* text, uint64_t text_bytes, const uint64_t* offsets, size_t n,
                                       uint64_t* counts) noexcept
{
    if(t == nullptr || !offsets_ok(text, text_bytes, offsets, n) || (n > 0 && counts == nullptr))
        return KACHE_EINVAL;
    if(n > STATIC_MAX_SEQS || (n > 0 && offsets[n] > STATIC_MAX_TEXT))
    {
        last_error = "kache_amd_count_seqs_device_static: " + std::to_string(n) + " sequences over " +
                     std::to_string(n ? offsets[n] : 0) + " text bytes exceed its scratch (at most " +
                     std::to_string(STATIC_MAX_SEQS) + " sequences and " + std::to_string(STATIC_MAX_TEXT) + " bytes)";
        return KACHE_EINVAL;
    }
    return guard(t, [&]()
    {
        t->count_seqs_static(reinterpret_cast<const uint8_t*>(text), offsets, n, counts);
        return static_cast<int>(KACHE_OK);
    });
}

int kache_amd_lookup_groups_device(const kache_table* t, const char* text, uint64_t text_bytes, const uint64_t* offsets, size_t n,
                                   const uint64_t* group_first, size_t n_groups, uint8_t* const* found, void* const* values,
                                   uint8_t* const* fwd) noexcept
{
    if(t == nullptr || !offsets_ok(text, text_bytes, offsets, n) || (n_groups > 0 && (group_first == nullptr || found == nullptr)))
        return KACHE_EINVAL;
    for(size_t i = 0; i < n_groups; ++i)
        if(group_first[i] > n || (i > 0 && group_first[i] < group_first[i - 1]) || found[i] == nullptr ||
           (values && values[i] && t->value_bytes == 0))
            return KACHE_EINVAL;
    if(n_groups > 0 && group_first[0] != 0)
        return KACHE_EINVAL;
    if(n_groups == 0)
        return n == 0 ? static_cast<int>(KACHE_OK) : static_cast<int>(KACHE_EINVAL);
    return guard(t, [&]()
    {
        std::vector<kache_table::LookupDest> d(n_groups);
        for(size_t i = 0; i < n_groups; ++i)
            d[i] = {group_first[i], found[i], values ? values[i] : nullptr, fwd ? fwd[i] : nullptr};
        t->lookup_groups(nullptr, nullptr, n, d.data(), n_groups, reinterpret_cast<const uint8_t*>(text), offsets);
        return static_cast<int>(KACHE_OK);
    });
}

int kache_amd_lookup_groups(const kache_table* t, const char* const* seqs, const size_t* lens, size_t n, const uint64_t* group_first,
                            size_t n_groups, uint8_t* const* found, void* const* values, uint8_t* const* fwd) noexcept
{
    if(t == nullptr || !valid(seqs, lens, n) || (n_groups > 0 && (group_first == nullptr || found == nullptr)))
        return KACHE_EINVAL;
    for(size_t i = 0; i < n_groups; ++i)
        if(group_first[i] > n || (i > 0 && group_first[i] < group_first[i - 1]) || found[i] == nullptr ||
           (values && values[i] && t->value_bytes == 0))
            return KACHE_EINVAL;
    if(n_groups > 0 && group_first[0] != 0)
        return KACHE_EINVAL;
    if(n_groups == 0)
        return n == 0 ? static_cast<int>(KACHE_OK) : static_cast<int>(KACHE_EINVAL);
    return guard(t, [&]()
    {
        std::vector<kache_table::LookupDest> d(n_groups);
        for(size_t i = 0; i < n_groups; ++i)
            d[i] = {group_first[i], found[i], values ? values[i] : nullptr, fwd ? fwd[i] : nullptr};
        t->lookup_groups(seqs, lens, n, d.data(), n_groups);
        return static_cast<int>(KACHE_OK);
    });
}

int kache_amd_insert_seqs(kache_table* t, uint32_t user, const char* const* seqs, const size_t* lens, size_t n, uint32_t op, uint64_t x, uint64_t* inserted) noexcept
{
    if(t == nullptr || !registered(t, user) || !valid(seqs, lens, n) ||
       (t->map && ((op > KACHE_MAX && op != KACHE_AMD_ADD_WRAP) || !fits(t, x))))
        return KACHE_EINVAL;
    return guard(t, [&]()
    {
        const uint64_t c = t->insert_host(user, seqs, lens, n, t->map ? op : 0, false, t->map ? x : 0, 0);
        if(inserted != nullptr)
            *inserted = c;
        return static_cast<int>(KACHE_OK);
    });
}

int kache_amd_update_seqs(kache_table* t, const char* const* seqs, const size_t* lens, size_t n, uint32_t op, uint64_t x, uint64_t* present) noexcept
{
    if(t == nullptr || !valid(seqs, lens, n) || (op > KACHE_MAX && op != KACHE_AMD_ADD_WRAP) || !fits(t, x))
        return KACHE_EINVAL;
    return guard(t, [&]()
    {
        const uint64_t c = t->update_host(seqs, lens, n, op, x);
        if(present != nullptr)
            *present = c;
        return static_cast<int>(KACHE_OK);
    });
}

int kache_amd_insert_text(kache_table* t, uint32_t user, const char* text, uint64_t text_bytes, uint32_t op, uint64_t x, uint64_t* inserted) noexcept
{
    if(t == nullptr || !registered(t, user) || (text == nullptr && text_bytes > 0) ||
       (t->map && ((op > KACHE_MAX && op != KACHE_AMD_ADD_WRAP) || !fits(t, x))))
        return KACHE_EINVAL;
    return guard(t, [&]()
    {
        const uint64_t c = text_bytes ? t->insert_host(user, nullptr, nullptr, 0, t->map ? op : 0, false, t->map ? x : 0, 0, text, text_bytes) : 0;
        if(inserted != nullptr)
            *inserted = c;
        return static_cast<int>(KACHE_OK);
    });
}

int kache_amd_update_text(kache_table* t, const char* text, uint64_t text_bytes, uint32_t op, uint64_t x, uint64_t* present) noexcept
{
    if(t == nullptr || (text == nullptr && text_bytes > 0) || (op > KACHE_MAX && op != KACHE_AMD_ADD_WRAP) || !fits(t, x))
        return KACHE_EINVAL;
    return guard(t, [&]()
    {
        const uint64_t c = text_bytes ? t->update_host(nullptr, nullptr, 0, op, x, text, text_bytes) : 0;
        if(present != nullptr)
            *present = c;
        return static_cast<int>(KACHE_OK);
    });
}

int kache_amd_save(const kache_table* t, const char* path) noexcept
{
    if(t == nullptr || path == nullptr)
        return KACHE_EINVAL;
    if(t->damaged.load(std::memory_order_acquire))
    {
        last_error = t->damage;
        return KACHE_EINTERNAL;
    }
    try
    {
        t->save(path);
        return KACHE_OK;
    }
    catch(const std::invalid_argument& e)
    {
        last_error = e.what();
        return KACHE_EINVAL;
    }
    catch(...)
    {
        return guard([]() -> int { throw; });
    }
}

int kache_amd_load(const char* path, kache_table** out) noexcept
{
    if(path == nullptr || out == nullptr)
        return KACHE_EINVAL;
    *out = nullptr;
    FILE* f = std::fopen(path, "rb");
    if(!f)
    {
        last_error = std::string("cannot open ") + path + ": " + std::strerror(errno);
        return KACHE_EINVAL;
    }
    std::unique_ptr<FILE, int (*)(FILE*)> guard_f(f, std::fclose);
    kache_table::FileHeader h{};
    try
    {
        h = kache_table::read_header(f);
    }
    catch(const std::exception& e)
    {
        last_error = e.what();
        return KACHE_EINVAL;
    }
    kache_options o = kache_default_options(31, h.value_bytes);
    o.expected_keys = 1, o.load_factor = h.lf;
    kache_table* t = nullptr;
    const int st = kache_new(&o, &t);
    if(st != KACHE_OK)
        return st;
    try
    {
        t->load(f, h);
    }
    catch(const std::invalid_argument& e)
    {
        kache_free(t);
        last_error = e.what();
        return KACHE_EINVAL;
    }
    catch(...)
    {
        kache_free(t);
        return guard([]() -> int { throw; });
    }
    *out = t;
    return KACHE_OK;
}

int kache_amd_histogram(const kache_table* t, uint64_t max, uint64_t* out) noexcept
{
    if(t == nullptr || out == nullptr || t->value_bytes == 0 || max > (uint64_t(1) << 30))
        return KACHE_EINVAL;
    return guard(t, [&](){ t->histogram(max, out); return static_cast<int>(KACHE_OK); });
}

int kache_get(const kache_table* t, const char* kmer, uint8_t* found, uint64_t* value) noexcept
{
    if(t == nullptr || kmer == nullptr || found == nullptr)
        return KACHE_EINVAL;
    return guard(t, [&]()
    {
        // The k characters at `kmer`, looked up as a one-position sequence
        // (include/Streaming_Kmer_Hash_Table.hpp:2607-2645).
        const size_t len = K;
        uint8_t f = KACHE_NOT_KMER;
        uint64_t v = 0;
        t->lookup(&kmer, &len, 1, &f, t->map ? &v : nullptr, nullptr);
        *found = (f == KACHE_PRESENT);
        if(value != nullptr)
            *value = (f == KACHE_PRESENT) ? v : 0;
        return static_cast<int>(KACHE_OK);
    });
}

int kache_size(const kache_table* t, uint64_t* keys) noexcept
{
    if(t == nullptr || keys == nullptr)
        return KACHE_EINVAL;
    return guard(t, [&](){ std::lock_guard<std::mutex> lock(t->mu); *keys = t->size; return static_cast<int>(KACHE_OK); });
}

int kache_reduce(const kache_table* t, uint32_t what, uint64_t mask, uint32_t workers, uint64_t* out) noexcept
{
    if(t == nullptr || out == nullptr || workers < 1 || what > KACHE_REDUCE_SUM || (t->value_bytes == 0 && what != KACHE_REDUCE_KEYS))
        return KACHE_EINVAL;
    return guard(t, [&](){ *out = t->reduce(what, mask & t->vmax()); return static_cast<int>(KACHE_OK); });
}

int kache_for_each(const kache_table* t, uint32_t shards, uint32_t shard, kache_keys_fn f, void* ctx) noexcept
{
    if(t == nullptr || f == nullptr || shard >= shards)
        return KACHE_EINVAL;
    return guard(t, [&](){ t->for_each(shards, shard, f, ctx); return static_cast<int>(KACHE_OK); });
}

int kache_get_stats(const kache_table* t, kache_stats* out) noexcept
{
    if(t == nullptr || out == nullptr)
        return KACHE_EINVAL;
    return guard(t, [&](){ t->stats(*out); return static_cast<int>(KACHE_OK); });
}

const char* kache_strerror(int code) noexcept
{
    switch(code)
    {
    case KACHE_OK:          return "success";
    case KACHE_EINVAL:      return "invalid argument";
    case KACHE_EUSERS:      return "too many users registered";
    case KACHE_ENOMEM:      return "out of memory";
    case KACHE_EINTERNAL:   return "internal error";
    default:                return "unknown status";
    }
}

// ---------------------------------------------------------------------------
// Extensions (include/kache_amd.h).

int kache_amd_insert_device(kache_table* t, uint32_t user, const char* text, uint64_t text_bytes, uint32_t op, uint64_t x, uint64_t* inserted) noexcept
{
    if(t == nullptr || !registered(t, user) || !device_text_ok(text, text_bytes) ||
       (t->map && ((op > KACHE_MAX && op != KACHE_AMD_ADD_WRAP) || !fits(t, x))))
        return KACHE_EINVAL;
    return guard(t, [&]()
    {
        const uint64_t c = text_bytes ? t->insert_device(user, reinterpret_cast<const uint8_t*>(text), text_bytes, t->map ? op : 0, false, t->map ? x : 0, 0) : 0;
        if(inserted != nullptr)
            *inserted = c;
        return static_cast<int>(KACHE_OK);
    });
}

int kache_amd_insert_oriented_device(kache_table* t, uint32_t user, const char* text, uint64_t text_bytes, uint32_t op, uint64_t x_fwd, uint64_t x_rev, uint64_t* inserted) noexcept
{
    const bool op_ok = (op == KACHE_ASSIGN || op == KACHE_ADD || op == KACHE_OR || op == KACHE_MAX);
    if(t == nullptr || !registered(t, user) || !device_text_ok(text, text_bytes) || !op_ok || !fits(t, x_fwd) || !fits(t, x_rev))
        return KACHE_EINVAL;
    return guard(t, [&]()
    {
        const uint64_t c = text_bytes ? t->insert_device(user, reinterpret_cast<const uint8_t*>(text), text_bytes, op, true, x_fwd, x_rev) : 0;
        if(inserted != nullptr)
            *inserted = c;
        return static_cast<int>(KACHE_OK);
    });
}

int kache_amd_insert_batches(kache_table* t, uint32_t user, const char* text, const uint64_t* bounds, uint64_t n_batches, uint32_t op,
                             uint64_t x, uint64_t* inserted) noexcept
{
    if(t == nullptr || !registered(t, user) || (n_batches > 0 && bounds == nullptr) || !device_text_ok(text, n_batches ? bounds[n_batches] : 0) ||
       (t->map && ((op > KACHE_MAX && op != KACHE_AMD_ADD_WRAP) || !fits(t, x))))
        return KACHE_EINVAL;
    for(uint64_t i = 0; i < n_batches; ++i)
        if(bounds[i] > bounds[i + 1])
            return KACHE_EINVAL;
    return guard(t, [&]()
    {
        t->insert_batches(user, reinterpret_cast<const uint8_t*>(text), bounds, n_batches, t->map ? op : 0, false, t->map ? x : 0, 0, inserted);
        return static_cast<int>(KACHE_OK);
    });
}

int kache_amd_count_batches(const kache_table* t, const char* text, const uint64_t* bounds, uint64_t n_batches, uint64_t* hits, uint64_t* positions) noexcept
{
    if(t == nullptr || hits == nullptr || (n_batches > 0 && bounds == nullptr) || !device_text_ok(text, n_batches ? bounds[n_batches] : 0))
        return KACHE_EINVAL;
    for(uint64_t i = 0; i < n_batches; ++i)
        if(bounds[i] > bounds[i + 1])
            return KACHE_EINVAL;
    return guard(t, [&]()
    {
        if(n_batches)
            t->count_batches(reinterpret_cast<const uint8_t*>(text), bounds, n_batches, hits, positions);
        return static_cast<int>(KACHE_OK);
    });
}

int kache_amd_update_device(kache_table* t, const char* text, uint64_t text_bytes, uint32_t op, uint64_t x, uint64_t* present) noexcept
{
    if(t == nullptr || !device_text_ok(text, text_bytes) || (op > KACHE_MAX && op != KACHE_AMD_ADD_WRAP) || !fits(t, x))
        return KACHE_EINVAL;
    return guard(t, [&]()
    {
        uint64_t c = 0;
        if(text_bytes)
        {
            std::lock_guard<std::mutex> lock(t->mu);
            c = t->update_text(op, reinterpret_cast<const uint8_t*>(text), text_bytes, x);
        }
        if(present != nullptr)
            *present = c;
        return static_cast<int>(KACHE_OK);
    });
}

int kache_amd_count_device(const kache_table* t, const char* text, uint64_t text_bytes, uint64_t* hits, uint64_t* positions) noexcept
{
    if(t == nullptr || hits == nullptr || !device_text_ok(text, text_bytes))
        return KACHE_EINVAL;
    return guard(t, [&]()
    {
        Counters c{};
        if(text_bytes)
            c = t->count_device(reinterpret_cast<const uint8_t*>(text), text_bytes);
        *hits = c.present;
        if(positions != nullptr)
            *positions = c.positions;
        return static_cast<int>(KACHE_OK);
    });
}

int kache_amd_describe_device(const char* text, uint64_t text_bytes, uint64_t* key, uint64_t* h, uint8_t* c, uint8_t* fwd, uint8_t* valid) noexcept
{
    if(!device_text_ok(text, text_bytes) || (text_bytes > 0 && (!key || !h || !c || !fwd || !valid)))
        return KACHE_EINVAL;
    return guard([&]()
    {
        if(text_bytes == 0)
            return static_cast<int>(KACHE_OK);
        const uint64_t tiles = (text_bytes + TILE - 1) / TILE;
        const uint32_t g = static_cast<uint32_t>(std::min<uint64_t>(tiles, 4096));
        hipLaunchKernelGGL(k_describe, dim3(g), dim3(TILE), 0, 0, reinterpret_cast<const uint8_t*>(text), text_bytes, 0, text_bytes, key, h, c, fwd, valid);
        HIP(hipGetLastError());
        HIP(hipDeviceSynchronize());
        return static_cast<int>(KACHE_OK);
    });
}

int kache_amd_hashes_device(const uint64_t* x, uint64_t n, uint64_t cap, uint64_t* out) noexcept
{
    if(n > 0 && (!x || !out))
        return KACHE_EINVAL;
    if(cap == 0 || cap > MAX_CAP || n > (uint64_t(1) << 32))
        return KACHE_EINVAL;
    return guard([&]()
    {
        if(n == 0)
            return static_cast<int>(KACHE_OK);
        const uint32_t g = static_cast<uint32_t>(std::min<uint64_t>((n + TILE - 1) / TILE, 4096));
        hipLaunchKernelGGL(k_hashes, dim3(g), dim3(TILE), 0, 0, x, n, cap, out);
        HIP(hipGetLastError());
        HIP(hipDeviceSynchronize());
        return static_cast<int>(KACHE_OK);
    });
}

int kache_amd_profile(kache_table* t, int on) noexcept
{
    if(t == nullptr)
        return KACHE_EINVAL;
    std::lock_guard<std::mutex> lock(t->mu);
    t->profiling = (on != 0);
    if(on)
        t->prof.clear();
    return KACHE_OK;
}

int kache_amd_kernel_time(const kache_table* t, const char* name, double* ms, uint64_t* launches) noexcept
{
    if(t == nullptr || name == nullptr)
        return KACHE_EINVAL;
    std::lock_guard<std::mutex> lock(t->mu);
    const auto it = t->prof.find(name);
    if(ms)
        *ms = it == t->prof.end() ? 0 : it->second.first;
    if(launches)
        *launches = it == t->prof.end() ? 0 : it->second.second;
    return KACHE_OK;
}

int kache_amd_set_blocks_per_cu(kache_table* t, uint32_t blocks) noexcept
{
    if(t == nullptr || blocks > 64)
        return KACHE_EINVAL;
    std::lock_guard<std::mutex> lock(t->mu);
    t->blocks_per_cu = blocks;
    return KACHE_OK;
}

int kache_amd_trim_cache(uint64_t* released) noexcept
{
    return guard([&]()
    {
        const uint64_t n = ArrayCache::get().trim(current_device());
        if(released)
            *released = n;
        return static_cast<int>(KACHE_OK);
    });
}

int kache_amd_set_hbm_budget(kache_table* t, uint64_t bytes) noexcept
{
    if(t == nullptr)
        return KACHE_EINVAL;
    std::lock_guard<std::mutex> lock(t->mu);
    t->hbm_budget = bytes;
    // A store taken from kept memory may hold more than its buckets (room
    // to grow into, already mapped); under a budget that room goes back.
    if(bytes != 0 && t->stream)
    {
        (void)hipStreamSynchronize(t->stream);
        VmSpace::get().shrink(t->buckets, t->cap * t->bucket_bytes());
    }
    return KACHE_OK;
}

int kache_amd_set_tuning(kache_table* t, const kache_amd_tuning* cfg) noexcept
{
    if(t == nullptr || cfg == nullptr || cfg->blocks_per_cu > 64 || cfg->launch_log2 < 16 || cfg->launch_log2 > 30 || cfg->graph > 1 ||
       !(cfg->waves_per_simd == 0 || (cfg->waves_per_simd >= 6 && cfg->waves_per_simd <= 8)) ||
       !(cfg->threads_per_block == 0 || cfg->threads_per_block == 64 || cfg->threads_per_block == 128 || cfg->threads_per_block == 256))
        return KACHE_EINVAL;
    std::lock_guard<std::mutex> lock(t->mu);
    t->blocks_per_cu = cfg->blocks_per_cu, t->launch_log2 = cfg->launch_log2, t->use_graph = cfg->graph != 0;
    t->insert_waves = cfg->waves_per_simd ? cfg->waves_per_simd : KAMD_INSERT_WAVES;
    t->insert_threads = cfg->threads_per_block ? cfg->threads_per_block : TILE;
    t->tune_rounds = 0;
    return KACHE_OK;
}

int kache_amd_get_tuning(const kache_table* t, kache_amd_tuning* cfg) noexcept
{
    if(t == nullptr || cfg == nullptr)
        return KACHE_EINVAL;
    std::lock_guard<std::mutex> lock(t->mu);
    *cfg = kache_amd_tuning{t->blocks_per_cu, t->launch_log2, t->use_graph ? 1u : 0u, t->tune_rounds, t->insert_waves, t->insert_threads};
    return KACHE_OK;
}

int kache_amd_autotune(kache_table* t, uint32_t rounds) noexcept
{
    if(t == nullptr || rounds > 16)
        return KACHE_EINVAL;
    std::lock_guard<std::mutex> lock(t->mu);
    return guard([&](){ t->start_tuning(rounds); return static_cast<int>(KACHE_OK); });
}

int kache_amd_tuning_trials(const kache_table* t, kache_amd_tuning_trial* out, size_t cap, size_t* n) noexcept
{
    if(t == nullptr || n == nullptr || (cap > 0 && out == nullptr))
        return KACHE_EINVAL;
    std::lock_guard<std::mutex> lock(t->mu);
    *n = t->trials.size();
    for(size_t i = 0; i < std::min(cap, t->trials.size()); ++i)
    {
        std::vector<double> v = t->trials[i].ns_per_pos, r = t->trials[i].rel;
        std::sort(v.begin(), v.end());
        std::sort(r.begin(), r.end());
        out[i] = kache_amd_tuning_trial{t->trials[i].cfg, static_cast<uint32_t>(v.size()), v.empty() ? 0.0 : v[v.size() / 2],
                                        r.empty() ? 0.0 : r[r.size() / 2]};
    }
    return KACHE_OK;
}

int kache_amd_set_growth(kache_table* t, uint32_t mode, uint64_t away_entries) noexcept
{
    if(t == nullptr || mode > KACHE_AMD_GROW_OUT_OF_PLACE)
        return KACHE_EINVAL;
    std::lock_guard<std::mutex> lock(t->mu);
    t->growth_mode = mode;
    t->away_min = away_entries;
    return KACHE_OK;
}

int kache_amd_get_info(const kache_table* t, kache_amd_info* out) noexcept
{
    if(t == nullptr || out == nullptr)
        return KACHE_EINVAL;
    std::lock_guard<std::mutex> lock(t->mu);
    out->buckets = t->cap;
    out->bucket_bytes = t->bucket_bytes();
    out->overflow_slots = t->ov_slots;
    out->overflow_keys = t->ov_count;
    out->resizes = t->resizes;
    out->device_bytes = t->device_bytes();
    out->value_bytes = t->value_bytes;
    out->next_doubling_bytes = t->next_doubling_bytes();
    size_t free_b = 0, total_b = 0;
    out->device_free_bytes = (hipMemGetInfo(&free_b, &total_b) == hipSuccess ? free_b : 0) + ArrayCache::get().bytes_kept(t->device);
    out->hbm_budget = t->hbm_budget;
    out->store_mapped_bytes = VmSpace::get().mapped(t->buckets);
    if(out->store_mapped_bytes == 0)
        out->store_mapped_bytes = t->cap * t->bucket_bytes();
    out->store_reserved_bytes = VmSpace::get().reserved(t->buckets);
    out->in_place_doublings = t->in_place_doublings;
    out->away_list_regrows = t->away_regrows;
    out->last_away_keys = t->last_away;
    out->growth_mode = t->growth_mode;
    out->reserved0 = 0;
    return KACHE_OK;
}

const char* kache_amd_last_error(void) noexcept
{
    return last_error.c_str();
}

uint64_t kache_amd_host_minimizer_hash(uint64_t key) noexcept { return minimizer_hash(key); }
uint8_t kache_amd_host_checksum(uint64_t key) noexcept { return checksum(key); }
uint64_t kache_amd_host_xxh3(uint64_t x, uint64_t seed) noexcept { return xxh3_8(x, seed); }
uint64_t kache_amd_host_cap_for(uint64_t n, double lf, uint32_t pow2) noexcept { return cap_for(n, lf, pow2 != 0); }

int kache_amd_host_fits(uint64_t n, double lf, uint32_t pow2) noexcept
{
    return lf > 0 && lf < 1 && size_fits(n, lf, pow2 != 0);
}

int kache_amd_host_describe(const char* text, uint64_t text_bytes, uint64_t* key, uint64_t* h, uint8_t* c, uint8_t* fwd, uint8_t* valid) noexcept
{
    if(text_bytes > 0 && (!text || !key || !h || !c || !fwd || !valid))
        return KACHE_EINVAL;
    // The device describer's steps, on the host: 32-base words packed 8
    // characters at a time with pack8, 32-base windows from two words, l-mer
    // hashes, window minima.
    const uint64_t words = (text_bytes + TILE_WORDS * 32) / 32 + 2;
    std::vector<uint64_t> pk(words, 0);
    std::vector<uint32_t> bad(words, ~0u);
    for(uint64_t w = 0; w < words; ++w)
    {
        pk[w] = 0, bad[w] = 0;
        for(int q = 0; q < 4; ++q)
        {
            uint64_t v = 0;
            for(int b = 0; b < 8; ++b)
            {
                const uint64_t at = w * 32 + 8 * q + b;
                const uint64_t ch = at < text_bytes ? static_cast<uint8_t>(text[at]) : '\n';
                v |= ch << (8 * b);
            }
            uint32_t code, nb;
            pack8(v, code, nb);
            pk[w] |= static_cast<uint64_t>(code) << (16 * (3 - q));
            bad[w] |= nb << (8 * q);
        }
    }
    const auto window = [&](uint64_t u)
    {
        const uint64_t wi = u >> 5;
        const int sh = static_cast<int>(u & 31) * 2;
        return sh ? (pk[wi] << sh) | (pk[wi + 1] >> (64 - sh)) : pk[wi];
    };
    std::vector<uint64_t> lh(text_bytes + W);
    for(uint64_t p = 0; p < text_bytes + W - 1; ++p)
    {
        uint64_t kf, kr;
        window_parts(window(p), kf, kr, lh[p]);
    }
    for(uint64_t p = 0; p < text_bytes; ++p)
    {
        uint64_t kf, kr, l0;
        window_parts(window(p), kf, kr, l0);
        const uint64_t wi = p >> 5;
        const uint64_t b64 = static_cast<uint64_t>(bad[wi]) | (static_cast<uint64_t>(bad[wi + 1]) << 32);
        const bool ok = ((b64 >> (p & 31)) & 0x7FFFFFFFull) == 0 && p + K <= text_bytes;
        uint64_t m = lh[p];
        for(int o = 1; o < W; ++o)
            m = std::min(m, lh[p + o]);
        valid[p] = ok;
        key[p] = ok ? std::min(kf, kr) : 0;
        h[p] = ok ? m : 0;
        c[p] = ok ? checksum(std::min(kf, kr)) : 0;
        fwd[p] = ok ? (kf <= kr) : 0;
    }
    return KACHE_OK;
}

}

// Copies n spans (each followed by a '\n' at dst[i] + len[i]), split over
// the packing threads by bytes, of the second pool if `aside` (other sources
// of the library).
__attribute__((visibility("hidden"))) void kamd_copy_spans(char* const* dst, const char* const* src, const uint64_t* len, size_t n, bool aside)
{
    CopyPool& pool = aside ? CopyPool::aside() : CopyPool::get();
    uint64_t total = 0;
    for(size_t i = 0; i < n; ++i)
        total += len[i] + 1;
    const auto copy = [&](size_t a, size_t b)
    {
        for(size_t i = a; i < b; ++i)
        {
            if(len[i])
                std::memcpy(dst[i], src[i], len[i]);
            dst[i][len[i]] = '\n';
        }
    };
    if(total < (uint64_t(4) << 20) || n < 2)
    {
        if(n == 1 && total >= (uint64_t(4) << 20))
        {
            par_copy(dst[0], src[0], len[0], pool);
            dst[0][len[0]] = '\n';
        }
        else
            copy(0, n);
        return;
    }
    const size_t parts = pool.threads;
    std::vector<size_t> cut(parts + 1, n);
    cut[0] = 0;
    uint64_t acc = 0;
    for(size_t p = 1, i = 0; p < parts; ++p)
    {
        const uint64_t target = total * p / parts;
        while(i < n && acc < target)
            acc += len[i] + 1, ++i;
        cut[p] = i;
    }
    pool.run(parts, [&](size_t p){ copy(cut[p], std::max(cut[p], cut[p + 1])); });
}

// Sets the calling thread's last error (other sources of the library).
__attribute__((visibility("hidden"))) void kamd_set_last_error(const std::string& what)
{
    last_error = what;
}
