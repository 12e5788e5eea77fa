// Memory roofs of one MI355X for the k-mer table's roofline, after the
// reference's tools/memory-roof.cpp (phases :222-354): over an array as large
// as the tables it bounds,
//   stream          sequential 16-byte reads of every line (GB/s)
//   copy, write     a sequential copy of half the array onto the other half
//                   (bytes read + written), sequential 16-byte writes (GB/s)
//   random_read     independent reads of one word of random lines (lines/s)
//   random_update   independent read-modify-writes of one word of random lines
//   random_atomic   independent 64-bit atomic adds to random lines, no return
//   random_atomic_ret  the same, returning the old value (as CAS / fetch-add)
//   chase           dependent random reads: latency in ns
// each at 64-byte and 128-byte line granules.  Random line indices come from
// a counter-based hash, so every access is independent of the others, and
// each lane keeps DEPTH of them in flight (the reference prefetches 32 ahead).
// Prints one JSON object.
//
// Usage: memory_roof [--gib N] [--iters N]

#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#define CHECK(x) do { hipError_t e_ = (x); if(e_ != hipSuccess) { std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_)); std::exit(1); } } while(0)

namespace
{

constexpr int DEPTH = 8;

__device__ __forceinline__ uint64_t mix(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ void k_init(uint64_t* a, uint64_t words, uint64_t line_words, uint64_t lines)
{
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for(uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < words; i += stride)
    {
        // The first word of each line holds a random next line for the chase.
        a[i] = (i % line_words == 0) ? mix(i + 12345) & (lines - 1) : 0;
    }
}

__global__ void k_stream(const uint4* a, uint64_t n16, uint64_t* sink)
{
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    uint32_t s = 0;
    for(uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n16; i += stride)
    {
        const uint4 v = a[i];
        s ^= v.x ^ v.y ^ v.z ^ v.w;
    }
    if(s == 0x12345678u)
        atomicAdd(reinterpret_cast<unsigned long long*>(sink), 1ull);
}

// Copies n16 16-byte words from src to dst (streams both ways).
__global__ void k_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, uint64_t n16)
{
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for(uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n16; i += stride)
        dst[i] = src[i];
}

// Writes n16 16-byte words.
__global__ void k_write(uint4* a, uint64_t n16)
{
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for(uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n16; i += stride)
        a[i] = make_uint4(uint32_t(i), 0u, 0u, 0u);
}

// MODE 0 read, 1 read-modify-write, 2 atomic add, 3 atomic add with return,
// 4 atomic add by groups of 4 lanes on consecutive 16-byte slots of one line
// (the pattern of a minimizer run's counters).
template <int MODE>
__global__ void k_random(uint64_t* a, uint64_t lines, uint32_t line_words, uint32_t touch, uint64_t iters, uint64_t seed, uint64_t* sink)
{
    const uint64_t gtid = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
    const uint64_t tid = MODE == 4 ? gtid / 4 : gtid;
    const uint32_t sub = MODE == 4 ? static_cast<uint32_t>(gtid % 4) * 2 : 0;
    uint64_t acc = 0;
    for(uint64_t it = 0; it < iters; it += DEPTH)
    {
        uint64_t* p[DEPTH];
#pragma unroll
        for(int d = 0; d < DEPTH; ++d)
            p[d] = a + (mix(seed + tid * 0x9E3779B97F4A7C15ull + (it + d) * 0xD1B54A32D192ED03ull) & (lines - 1)) * line_words;
        if constexpr(MODE == 0)
        {
            uint64_t v[DEPTH];
#pragma unroll
            for(int d = 0; d < DEPTH; ++d)
            {
                v[d] = p[d][0];
                if(touch > 1)
                    v[d] += p[d][8];
            }
#pragma unroll
            for(int d = 0; d < DEPTH; ++d)
                acc += v[d];
        }
        else if constexpr(MODE == 1)
        {
            uint64_t v[DEPTH];
#pragma unroll
            for(int d = 0; d < DEPTH; ++d)
                v[d] = p[d][1];
#pragma unroll
            for(int d = 0; d < DEPTH; ++d)
                p[d][1] = v[d] + 1;
        }
        else if constexpr(MODE == 2 || MODE == 4)
        {
#pragma unroll
            for(int d = 0; d < DEPTH; ++d)
                __hip_atomic_fetch_add(p[d] + 1 + sub, uint64_t(1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        else
        {
            uint64_t v[DEPTH];
#pragma unroll
            for(int d = 0; d < DEPTH; ++d)
                v[d] = __hip_atomic_fetch_add(p[d] + 1, uint64_t(1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
            for(int d = 0; d < DEPTH; ++d)
                acc += v[d];
        }
    }
    if(acc == 0x123456789ull)
        atomicAdd(reinterpret_cast<unsigned long long*>(sink), 1ull);
}

__global__ void k_chase(const uint64_t* a, uint64_t lines, uint32_t line_words, uint64_t steps, uint64_t* sink)
{
    const uint64_t tid = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
    uint64_t i = mix(tid + 777) & (lines - 1);
    for(uint64_t s = 0; s < steps; ++s)
        i = a[i * line_words];
    if(i == 0xFFFFFFFFFFFFull)
        atomicAdd(reinterpret_cast<unsigned long long*>(sink), 1ull);
}

template <typename F>
double timed(F f)
{
    hipEvent_t a, b;
    CHECK(hipEventCreate(&a));
    CHECK(hipEventCreate(&b));
    CHECK(hipEventRecord(a));
    f();
    CHECK(hipGetLastError());
    CHECK(hipEventRecord(b));
    CHECK(hipEventSynchronize(b));
    float ms = 0;
    CHECK(hipEventElapsedTime(&ms, a, b));
    return ms * 1e-3;
}

}

int main(int argc, char** argv)
{
    uint64_t gib = 32, iters = 512;
    for(int i = 1; i + 1 < argc; i += 2)
    {
        if(!std::strcmp(argv[i], "--gib"))
            gib = std::strtoull(argv[i + 1], nullptr, 10);
        else if(!std::strcmp(argv[i], "--iters"))
            iters = std::strtoull(argv[i + 1], nullptr, 10);
    }
    iters = (iters + DEPTH - 1) / DEPTH * DEPTH;
    const uint64_t bytes = gib << 30;
    int dev = 0, cus = 0;
    CHECK(hipGetDevice(&dev));
    CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    uint64_t* a = nullptr;
    uint64_t* sink = nullptr;
    CHECK(hipMalloc(&a, bytes));
    CHECK(hipMalloc(&sink, 8));
    const uint32_t block = 256;
    const uint32_t grid = static_cast<uint32_t>(cus) * 8;   // resident blocks at 8 waves per SIMD
    const uint64_t lanes = uint64_t(grid) * block;

    std::printf("{\"tool\": \"memory_roof\", \"array_bytes\": %llu, \"cus\": %d, \"lanes\": %llu, \"depth\": %d, \"iters_per_lane\": %llu",
                (unsigned long long)bytes, cus, (unsigned long long)lanes, DEPTH, (unsigned long long)iters);

    for(const uint32_t line : {64u, 128u})
    {
        const uint64_t lines = bytes / line, words = bytes / 8, lw = line / 8;
        hipLaunchKernelGGL(k_init, dim3(grid), dim3(block), 0, 0, a, words, lw, lines);
        CHECK(hipDeviceSynchronize());
        if(line == 64)
        {
            timed([&]{ hipLaunchKernelGGL(k_stream, dim3(grid), dim3(block), 0, 0, reinterpret_cast<const uint4*>(a), bytes / 16, sink); });
            const double t = timed([&]{ hipLaunchKernelGGL(k_stream, dim3(grid), dim3(block), 0, 0, reinterpret_cast<const uint4*>(a), bytes / 16, sink); });
            std::printf(", \"stream_GBps\": %.1f", bytes / t / 1e9);
            // Copy of one half onto the other (bytes read + bytes written), then a
            // stream of writes over the whole array; the array is set up again after.
            const uint64_t half16 = bytes / 32;
            const auto copy = [&]{ hipLaunchKernelGGL(k_copy, dim3(grid), dim3(block), 0, 0, reinterpret_cast<const uint4*>(a),
                                                      reinterpret_cast<uint4*>(a) + half16, half16); };
            timed(copy);
            const double tcp = timed(copy);
            std::printf(", \"copy_GBps\": %.1f", 2.0 * half16 * 16 / tcp / 1e9);
            const auto write = [&]{ hipLaunchKernelGGL(k_write, dim3(grid), dim3(block), 0, 0, reinterpret_cast<uint4*>(a), bytes / 16); };
            timed(write);
            const double tw = timed(write);
            std::printf(", \"write_GBps\": %.1f", bytes / tw / 1e9);
            hipLaunchKernelGGL(k_init, dim3(grid), dim3(block), 0, 0, a, words, lw, lines);
            CHECK(hipDeviceSynchronize());
        }
        const uint32_t touch = line == 128 ? 2 : 1;     // words read per line: both 64-byte halves of a 128-byte line
        const double n = double(lanes) * iters;
        const auto phase = [&](const char* name, auto kern, int mode)
        {
            (void)mode;
            timed([&]{ hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, 0, a, lines, uint32_t(lw), touch, uint64_t(DEPTH * 4), 99ull, sink); });
            const double t = timed([&]{ hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, 0, a, lines, uint32_t(lw), touch, iters, 1ull + line, sink); });
            std::printf(", \"%s_%u_lines_per_s\": %.4g, \"%s_%u_GBps\": %.1f", name, line, n / t, name, line, n * line / t / 1e9);
        };
        phase("random_read", k_random<0>, 0);
        phase("random_update", k_random<1>, 1);
        phase("random_atomic", k_random<2>, 2);
        phase("random_atomic_ret", k_random<3>, 3);
        if(line == 64)
            phase("run4_atomic", k_random<4>, 4);   // counts operations: 4 per line
        const uint64_t steps = 256;
        const double tc = timed([&]{ hipLaunchKernelGGL(k_chase, dim3(grid), dim3(block), 0, 0, a, lines, uint32_t(lw), steps, sink); });
        std::printf(", \"chase_%u_ns\": %.1f, \"chase_%u_lines_per_s\": %.4g", line, tc / steps * 1e9, line, lanes * double(steps) / tc);
    }
    std::printf("}\n");
    CHECK(hipFree(a));
    CHECK(hipFree(sink));
    return 0;
}
