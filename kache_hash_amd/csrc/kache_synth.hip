// Device memory helpers and the seeded synthetic workload generator of
// include/kache_amd.h: genomes and reads for tests, benchmarks and the
// serving loop, made on the device so large workloads never cross PCIe.
//
// Every value comes from a counter-based hash (the splitmix64 finalizer of
// seed + (i + 1) * golden gamma), so the host reproduces any read exactly
// (kache_amd_host_synth_reads runs the same code).

#include "kache_amd.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>

namespace
{

__host__ __device__ __forceinline__ uint64_t mix(uint64_t seed, uint64_t i)
{
    uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__host__ __device__ __forceinline__ uint8_t base_char(uint32_t code)
{
    return static_cast<uint8_t>(code == 0 ? 'A' : code == 1 ? 'C' : code == 2 ? 'G' : 'T');
}

__host__ __device__ __forceinline__ uint32_t genome_code(uint64_t seed, uint64_t i)
{
    return static_cast<uint32_t>((mix(seed, i >> 5) >> (2 * (i & 31))) & 3);
}

// Parameters of a read set.
struct ReadParams
{
    uint64_t genome_seed, genome_len, seed, n, len;
    uint32_t sub_th, n_th, lower_th;    // thresholds out of 2^24
};

// Character j (0 <= j <= len) of read r: the genome window at a uniform
// start, substitutions to another base, reverse complement for half the
// reads, N at rate n_th, a lowercase run in some reads; '\n' ends the read.
__host__ __device__ __forceinline__ uint8_t read_char(const ReadParams& p, uint64_t r, uint64_t j)
{
    if(j == p.len)
        return '\n';
    const uint64_t h0 = mix(p.seed, 4 * r), h1 = mix(p.seed, 4 * r + 1), h2 = mix(p.seed, 4 * r + 2);
    const uint64_t start = h0 % (p.genome_len - p.len + 1);
    const bool rev = h1 & 1;
    const uint64_t w = rev ? p.len - 1 - j : j;      // window offset of read position j
    uint32_t code = genome_code(p.genome_seed, start + w);
    const uint64_t u = mix(p.seed ^ 0x5DEECE66Dull, r * p.len + w);
    if((u & 0xFFFFFF) < p.sub_th)
        code = (code + 1 + static_cast<uint32_t>((u >> 24) % 3)) & 3;
    if(rev)
        code = 3 - code;
    uint8_t ch = base_char(code);
    const uint64_t v = mix(p.seed ^ 0xB5AD4ECEDA1CE2A9ull, r * p.len + j);
    if((v & 0xFFFFFF) < p.n_th)
        ch = 'N';
    if(((h2 & 0xFFFFFF) < p.lower_th))
    {
        const uint64_t lo = (h2 >> 24) % p.len, span = 1 + (h2 >> 40) % p.len;
        if(j >= lo && j < lo + span)
            ch |= 0x20;
    }
    return ch;
}

// Character j (0 <= j < len) of sequence i of length len at text offset
// `base`: as read_char, with per-sequence lengths; the position hashes use
// the text offset, unique to each character.
__host__ __device__ __forceinline__ uint8_t seq_char(const ReadParams& p, uint64_t i, uint64_t len, uint64_t base, uint64_t j)
{
    const uint64_t h0 = mix(p.seed, 4 * i), h1 = mix(p.seed, 4 * i + 1), h2 = mix(p.seed, 4 * i + 2);
    const uint64_t start = h0 % (p.genome_len - len + 1);
    const bool rev = h1 & 1;
    const uint64_t w = rev ? len - 1 - j : j;
    uint32_t code = genome_code(p.genome_seed, start + w);
    const uint64_t u = mix(p.seed ^ 0x5DEECE66Dull, base + w);
    if((u & 0xFFFFFF) < p.sub_th)
        code = (code + 1 + static_cast<uint32_t>((u >> 24) % 3)) & 3;
    if(rev)
        code = 3 - code;
    uint8_t ch = base_char(code);
    const uint64_t v = mix(p.seed ^ 0xB5AD4ECEDA1CE2A9ull, base + j);
    if((v & 0xFFFFFF) < p.n_th)
        ch = 'N';
    if((h2 & 0xFFFFFF) < p.lower_th)
    {
        const uint64_t lo = (h2 >> 24) % len, span = 1 + (h2 >> 40) % len;
        if(j >= lo && j < lo + span)
            ch |= 0x20;
    }
    return ch;
}

// One block per sequence (grid-stride): sequence i fills [off[i], off[i+1] - 1)
// and '\n' ends it.
__global__ void k_synth_seqs(uint8_t* out, const uint64_t* off, ReadParams p)
{
    for(uint64_t i = blockIdx.x; i < p.n; i += gridDim.x)
    {
        const uint64_t b = off[i], len = off[i + 1] - b - 1;
        for(uint64_t j = threadIdx.x; j <= len; j += blockDim.x)
            out[b + j] = j == len ? uint8_t('\n') : seq_char(p, i, len, b, j);
    }
}

__global__ void k_synth_genome(uint8_t* out, uint64_t n, uint64_t seed)
{
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for(uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += stride)
        out[i] = base_char(genome_code(seed, i));
}

__global__ void k_synth_reads(uint8_t* out, ReadParams p)
{
    const uint64_t total = p.n * (p.len + 1), stride = uint64_t(gridDim.x) * blockDim.x;
    for(uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < total; i += stride)
        out[i] = read_char(p, i / (p.len + 1), i % (p.len + 1));
}

thread_local std::string synth_error;

int fail(hipError_t e)
{
    synth_error = hipGetErrorString(e);
    (void)hipGetLastError();
    return e == hipErrorOutOfMemory ? KACHE_ENOMEM : KACHE_EINTERNAL;
}

ReadParams params(uint64_t genome_seed, uint64_t genome_len, uint64_t n_reads, uint64_t len, uint64_t seed, double sub, double nrate, double lower)
{
    const auto th = [](double x) { return static_cast<uint32_t>(std::clamp(x, 0.0, 1.0) * 16777216.0); };
    return ReadParams{genome_seed, genome_len, seed, n_reads, len, th(sub), th(nrate), th(lower)};
}

}

extern "C"
{

int kache_amd_device_alloc(uint64_t bytes, void** ptr) noexcept
{
    if(ptr == nullptr)
        return KACHE_EINVAL;
    *ptr = nullptr;
    const hipError_t e = hipMalloc(ptr, std::max<uint64_t>(bytes, 1));
    return e == hipSuccess ? KACHE_OK : fail(e);
}

int kache_amd_device_free(void* ptr) noexcept
{
    const hipError_t e = hipFree(ptr);
    return e == hipSuccess ? KACHE_OK : fail(e);
}

int kache_amd_host_alloc(uint64_t bytes, void** ptr) noexcept
{
    if(ptr == nullptr)
        return KACHE_EINVAL;
    *ptr = nullptr;
    const hipError_t e = hipHostMalloc(ptr, std::max<uint64_t>(bytes, 1), hipHostMallocDefault);
    return e == hipSuccess ? KACHE_OK : fail(e);
}

int kache_amd_host_free(void* ptr) noexcept
{
    const hipError_t e = hipHostFree(ptr);
    return e == hipSuccess ? KACHE_OK : fail(e);
}

int kache_amd_copy(void* dst, const void* src, uint64_t bytes, int to_device) noexcept
{
    if(bytes == 0)
        return KACHE_OK;
    if(dst == nullptr || src == nullptr)
        return KACHE_EINVAL;
    const hipError_t e = hipMemcpy(dst, src, bytes, to_device ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost);
    return e == hipSuccess ? KACHE_OK : fail(e);
}

int kache_amd_device_fill(void* ptr, uint8_t value, uint64_t bytes) noexcept
{
    if(bytes == 0)
        return KACHE_OK;
    hipError_t e = hipMemset(ptr, value, bytes);
    if(e == hipSuccess)
        e = hipDeviceSynchronize();
    return e == hipSuccess ? KACHE_OK : fail(e);
}

int kache_amd_device_sync(void) noexcept
{
    const hipError_t e = hipDeviceSynchronize();
    return e == hipSuccess ? KACHE_OK : fail(e);
}

int kache_amd_synth_genome(char* out, uint64_t n, uint64_t seed) noexcept
{
    if(n > 0 && out == nullptr)
        return KACHE_EINVAL;
    if(n == 0)
        return KACHE_OK;
    const uint32_t g = static_cast<uint32_t>(std::min<uint64_t>((n + 255) / 256, 65536));
    hipLaunchKernelGGL(k_synth_genome, dim3(g), dim3(256), 0, 0, reinterpret_cast<uint8_t*>(out), n, seed);
    hipError_t e = hipGetLastError();
    if(e == hipSuccess)
        e = hipDeviceSynchronize();
    return e == hipSuccess ? KACHE_OK : fail(e);
}

int kache_amd_synth_reads(char* out, uint64_t n_reads, uint64_t read_len, uint64_t genome_seed, uint64_t genome_len,
                          uint64_t seed, double sub, double n_rate, double lower) noexcept
{
    if((n_reads > 0 && out == nullptr) || read_len == 0 || genome_len < read_len)
        return KACHE_EINVAL;
    if(n_reads == 0)
        return KACHE_OK;
    const ReadParams p = params(genome_seed, genome_len, n_reads, read_len, seed, sub, n_rate, lower);
    const uint64_t total = n_reads * (read_len + 1);
    const uint32_t g = static_cast<uint32_t>(std::min<uint64_t>((total + 255) / 256, 65536));
    hipLaunchKernelGGL(k_synth_reads, dim3(g), dim3(256), 0, 0, reinterpret_cast<uint8_t*>(out), p);
    hipError_t e = hipGetLastError();
    if(e == hipSuccess)
        e = hipDeviceSynchronize();
    return e == hipSuccess ? KACHE_OK : fail(e);
}

int kache_amd_synth_seqs(char* out, const uint64_t* offsets, uint64_t n, uint64_t max_len, uint64_t genome_seed, uint64_t genome_len,
                         uint64_t seed, double sub, double n_rate, double lower) noexcept
{
    if((n > 0 && (out == nullptr || offsets == nullptr)) || genome_len < max_len || max_len == 0)
        return KACHE_EINVAL;
    if(n == 0)
        return KACHE_OK;
    const ReadParams p = params(genome_seed, genome_len, n, 0, seed, sub, n_rate, lower);
    const uint32_t g = static_cast<uint32_t>(std::min<uint64_t>(n, 65536));
    hipLaunchKernelGGL(k_synth_seqs, dim3(g), dim3(256), 0, 0, reinterpret_cast<uint8_t*>(out), offsets, p);
    hipError_t e = hipGetLastError();
    if(e == hipSuccess)
        e = hipDeviceSynchronize();
    return e == hipSuccess ? KACHE_OK : fail(e);
}

int kache_amd_host_synth_reads(char* out, uint64_t first_read, uint64_t n_reads, uint64_t read_len, uint64_t genome_seed,
                               uint64_t genome_len, uint64_t seed, double sub, double n_rate, double lower) noexcept
{
    if((n_reads > 0 && out == nullptr) || read_len == 0 || genome_len < read_len)
        return KACHE_EINVAL;
    const ReadParams p = params(genome_seed, genome_len, first_read + n_reads, read_len, seed, sub, n_rate, lower);
    for(uint64_t r = 0; r < n_reads; ++r)
        for(uint64_t j = 0; j <= read_len; ++j)
            out[r * (read_len + 1) + j] = static_cast<char>(read_char(p, first_read + r, j));
    return KACHE_OK;
}

const char* kache_amd_synth_error(void) noexcept
{
    return synth_error.c_str();
}

}
