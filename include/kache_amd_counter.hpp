#ifndef KACHE_AMD_COUNTER_HPP
#define KACHE_AMD_COUNTER_HPP

// The reference's Kmer_Set_Counter (counter/Kmer_Set_Counter.hpp:14-79) on
// the MI355X table: a set of 31-mers fixed by prime(), whose occurrences in
// reads count() counts, with 32-bit counts.  Same operations and results:
//   prime(path)      upsert with v -> v, absent k-mers at 0      (:64-75)
//   count(path)      update v -> v + 1 (wrapping) of present k-mers; returns
//                    the occurrences counted                      (:78-94)
//   histogram(max)   keys with count c at [c], c in [1, max]; [0] = 0 (:97-120)
//   count_of(kmer)   the count, 0 if absent or not a k-mer       (:123-127)
// Inputs are FASTA / FASTQ files read as the reference reads them
// (kache_amd_insert_fastx), or in-memory texts of sequences separated by
// non-ACGT bytes.  `thread_c` only sizes the table's user count, as the
// reference's does; the device does the work.  Operations must not overlap,
// but for histograms and count_ofs, as in the reference.

#include "kache_amd.h"

#include <chrono>
#include <cstddef>
#include <cstdint>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

namespace kache_amd
{

class Kmer_Set_Counter
{
public:
    struct Timings
    {
        double alloc = 0;
        double prime = 0;
        double count = 0;
        double histogram = 0;
    };

    static bool supports(uint16_t k) { return k == 31; }

    // Room for n k-mers at load factor lf, the bucket count rounded up to a
    // power of 2 (counter/Kmer_Set_Counter.cpp:29-38).
    Kmer_Set_Counter(uint16_t k, std::size_t n, std::size_t thread_c, double lf = 0.9)
    {
        if(!supports(k))
            throw std::invalid_argument("k-mers of length " + std::to_string(k) + " are not supported, only of 31");
        if(thread_c < 1 || thread_c > KACHE_MAX_USERS)
            throw std::invalid_argument("the thread count must be in [1, " + std::to_string(KACHE_MAX_USERS) + "]");
        const auto t0 = now();
        kache_options o = kache_default_options(31, 4);
        o.expected_keys = n, o.load_factor = lf, o.resize_workers = static_cast<uint32_t>(thread_c), o.pow2_capacity = 1;
        check(kache_new(&o, &t), "kache_new");
        check(kache_register(t, &user), "kache_register");
        timings_.alloc = since(t0);
    }

    Kmer_Set_Counter(const Kmer_Set_Counter&) = delete;
    Kmer_Set_Counter& operator=(const Kmer_Set_Counter&) = delete;
    ~Kmer_Set_Counter() { kache_free(t); }

    void prime(const std::string& path)
    {
        const auto t0 = now();
        check(kache_amd_insert_fastx(t, user, path.c_str(), KACHE_KEEP, 0, nullptr), "prime");
        timings_.prime += since(t0);
    }

    // The same over an in-memory text of sequences separated by non-ACGT bytes.
    void prime_text(const char* text, std::size_t bytes)
    {
        const auto t0 = now();
        check(kache_amd_insert_text(t, user, text, bytes, KACHE_KEEP, 0, nullptr), "prime");
        timings_.prime += since(t0);
    }

    std::size_t count(const std::string& path)
    {
        const auto t0 = now();
        uint64_t found = 0;
        check(kache_amd_update_fastx(t, path.c_str(), KACHE_AMD_ADD_WRAP, 1, &found), "count");
        timings_.count += since(t0);
        return found;
    }

    std::size_t count_text(const char* text, std::size_t bytes)
    {
        const auto t0 = now();
        uint64_t found = 0;
        check(kache_amd_update_text(t, text, bytes, KACHE_AMD_ADD_WRAP, 1, &found), "count");
        timings_.count += since(t0);
        return found;
    }

    std::vector<std::size_t> histogram(std::size_t max)
    {
        const auto t0 = now();
        std::vector<uint64_t> h(max + 1);
        check(kache_amd_histogram(t, max, h.data()), "histogram");
        timings_.histogram += since(t0);
        return std::vector<std::size_t>(h.begin(), h.end());
    }

    uint64_t count_of(const char* kmer) const
    {
        uint8_t f = 0;
        uint64_t v = 0;
        check(kache_get(t, kmer, &f, &v), "count_of");
        return f ? v : 0;
    }

    std::size_t size() const
    {
        uint64_t n = 0;
        check(kache_size(t, &n), "size");
        return n;
    }

    const Timings& timings() const { return timings_; }

    // The C ABI handle, for the extensions of kache_amd.h.
    kache_table* handle() const { return t; }

private:
    kache_table* t = nullptr;
    uint32_t user = 0;
    Timings timings_;

    static std::chrono::steady_clock::time_point now() { return std::chrono::steady_clock::now(); }
    static double since(std::chrono::steady_clock::time_point t0) { return std::chrono::duration<double>(now() - t0).count(); }

    static void check(int st, const char* what)
    {
        if(st == KACHE_OK)
            return;
        const std::string msg = std::string(what) + ": " + kache_strerror(st) + " (" + kache_amd_last_error() + ")";
        if(st == KACHE_EINVAL)
            throw std::runtime_error(msg);   // the reference's reader throws runtime_error on bad input
        if(st == KACHE_ENOMEM)
            throw std::bad_alloc();
        throw std::runtime_error(msg);
    }
};

}

#endif
