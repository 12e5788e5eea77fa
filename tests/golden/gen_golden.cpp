// Fixture writer for the describer and hash golden vectors.
//
// Built against the reference headers (include/Kmer_Batch.hpp, include/Kmer_Lanes.hpp,
// include/DNA_Utility.hpp, include/xxHash/xxhash.h) by `make -C oracle golden`, which
// compiles it twice: once for AVX2 (the multiply checksum, include/Kmer_Batch.hpp:49-50)
// and once for AVX-512 (the Wang checksum, include/Kmer_Batch.hpp:47-48).  It writes
// plain little-endian binary records; tests/golden/README.md gives their layout.
//
// Usage: gen_golden <out_dir> <tag>

#include "DNA_Utility.hpp"
#include "Kmer_Batch.hpp"
#include "Kmer_Lanes.hpp"
#include "xxHash/xxhash.h"

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

using namespace kache_hash;

namespace
{

constexpr uint16_t k = 31, l = 22;

struct SplitMix
{
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
};

// Seeded test sequences: uniform ACGT with lowercase runs, non-ACGT characters
// of several kinds, and ragged lengths, including the edge lengths k-1, k, k+1.
std::vector<std::string> sequences()
{
    SplitMix r{0x601DE4ull};
    std::vector<std::string> S;
    const char* const acgt = "ACGT";
    const char bad[] = {'N', 'n', 'R', '-', '.', 'X', '\0', 'U'};
    const std::size_t fixed[] = {0, 1, 30, 31, 32, 40, 41, 50, 63, 64, 65, 100, 150, 151, 257, 1000, 4096, 5000};
    for(const auto len : fixed)
    {
        std::string s(len, 'A');
        for(auto& c : s)
            c = acgt[r.next() & 3];
        S.push_back(s);
    }

    for(int i = 0; i < 60; ++i)
    {
        const std::size_t len = 1 + r.next() % 1200;
        std::string s(len, 'A');
        for(auto& c : s)
            c = acgt[r.next() & 3];
        if(i % 3 == 0)  // A lowercase run.
        {
            const std::size_t a = r.next() % len, b = a + r.next() % (len - a + 1);
            for(std::size_t j = a; j < b; ++j)
                s[j] |= 0x20;
        }
        if(i % 2 == 1)  // Some non-ACGT characters.
            for(std::size_t m = r.next() % 6; m > 0; --m)
                s[r.next() % len] = bad[r.next() % sizeof(bad)];
        if(i % 7 == 3)  // A homopolymer, whose k-mers repeat.
            for(std::size_t j = 0; j < std::min<std::size_t>(len, 200); ++j)
                s[j] = 'A';
        S.push_back(s);
    }

    S.push_back(std::string(40, 'N'));
    S.push_back(std::string(31, 'a'));
    S.push_back("NNNNACGTACGTACGTACGTACGTACGTACGTACGTNNNN");
    S.push_back("ATGCGTACCTTAGGCATCCAGTTAGACGGATCA");
    return S;
}

template <typename T>
void put(std::FILE* f, const T& x) { std::fwrite(&x, sizeof(x), 1, f); }

}


int main(const int argc, const char* const argv[])
{
    if(argc < 3)
    {
        std::fprintf(stderr, "Usage: %s <out_dir> <tag>\n", argv[0]);
        return 2;
    }

    const std::string dir = argv[1], tag = argv[2];

    // 1. Sequences, and the describer's records for each k-mer of each stretch,
    //    split as the table's streams split them (include/Streaming_Kmer_Hash_Table.hpp:1693-1710).
    const auto S = sequences();
    {
        std::FILE* f = std::fopen((dir + "/seqs.bin").c_str(), "wb");
        put<uint64_t>(f, S.size());
        for(const auto& s : S)
            put<uint64_t>(f, s.size());
        for(const auto& s : S)
            std::fwrite(s.data(), 1, s.size(), f);
        std::fclose(f);
    }
    {
        std::FILE* f = std::fopen((dir + "/describe_" + tag + ".bin").c_str(), "wb");
        Kmer_Batch<k, l> kb;
        typename Kmer_Batch<k, l>::Oriented_Chunk ch;
        for(uint32_t q = 0; q < S.size(); ++q)
        {
            const char* const seq = S[q].data();
            const std::size_t len = S[q].size();
            for(std::size_t s = 0; s + k <= len; )
            {
                const auto e = s + DNA_Utility::acgt_prefix(seq + s, len - s);
                if(e - s >= k)
                {
                    kb.reset(seq + s, e - s - (k - 1));
                    std::size_t i = s;
                    while(kb.next(ch))
                        for(std::size_t j = 0; j < ch.n; ++j, ++i)
                        {
                            put<uint32_t>(f, q), put<uint32_t>(f, static_cast<uint32_t>(i));
                            put<uint64_t>(f, ch.key[j]), put<uint64_t>(f, ch.h[j]);
                            put<uint8_t>(f, ch.c[j]), put<uint8_t>(f, ch.o[j]);
                            put<uint16_t>(f, 0), put<uint32_t>(f, 0);
                        }
                }
                s = e + 1;
            }
        }
        std::fclose(f);
    }

    // 2. Scalar functions over seeded and edge inputs: wang_u64, kmer_checksum,
    //    minimizer_hashes, XXH3_64bits_withSeed(&x, 8, {0, 2^63}), XXH3_64bits(&x, 8).
    {
        SplitMix r{0xC0FFEEull};
        std::vector<uint64_t> X{0, 1, 2, 3, 0x1234, (uint64_t(1) << 62) - 1, (uint64_t(1) << 62) - 2, ~uint64_t(0), uint64_t(1) << 63};
        while(X.size() < 4096)
            X.push_back(X.size() % 2 ? r.next() : r.next() & ((uint64_t(1) << 62) - 1));

        std::vector<uint64_t> H(X.size());
        std::vector<uint64_t> C(X.size());
        for(std::size_t i = 0; i < X.size(); ++i)
            C[i] = X[i] & ((uint64_t(1) << 62) - 1);
        Kmer_Lanes<k, l>::minimizer_hashes(C.data(), C.size(), H.data());

        std::FILE* f = std::fopen((dir + "/scalars_" + tag + ".bin").c_str(), "wb");
        put<uint64_t>(f, X.size());
        for(std::size_t i = 0; i < X.size(); ++i)
        {
            const uint64_t x = X[i];
            put<uint64_t>(f, x);
            put<uint64_t>(f, wang_u64(x));
            put<uint64_t>(f, kmer_checksum(x));
            put<uint64_t>(f, C[i]), put<uint64_t>(f, H[i]);
            put<uint64_t>(f, XXH3_64bits_withSeed(&x, sizeof(x), 0));
            put<uint64_t>(f, XXH3_64bits_withSeed(&x, sizeof(x), uint64_t(1) << 63));
            put<uint64_t>(f, XXH3_64bits(&x, sizeof(x)));
        }
        std::fclose(f);
    }

    std::printf("gen_golden %s: %zu sequences written to %s\n", tag.c_str(), S.size(), dir.c_str());
    return 0;
}
