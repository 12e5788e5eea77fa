// Fixture writer for the describer, hash and placement golden vectors.
//
// Built against the reference headers (include/Kmer_Batch.hpp, include/Kmer_Lanes.hpp,
// include/DNA_Utility.hpp, include/xxHash/xxhash.h, include/Streaming_Kmer_Hash_Table.hpp,
// include/Concurrent_Hash_Table.hpp) by `make -C oracle golden`, which compiles it twice:
// once for AVX2 (the multiply checksum, include/Kmer_Batch.hpp:49-50) and once for AVX-512
// (the Wang checksum, include/Kmer_Batch.hpp:47-48).  It writes plain little-endian binary
// files, which tests/golden/make_golden.py packs into .npz fixtures:
//   seqs.bin            u64 n, u64 lens[n], then the n sequences' bytes back to back
//   describe_<tag>.bin  one 32-byte record per k-mer: u32 q, u32 i, u64 key, u64 h, u8 c, u8 fwd, 6 pad bytes
//   scalars_<tag>.bin   u64 n, then n rows of 8 u64: x, wang_u64(x), kmer_checksum(x), canonical 62-bit x,
//                       its minimizer hash, XXH3_64bits_withSeed(&x, 8, 0), the same with seed 2^63,
//                       XXH3_64bits(&x, 8)
//   placement_<tag>.bin the table's private placement helpers, called through their own code:
//                       u64 n_cap, caps[n_cap], u64 n_h, then n_h rows of (h, filter_pos(h),
//                       n_cap x (bucket_of(h, cap), pick_bucket(h, cap))); u64 n_tab, then per table
//                       (u64 max_sz, u64 buckets, u64 overflow slots, u64 n_sec, n_sec x (h, p, q) of
//                       secondaries(h)); u64 n_ov, then n_ov pairs (max_n, overflow-table capacity)
//
// Usage: gen_golden <out_dir> <tag>

#include "DNA_Utility.hpp"
#include "Kmer_Batch.hpp"
#include "Kmer_Lanes.hpp"
#include "Streaming_Kmer_Hash_Table.hpp"
#include "Concurrent_Hash_Table.hpp"
#include "xxHash/xxhash.h"

#include <cstdint>
#include <cstdio>
#include <string>
#include <utility>
#include <vector>

using namespace kache_hash;

namespace
{

constexpr uint16_t k = 31, l = 22;

typedef Streaming_Kmer_Hash_Table<k, true, void, l> Set_Table;

// Access to the table's private placement helpers for the fixtures: the names
// of an explicit instantiation are not access-checked ([temp.spec]/6), so
// Reach<Tag, &Set_Table::f> hands out a pointer to f through get(Tag).
template <typename Tag, typename Tag::type M>
struct Reach
{
    friend typename Tag::type get(Tag) { return M; }
};

struct Bucket_Of { typedef std::size_t (*type)(uint64_t, std::size_t); friend type get(Bucket_Of); };
struct Pick_Bucket { typedef std::size_t (*type)(uint64_t, std::size_t); friend type get(Pick_Bucket); };
struct Filter_Pos { typedef std::size_t (*type)(uint64_t); friend type get(Filter_Pos); };
struct Secondaries { typedef std::pair<std::size_t, std::size_t> (Set_Table::*type)(uint64_t) const; friend type get(Secondaries); };

template struct Reach<Bucket_Of, &Set_Table::bucket_of>;
template struct Reach<Pick_Bucket, &Set_Table::pick_bucket>;
template struct Reach<Filter_Pos, &Set_Table::filter_pos>;
template struct Reach<Secondaries, &Set_Table::secondaries>;

// The overflow table's capacity, through its own member.
template <typename Tag, auto M>
struct Reach_Auto
{
    friend auto get(Tag) { return M; }
};
struct Ov_Member { friend auto get(Ov_Member); };
template struct Reach_Auto<Ov_Member, &Set_Table::ov>;

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

    // 3. Placement: bucket_of, pick_bucket and filter_pos over seeded hashes and
    //    capacities (include/Streaming_Kmer_Hash_Table.hpp:1014-1019, 1055-1068),
    //    secondaries of live tables (:2111-2115), and overflow capacities, of
    //    the overflow table alone (include/Concurrent_Hash_Table.hpp:158-164) and
    //    as a table sizes it (include/Streaming_Kmer_Hash_Table.hpp:775).
    {
        SplitMix r{0x9A7E5ull};
        const std::vector<uint64_t> caps{1, 2, 3, 7, 40, 64, 1000, 4095, 3906251, 78125001, (uint64_t(1) << 31) + 11,
                                         (uint64_t(1) << 32) - 1};
        std::vector<uint64_t> H{0, 1, 0xFFFFFFFFull, 0x100000000ull, 0xFF00000000ull, ~uint64_t(0), uint64_t(1) << 63};
        while(H.size() < 2048)
            H.push_back(r.next());

        std::FILE* f = std::fopen((dir + "/placement_" + tag + ".bin").c_str(), "wb");
        put<uint64_t>(f, caps.size());
        for(const auto c : caps)
            put<uint64_t>(f, c);
        put<uint64_t>(f, H.size());
        for(const auto h : H)
        {
            put<uint64_t>(f, h), put<uint64_t>(f, get(Filter_Pos{})(h));
            for(const auto c : caps)
                put<uint64_t>(f, get(Bucket_Of{})(h, c)), put<uint64_t>(f, get(Pick_Bucket{})(h, c));
        }

        const std::vector<uint64_t> sizes{10, 1000, 40000, 1 << 20};
        put<uint64_t>(f, sizes.size());
        for(const auto n : sizes)
        {
            const Set_Table t(n, 1);
            put<uint64_t>(f, n), put<uint64_t>(f, t.capacity() / 32), put<uint64_t>(f, (t.*get(Ov_Member{}))->capacity());
            const std::size_t n_sec = 256;
            put<uint64_t>(f, n_sec);
            for(std::size_t i = 0; i < n_sec; ++i)
            {
                const auto [p, q] = (t.*get(Secondaries{}))(H[i]);
                put<uint64_t>(f, H[i]), put<uint64_t>(f, p), put<uint64_t>(f, q);
            }
        }

        const std::vector<uint64_t> ov_n{1, 2, 800, 819, 820, 1023, 1024, 1025, 3906251, 5000000};
        put<uint64_t>(f, ov_n.size());
        for(const auto n : ov_n)
        {
            const Concurrent_Hash_Table<Kmer<k>, uint64_t, Hash<Kmer<k>>> ov(n);
            put<uint64_t>(f, n), put<uint64_t>(f, ov.capacity());
        }
        std::fclose(f);
    }

    std::printf("gen_golden %s: %zu sequences written to %s\n", tag.c_str(), S.size(), dir.c_str());
    return 0;
}
