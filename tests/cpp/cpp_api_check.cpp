// Drives the C++ API headers (include/kache_amd.hpp, include/kache_amd_counter.hpp)
// over two FASTA/FASTQ files and writes what they return, for the GPU tests to
// compare with the oracle:
//   cpp_api_check <reads> <segments> <out_dir>
// writes <out_dir>/<name>.u64 arrays (little-endian) and <out_dir>/scalars.txt.

#include "kache_amd.hpp"
#include "kache_amd_counter.hpp"

#include <algorithm>
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

namespace
{

std::vector<std::string> read_file(const std::string& path)
{
    std::string text;
    const auto cb = [](void* ctx, const char* p, uint64_t n) -> int
    {
        static_cast<std::string*>(ctx)->append(p, n);
        return 0;
    };
    if(kache_amd_read_fastx(path.c_str(), cb, &text) != KACHE_OK)
        throw std::runtime_error("cannot read " + path + ": " + kache_amd_last_error());
    std::vector<std::string> out;
    size_t b = 0;
    for(size_t i = 0; i < text.size(); ++i)
        if(text[i] == '\n')
            out.emplace_back(text, b, i - b), b = i + 1;
    return out;
}

struct Batches
{
    std::vector<std::vector<const char*>> seqs;
    std::vector<std::vector<size_t>> lens;
    explicit Batches(const std::vector<std::string>& S, size_t per = 1024)
    {
        for(size_t i = 0; i < S.size(); i += per)
        {
            seqs.emplace_back(), lens.emplace_back();
            for(size_t j = i; j < std::min(S.size(), i + per); ++j)
                seqs.back().push_back(S[j].data()), lens.back().push_back(S[j].size());
        }
    }
};

void write(const std::string& path, const std::vector<uint64_t>& v)
{
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char*>(v.data()), static_cast<std::streamsize>(v.size() * 8));
}

}

int main(int argc, char** argv)
{
    if(argc != 4)
    {
        std::fprintf(stderr, "usage: %s <reads> <segments> <out_dir>\n", argv[0]);
        return 2;
    }
    const std::string dir = argv[3];
    const auto R = read_file(argv[1]), S = read_file(argv[2]);
    const Batches RB(R), SB(S);
    std::ofstream sc(dir + "/scalars.txt");

    // A set: insert the reads; count the segments.
    {
        kache_amd::Streaming_Kmer_Hash_Table<31, true, void, 22> M(100'000);
        const auto tok = M.register_user();
        size_t ins = 0;
        for(size_t b = 0; b < RB.seqs.size(); ++b)
            ins += M.insert(RB.seqs[b].data(), RB.lens[b].data(), RB.seqs[b].size(), tok);
        size_t cnt = 0;
        for(size_t b = 0; b < SB.seqs.size(); ++b)
            cnt += M.count(SB.seqs[b].data(), SB.lens[b].data(), SB.seqs[b].size());
        std::vector<uint64_t> keys;
        for(size_t w = 0; w < 3; ++w)
            M.for_each([&](const kache_amd::Kmer<31>& x){ keys.push_back(x.as_int()); }, 3, w);
        std::sort(keys.begin(), keys.end());
        write(dir + "/set_keys.u64", keys);
        sc << "set_inserted " << ins << "\nset_size " << M.size() << "\nset_count " << cnt << "\nset_capacity " << M.capacity() << "\n";
        sc << "set_find_first " << M.find(S[0].data()) << "\n";
    }

    // A u32 map: upsert v + 1 from 1 (wrapping) per read; update max(v, 7) over
    // the segments; lookups of the segments with orientations.
    {
        kache_amd::Streaming_Kmer_Hash_Table<31, true, uint32_t, 22> M(100'000);
        const auto tok = M.register_user();
        size_t ins = 0, pres = 0;
        for(size_t b = 0; b < RB.seqs.size(); ++b)
            ins += M.upsert(RB.seqs[b].data(), RB.lens[b].data(), RB.seqs[b].size(), kache_amd::Add_Wrap{1}, uint32_t(1), tok);
        for(size_t b = 0; b < SB.seqs.size(); ++b)
            pres += M.update(SB.seqs[b].data(), SB.lens[b].data(), SB.seqs[b].size(), kache_amd::Max{7});
        std::vector<uint64_t> found;     // per k-mer position of the segments: q, i, value + 1 (0 absent), fwd
        const auto& s0 = SB.seqs[0];
        M.find(s0.data(), SB.lens[0].data(), s0.size(), [&](size_t q, size_t i, const std::optional<uint32_t>& r, bool fwd)
        {
            found.insert(found.end(), {q, i, r ? uint64_t(*r) + 1 : 0, uint64_t(fwd)});
        });
        write(dir + "/map_find.u64", found);
        std::vector<uint64_t> kv;
        M.for_each([&](const kache_amd::Kmer<31>& x, uint32_t v){ kv.push_back(x.as_int()), kv.push_back(v); }, 1, 0);
        write(dir + "/map_kv.u64", kv);
        sc << "map_inserted " << ins << "\nmap_present " << pres << "\nmap_size " << M.size() << "\n";
        try
        {
            M.upsert(RB.seqs[0].data(), RB.lens[0].data(), 1, kache_amd::Add_Sat{2}, uint32_t(3), tok);
            sc << "map_val_mismatch_refused 0\n";
        }
        catch(const std::invalid_argument&)
        {
            sc << "map_val_mismatch_refused 1\n";
        }
    }

    // A u8 map: oriented OR of 1 (forward) / 2 (reverse complement).
    {
        kache_amd::Streaming_Kmer_Hash_Table<31, true, uint8_t, 22> M(100'000);
        const auto tok = M.register_user();
        size_t ins = 0;
        for(size_t b = 0; b < RB.seqs.size(); ++b)
            ins += M.upsert_oriented(RB.seqs[b].data(), RB.lens[b].data(), RB.seqs[b].size(), kache_amd::Oriented<kache_amd::Or>{1, 2}, tok);
        std::vector<uint64_t> kv;
        M.for_each([&](const kache_amd::Kmer<31>& x, uint8_t v){ kv.push_back(x.as_int()), kv.push_back(v); }, 1, 0);
        write(dir + "/oriented_kv.u64", kv);
        sc << "oriented_inserted " << ins << "\n";
    }

    // The counter: prime with the segments, count the reads.
    {
        kache_amd::Kmer_Set_Counter C(31, 100'000, 16);
        C.prime(argv[2]);
        const size_t found = C.count(argv[1]);
        const auto h = C.histogram(50);
        write(dir + "/counter_hist.u64", std::vector<uint64_t>(h.begin(), h.end()));
        sc << "counter_size " << C.size() << "\ncounter_found " << found << "\ncounter_count_of_first " << C.count_of(S[0].data()) << "\n";
        bool threw = false;
        try
        {
            C.count(std::string(argv[3]) + "/absent.fa");
        }
        catch(const std::runtime_error&)
        {
            threw = true;
        }
        sc << "counter_absent_throws " << threw << "\n";
    }
    return 0;
}
