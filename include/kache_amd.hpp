#ifndef KACHE_AMD_HPP
#define KACHE_AMD_HPP

// The reference's C++ template API, Streaming_Kmer_Hash_Table<k, mt, T, l>
// (include/Streaming_Kmer_Hash_Table.hpp:618-746), over the MI355X table's C
// ABI.  The reference takes arbitrary functors f(v), f(v, fwd) and
// callbacks; a device table applies a fixed set of value operations, so `f`
// is one of the operation types below, and lookups call `g` on the host from
// the device's per-position outputs.  Method names, arguments and return
// values are the reference's; errors throw (std::invalid_argument for what
// the C ABI calls EINVAL, std::bad_alloc for ENOMEM, std::runtime_error for
// the rest).
//
//   kache_amd::Streaming_Kmer_Hash_Table<31, true, uint32_t, 22> M(n);
//   auto tok = M.register_user();
//   M.upsert(seqs, lens, c, kache_amd::Add_Wrap{1}, uint32_t(1), tok);
//   M.find(seqs, lens, c, [](std::size_t q, std::size_t i, std::optional<uint32_t> v){ ... });

#include "kache_amd.h"

#include <cstddef>
#include <cstdint>
#include <new>
#include <optional>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

namespace kache_amd
{

// Value operations: a present k-mer's value v becomes f(v); an absent one is
// inserted with the operand (the reference's `val`, which must equal it).
struct Keep      { uint64_t x = 0; static constexpr uint32_t op = KACHE_KEEP; };       // v
struct Assign    { uint64_t x;     static constexpr uint32_t op = KACHE_ASSIGN; };     // x
struct Add_Sat   { uint64_t x;     static constexpr uint32_t op = KACHE_ADD; };        // min(v + x, max)
struct Add_Wrap  { uint64_t x;     static constexpr uint32_t op = KACHE_AMD_ADD_WRAP; };  // v + x modulo 2^bits
struct Or        { uint64_t x;     static constexpr uint32_t op = KACHE_OR; };         // v | x
struct Min       { uint64_t x;     static constexpr uint32_t op = KACHE_MIN; };        // min(v, x)
struct Max       { uint64_t x;     static constexpr uint32_t op = KACHE_MAX; };        // max(v, x)

// Oriented operations for upsert_oriented: f(v, fwd) = Op(v, fwd ? x_fwd : x_rev),
// from v = T() for an absent k-mer.
template <typename Op> struct Oriented { uint64_t x_fwd, x_rev; };

namespace detail
{

inline void check(int st, const char* what)
{
    if(st == KACHE_OK)
        return;
    const std::string msg = std::string(what) + ": " + kache_strerror(st) + (kache_amd_last_error()[0] ? std::string(" (") + kache_amd_last_error() + ")" : std::string());
    if(st == KACHE_EINVAL)
        throw std::invalid_argument(msg);
    if(st == KACHE_ENOMEM)
        throw std::bad_alloc();
    throw std::runtime_error(msg);
}

template <typename T> struct Width { static constexpr uint32_t bytes = sizeof(T); };
template <> struct Width<void> { static constexpr uint32_t bytes = 0; };

template <typename T> struct is_oriented : std::false_type {};
template <typename Op> struct is_oriented<Oriented<Op>> : std::true_type {};

}

// A k-mer as the C ABI hands it out: 2-bit bases, first base most significant.
template <uint16_t k>
class Kmer
{
    uint64_t w = 0;

public:
    Kmer() = default;
    explicit Kmer(uint64_t word) : w(word) {}
    uint64_t as_int() const { return w; }
    std::string string_label() const
    {
        std::string s(k, 'A');
        for(uint16_t i = 0; i < k; ++i)
            s[i] = "ACGT"[(w >> (2 * (k - 1 - i))) & 3];
        return s;
    }
    bool operator==(const Kmer& o) const { return w == o.w; }
    bool operator<(const Kmer& o) const { return w < o.w; }
};

template <uint16_t k, bool mt_, typename T_, uint16_t l>
class Streaming_Kmer_Hash_Table
{
    static_assert(k == 31 && l == 22, "the device table supports k = 31, l = 22");
    static_assert(std::is_void_v<T_> || std::is_same_v<T_, uint8_t> || std::is_same_v<T_, uint16_t> ||
                  std::is_same_v<T_, uint32_t> || std::is_same_v<T_, uint64_t>, "values are void or unsigned of 1, 2, 4 or 8 bytes");

    static constexpr bool is_set_ = std::is_void_v<T_>;
    static constexpr bool is_map_ = !is_set_;
    static constexpr uint32_t width = detail::Width<T_>::bytes;

    kache_table* t = nullptr;

public:
    // The result of a lookup: the value in a map, presence in a set.
    using find_ret_t = std::conditional_t<is_set_, bool, std::optional<std::conditional_t<is_set_, int, T_>>>;

    class Token
    {
        friend class Streaming_Kmer_Hash_Table;
        uint32_t id;
        explicit Token(uint32_t id) : id(id) {}
    };

    explicit Streaming_Kmer_Hash_Table(std::size_t max_sz = 1 << 22, uint64_t resize_worker_c = 16, double lf = 0.8, bool pow2_cap = false)
    {
        kache_options o = kache_default_options(k, width);
        o.expected_keys = max_sz, o.load_factor = lf, o.resize_workers = static_cast<uint32_t>(resize_worker_c);
        o.pow2_capacity = pow2_cap;
        detail::check(kache_new(&o, &t), "kache_new");
    }

    static bool fits(std::size_t max_sz, double lf, bool pow2_cap) { return kache_amd_host_fits(max_sz, lf, pow2_cap); }

    Streaming_Kmer_Hash_Table(const Streaming_Kmer_Hash_Table&) = delete;
    Streaming_Kmer_Hash_Table& operator=(const Streaming_Kmer_Hash_Table&) = delete;
    ~Streaming_Kmer_Hash_Table() { kache_free(t); }

    // The C ABI handle, for the extensions of kache_amd.h.
    kache_table* handle() const { return t; }

    auto capacity() const { return stats().capacity; }
    std::size_t bucket_bytes() const { return stats().bytes; }
    std::size_t size() const
    {
        uint64_t n = 0;
        detail::check(kache_size(t, &n), "kache_size");
        return n;
    }
    std::size_t overflow_size() const { return stats().overflow_keys; }

    Token register_user()
    {
        uint32_t u = 0;
        detail::check(kache_register(t, &u), "kache_register");
        return Token(u);
    }

    std::size_t insert(const char* seq, std::size_t len, const Token& token) requires(is_set_) { return insert(&seq, &len, 1, token); }
    std::size_t insert(const char* const* seqs, const std::size_t* lens, std::size_t c, const Token& token) requires(is_set_)
    {
        uint64_t n = 0;
        detail::check(kache_insert(t, token.id, seqs, lens, c, KACHE_KEEP, 0, &n), "kache_insert");
        return n;
    }

    template <typename F, typename V = T_>
    std::size_t upsert(const char* seq, std::size_t len, const F& f, const V& val, const Token& token) requires(is_map_)
    {
        return upsert(&seq, &len, 1, f, val, token);
    }
    template <typename F, typename V = T_>
    std::size_t upsert(const char* const* seqs, const std::size_t* lens, std::size_t c, const F& f, const V& val, const Token& token) requires(is_map_)
    {
        static_assert(!detail::is_oriented<F>::value, "oriented operations go to upsert_oriented");
        if(static_cast<uint64_t>(val) != f.x)
            throw std::invalid_argument("upsert: the device table inserts absent k-mers with the operation's operand; val must equal it");
        uint64_t n = 0;
        detail::check(kache_amd_insert_seqs(t, token.id, seqs, lens, c, F::op, f.x, &n), "upsert");
        return n;
    }

    template <typename F>
    std::size_t upsert_oriented(const char* const* seqs, const std::size_t* lens, std::size_t c, const F& f, const Token& token) requires(is_map_)
    {
        static_assert(detail::is_oriented<F>::value, "upsert_oriented takes an Oriented<Assign|Add_Sat|Or|Max>");
        uint64_t n = 0;
        detail::check(kache_insert_oriented(t, token.id, seqs, lens, c, op_of(f), f.x_fwd, f.x_rev, &n), "upsert_oriented");
        return n;
    }

    template <typename F>
    std::size_t update(const char* seq, std::size_t len, const F& f) requires(is_map_) { return update(&seq, &len, 1, f); }
    template <typename F>
    std::size_t update(const char* const* seqs, const std::size_t* lens, std::size_t c, const F& f) requires(is_map_)
    {
        uint64_t n = 0;
        detail::check(kache_amd_update_seqs(t, seqs, lens, c, F::op, f.x, &n), "update");
        return n;
    }

    // g(i, r) per k-mer position i of `seq` (positions with a non-ACGT base
    // are skipped); in a map r is the value iff present.
    template <typename G>
    void find(const char* seq, std::size_t len, const G& g) const
    {
        find(&seq, &len, 1, [&](std::size_t, std::size_t i, const find_ret_t& r){ g(i, r); });
    }

    // g(q, i, r), or g(q, i, r, fwd) if it takes it, per k-mer position i of
    // sequence q (in sequence order, which the reference leaves unspecified).
    template <typename G>
    void find(const char* const* seqs, const std::size_t* lens, std::size_t c, const G& g) const
    {
        uint64_t p = 0;
        for(std::size_t q = 0; q < c; ++q)
            p += lens[q] >= k ? lens[q] - (k - 1) : 0;
        if(p == 0)
            return;
        std::vector<uint8_t> found(p), fwd(p);
        std::vector<uint8_t> vals(is_map_ ? p * width : 0);
        detail::check(kache_lookup(t, seqs, lens, c, found.data(), is_map_ ? vals.data() : nullptr, fwd.data()), "find");
        uint64_t o = 0;
        for(std::size_t q = 0; q < c; ++q)
        {
            const std::size_t n = lens[q] >= k ? lens[q] - (k - 1) : 0;
            for(std::size_t i = 0; i < n; ++i, ++o)
            {
                if(found[o] == KACHE_NOT_KMER)
                    continue;
                find_ret_t r{};
                if constexpr(is_set_)
                    r = (found[o] == KACHE_PRESENT);
                else if(found[o] == KACHE_PRESENT)
                {
                    T_ v;
                    __builtin_memcpy(&v, vals.data() + o * width, width);
                    r = v;
                }
                if constexpr(std::is_invocable_v<const G&, std::size_t, std::size_t, const find_ret_t&, bool>)
                    g(q, i, r, fwd[o] != 0);
                else
                    g(q, i, r);
            }
        }
    }

    find_ret_t find(const char* kmer) const
    {
        uint8_t f = 0;
        uint64_t v = 0;
        detail::check(kache_get(t, kmer, &f, &v), "find");
        if constexpr(is_set_)
            return f != 0;
        else
            return f ? find_ret_t(static_cast<T_>(v)) : find_ret_t();
    }

    std::size_t count(const char* seq, std::size_t len) const { return count(&seq, &len, 1); }
    std::size_t count(const char* const* seqs, const std::size_t* lens, std::size_t c) const
    {
        std::vector<uint64_t> n(c ? c : 1);
        detail::check(kache_count(t, seqs, lens, c, n.data()), "count");
        std::size_t s = 0;
        for(std::size_t q = 0; q < c; ++q)
            s += n[q];
        return s;
    }

    // f(x) per k-mer of a set, f(x, v) per k-mer and value of a map, in share
    // w_id of w_c (the shares partition the table).
    template <typename F>
    void for_each(const F& f, std::size_t w_c, std::size_t w_id) const
    {
        struct Ctx { const F* f; };
        Ctx ctx{&f};
        const auto cb = [](void* c, const uint64_t* keys, const void* values, size_t n) -> int
        {
            const F& fn = *static_cast<Ctx*>(c)->f;
            for(size_t i = 0; i < n; ++i)
                if constexpr(is_set_)
                    fn(Kmer<k>(keys[i]));
                else
                {
                    T_ v;
                    __builtin_memcpy(&v, static_cast<const uint8_t*>(values) + i * width, width);
                    fn(Kmer<k>(keys[i]), v);
                }
            return 0;
        };
        detail::check(kache_for_each(t, static_cast<uint32_t>(w_c), static_cast<uint32_t>(w_id), cb, &ctx), "for_each");
    }

    // Percentages of keys in their primary bucket, a secondary bucket, and the
    // overflow table.
    std::tuple<double, double, double> key_distribution() const
    {
        const kache_stats s = stats();
        return {s.primary_pct, s.secondary_pct, s.overflow_pct};
    }

private:
    kache_stats stats() const
    {
        kache_stats s{};
        detail::check(kache_get_stats(t, &s), "kache_get_stats");
        return s;
    }

    template <typename Op>
    static uint32_t op_of(const Oriented<Op>&)
    {
        static_assert(std::is_same_v<Op, Assign> || std::is_same_v<Op, Add_Sat> || std::is_same_v<Op, Or> || std::is_same_v<Op, Max>,
                      "oriented operations: Assign, Add_Sat, Or, Max (capi/Kmer_Table.hpp:50-61)");
        return Op::op;
    }
};

}

#endif
