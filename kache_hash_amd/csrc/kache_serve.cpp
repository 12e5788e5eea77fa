// Serving loop over a resident table (kache_amd_server_* in include/kache_amd.h).
// Requests (per-sequence counts, per-position lookups, present-only updates,
// upserts) arrive from any thread; one worker admits them in arrival order
// into waves, as many as fit the wave's staging budget in HBM, and runs each
// wave as: its reads merged into a few calls per kind (one staging copy;
// smaller requests in earlier calls, since reads commute), its updates, then
// its upserts, packed into one device text and
// inserted by one pipelined batch call per run of equal op and operand (each
// request one batch, so its count of new keys is its own).  Requests arriving
// while a wave runs join the next one; a lone request reaching an idle server
// runs at once by the plain calls, unstaged.  Each request sees exactly what
// a serial run in served order (wave, phase, ticket) gives it.
//
// It uses only the C ABI and the device extensions, like any other client.

#include "kache_amd.h"

// The library's packing threads (kache_table.hip); `aside`: its second pool,
// which a call in progress never waits for.
void kamd_copy_spans(char* const* dst, const char* const* src, const uint64_t* len, size_t n, bool aside);

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace
{

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Req
{
    kache_amd_request r;
    uint64_t ticket = 0;
    uint64_t bytes = 0;     // staged bytes: its sequences and separators
    int status = KACHE_OK;
    bool done = false;
    uint32_t phase = 0;
    uint64_t wave = 0;
    double t_submit = 0, t_start = 0, t_done = 0;
};

uint64_t staged_bytes(const kache_amd_request& r)
{
    uint64_t b = 0;
    for(size_t i = 0; i < r.n; ++i)
        b += r.lens[i] + 1;
    return b;
}

// Lookup positions of a request: sum over its sequences of max(0, len - 30)
// (capi/Kmer_Table.hpp:163-190).
uint64_t positions_of(const kache_amd_request& r)
{
    uint64_t p = 0;
    for(size_t i = 0; i < r.n; ++i)
        p += r.lens[i] >= 31 ? r.lens[i] - 30 : 0;
    return p;
}

// How long a server thread polls for work before it sleeps on its condition
// variable: a sleeping thread takes tens of microseconds to run again after a
// notify, which a request arriving at an idle server would pay at every
// hand-off between threads.
constexpr double POLL_S = 2e-3;

// cv.wait(lock, pred), polling pred for up to POLL_S first (the lock is
// released between polls).
template <class Pred>
void poll_wait(std::unique_lock<std::mutex>& lock, std::condition_variable& cv, Pred pred)
{
    const double until = now_s() + POLL_S;
    while(!pred() && now_s() < until)
    {
        lock.unlock();
        for(int i = 0; i < 32; ++i)
            __builtin_ia32_pause();
        lock.lock();
    }
    cv.wait(lock, pred);
}

uint32_t phase_of(uint32_t kind)
{
    return kind == KACHE_AMD_REQ_UPDATE ? 1 : kind == KACHE_AMD_REQ_INSERT ? 2 : 0;
}

}

// A staged wave runs as a list of device calls, one after another; the
// requests of one call run in ticket order within it, so each request's
// results are those of a serial run in served order (wave, call, ticket).
// A call holds requests of one kind: per-sequence counts, lookups (one
// output group per request), one present-only update, or upserts of one
// operation and operand (one batch per request, so each request's count of
// new keys is its own).
//
// Unless the server is saturated, the calls run shortest first across kinds:
// the wave's requests in order of size, consecutive ones of one kind (and for
// upserts one operation) merged into a call while no request is larger than
// the call so far.  Any serial order of the requests of one wave is a valid
// one (they were all pending together), and this one keeps a small request
// from waiting for a large one of another kind.  A wave taken while more
// requests wait (saturated: throughput, not one wave's latency, decides) runs
// its counts in one call, its lookups in one, each update, then its upserts
// in ticket order, one call per run of equal operation and operand.
struct Call
{
    uint32_t kind = 0, op = 0;
    uint64_t x = 0;
    std::vector<Req*> reqs;             // in ticket order
    uint64_t bytes = 0;                 // staged bytes of its requests
    uint64_t at = 0;                    // start of its region in the slot's text (16-byte aligned)
    std::vector<uint64_t> off;          // reads: sequence offsets from `at`, and the end; upserts: batch bounds
};

// A wave staged for the device: its requests, its calls and where their
// sequences sit in the slot's text (each sequence followed by '\n').
struct Staged
{
    uint64_t id = 0;
    int slot = 0;
    std::vector<std::shared_ptr<Req>> reqs;
    std::vector<Call> calls;
    uint64_t bytes = 0;                                        // staged text bytes
    bool direct = false;                                       // not staged: run by the plain calls
    bool saturated = false;                                    // more requests waited when it was taken
};

struct kache_amd_server
{
    kache_table* t;
    uint32_t user;
    uint64_t wave_bytes;
    uint32_t max_wave_requests;
    uint32_t width;

    std::mutex mu;
    std::condition_variable cv_in, cv_done, cv_staged, cv_slot;
    std::deque<std::shared_ptr<Req>> pending;
    std::map<uint64_t, std::shared_ptr<Req>> live;
    uint64_t next_ticket = 1;
    bool stopping = false;
    std::thread worker, stager;

    // Plain path: staging of a wave's upserts.
    char* h_text = nullptr;
    void* d_text = nullptr;
    uint64_t h_bytes = 0, d_bytes = 0;

    // Pipelined path: two slots, each a pinned host text and its device copy.
    // The stager packs and copies wave k + 1 into one while the executor
    // runs wave k from the other.
    static constexpr int SLOTS = 2;
    char* h_slot[SLOTS] = {};
    void* d_slot[SLOTS] = {};
    uint64_t slot_bytes[SLOTS] = {};
    bool slot_busy[SLOTS] = {};
    std::deque<std::unique_ptr<Staged>> staged;
    bool stager_done = false;
    bool exec_busy = false;     // the executor is running a wave

    kache_amd_server_stats stats{};
    std::atomic<uint64_t> hbm_budget{0};       // staging bytes of the slots together (0: no budget of its own)

    kache_amd_server(kache_table* t, uint32_t user, uint64_t wave_bytes, uint32_t max_req, uint32_t width)
        : t(t), user(user), wave_bytes(wave_bytes), max_wave_requests(max_req), width(width)
    {
    }

    // Allocates the pipeline's two slots at their wave size up front, so
    // that no wave waits for an allocation (a larger lone request still
    // grows its slot), then starts the threads.
    int start()
    {
        if(max_wave_requests == 1)
        {
            worker = std::thread([this]{ loop(); });
            return KACHE_OK;
        }
        const uint64_t want = wave_bytes + 4096 + KACHE_AMD_TEXT_PAD;
        for(int i = 0; i < SLOTS; ++i)
        {
            void* hp = nullptr;
            int st = kache_amd_host_alloc(want, &hp);
            if(st != KACHE_OK)
                return st;
            h_slot[i] = static_cast<char*>(hp);
            st = kache_amd_device_alloc(want, &d_slot[i]);
            if(st != KACHE_OK)
                return st;
            slot_bytes[i] = want;
        }
        stager = std::thread([this]{ stage_loop(); });
        worker = std::thread([this]{ exec_loop(); });
        return KACHE_OK;
    }

    ~kache_amd_server()
    {
        {
            std::lock_guard<std::mutex> lock(mu);
            stopping = true;
        }
        cv_in.notify_all();
        cv_slot.notify_all();
        if(stager.joinable())
            stager.join();
        if(worker.joinable())
            worker.join();
        if(d_text)
            kache_amd_device_free(d_text);
        if(h_text)
            kache_amd_host_free(h_text);
        for(int i = 0; i < SLOTS; ++i)
        {
            if(d_slot[i])
                kache_amd_device_free(d_slot[i]);
            if(h_slot[i])
                kache_amd_host_free(h_slot[i]);
        }
    }

    // Takes the pending requests of the next wave (the lock held): in
    // arrival order, as many as fit wave_bytes (at least one) and
    // max_wave_requests.
    std::vector<std::shared_ptr<Req>> take_wave(uint64_t& wave_id)
    {
        std::vector<std::shared_ptr<Req>> wave;
        uint64_t bytes = 0;
        while(!pending.empty() && wave.size() < max_wave_requests && (wave.empty() || bytes + pending.front()->bytes <= wave_bytes))
        {
            bytes += pending.front()->bytes;
            wave.push_back(pending.front());
            pending.pop_front();
        }
        wave_id = ++stats.waves;
        stats.requests += wave.size();
        stats.max_wave_requests = std::max<uint64_t>(stats.max_wave_requests, wave.size());
        stats.max_wave_bytes = std::max(stats.max_wave_bytes, bytes);
        return wave;
    }

    // --- Plain path: one worker, one request per wave, plain C-ABI calls ---

    void loop()
    {
        while(true)
        {
            std::vector<std::shared_ptr<Req>> wave;
            uint64_t wave_id = 0;
            {
                std::unique_lock<std::mutex> lock(mu);
                poll_wait(lock, cv_in, [&]{ return stopping || !pending.empty(); });
                if(pending.empty())
                    return;
                wave = take_wave(wave_id);
            }
            const double t0 = now_s();
            for(auto& q : wave)
                q->t_start = t0, q->wave = wave_id, q->phase = phase_of(q->r.kind);
            try
            {
                run(wave);
            }
            catch(...)
            {
                fail(wave);
            }
        }
    }

    // Completes the requests of a wave that could not run (host memory ran
    // out): each gets KACHE_ENOMEM unless it already completed.
    void fail(const std::vector<std::shared_ptr<Req>>& wave)
    {
        std::vector<Req*> open;
        {
            std::lock_guard<std::mutex> lock(mu);
            for(auto& q : wave)
                if(!q->done)
                    open.push_back(q.get());
        }
        for(Req* q : open)
            q->status = KACHE_ENOMEM;
        finish(open);
    }

    // Marks requests done as soon as their phase ends: later phases of the
    // wave cannot change their results.
    void finish(const std::vector<Req*>& rs)
    {
        if(rs.empty())
            return;
        const double t1 = now_s();
        {
            std::lock_guard<std::mutex> lock(mu);
            for(Req* q : rs)
                q->t_done = t1, q->done = true;
        }
        cv_done.notify_all();
    }

    void run(std::vector<std::shared_ptr<Req>>& wave)
    {
        std::vector<Req*> counts, lookups, updates, inserts;
        for(auto& q : wave)
            switch(q->r.kind)
            {
            case KACHE_AMD_REQ_COUNT: counts.push_back(q.get()); break;
            case KACHE_AMD_REQ_LOOKUP: lookups.push_back(q.get()); break;
            case KACHE_AMD_REQ_UPDATE: updates.push_back(q.get()); break;
            default: inserts.push_back(q.get()); break;
            }
        run_counts(counts);
        finish(counts);
        run_lookups(lookups);
        finish(lookups);
        for(Req* q : updates)
            q->status = kache_update(t, q->r.seqs, q->r.lens, q->r.n, q->r.op, q->r.x, q->r.counts);
        finish(updates);
        for(size_t a = 0; a < inserts.size();)
        {
            size_t b = a + 1;
            while(b < inserts.size() && inserts[b]->r.op == inserts[a]->r.op && inserts[b]->r.x == inserts[a]->r.x)
                b++;
            run_inserts(inserts, a, b);
            a = b;
        }
        finish(inserts);
    }

    // All sequences of the requests, in order, as one call's arrays.
    static void merge(const std::vector<Req*>& rs, std::vector<const char*>& seqs, std::vector<size_t>& lens)
    {
        for(Req* q : rs)
        {
            seqs.insert(seqs.end(), q->r.seqs, q->r.seqs + q->r.n);
            lens.insert(lens.end(), q->r.lens, q->r.lens + q->r.n);
        }
    }

    void run_counts(const std::vector<Req*>& rs)
    {
        if(rs.size() == 1)
        {
            rs[0]->status = kache_count(t, rs[0]->r.seqs, rs[0]->r.lens, rs[0]->r.n, rs[0]->r.counts);
            return;
        }
        if(rs.empty())
            return;
        std::vector<const char*> seqs;
        std::vector<size_t> lens;
        merge(rs, seqs, lens);
        std::vector<uint64_t> out(std::max<size_t>(seqs.size(), 1));
        const int st = kache_count(t, seqs.data(), lens.data(), seqs.size(), out.data());
        size_t o = 0;
        for(Req* q : rs)
        {
            q->status = st;
            if(st == KACHE_OK && q->r.n)
                std::memcpy(q->r.counts, out.data() + o, q->r.n * sizeof(uint64_t));
            o += q->r.n;
        }
    }

    void run_lookups(const std::vector<Req*>& rs)
    {
        if(rs.size() == 1)
        {
            const auto& r = rs[0]->r;
            rs[0]->status = kache_lookup(t, r.seqs, r.lens, r.n, r.found, r.values, r.fwd);
            return;
        }
        if(rs.empty())
            return;
        std::vector<const char*> seqs;
        std::vector<size_t> lens;
        merge(rs, seqs, lens);
        std::vector<uint64_t> first;
        std::vector<uint8_t*> found, fwd;
        std::vector<void*> values;
        uint64_t o = 0;
        for(Req* q : rs)
        {
            first.push_back(o), o += q->r.n;
            found.push_back(q->r.found), values.push_back(q->r.values), fwd.push_back(q->r.fwd);
        }
        const int st = kache_amd_lookup_groups(t, seqs.data(), lens.data(), seqs.size(), first.data(), rs.size(), found.data(),
                                               values.data(), fwd.data());
        for(Req* q : rs)
            q->status = st;
    }

    void run_inserts(const std::vector<Req*>& ins, size_t a, size_t b)
    {
        if(b - a == 1)
        {
            const auto& r = ins[a]->r;
            ins[a]->status = kache_insert(t, user, r.seqs, r.lens, r.n, r.op, r.x, r.counts);
            return;
        }
        std::vector<uint64_t> bounds{0};
        uint64_t bytes = 0;
        for(size_t i = a; i < b; ++i)
            bytes += ins[i]->bytes;
        const uint64_t need = bytes + KACHE_AMD_TEXT_PAD;
        const uint64_t want = std::max<uint64_t>(bytes, wave_bytes) + KACHE_AMD_TEXT_PAD;
        int st = KACHE_OK;
        if(need > h_bytes)
        {
            if(h_text)
                kache_amd_host_free(h_text), h_text = nullptr, h_bytes = 0;
            void* p = nullptr;
            st = kache_amd_host_alloc(want, &p);
            if(st == KACHE_OK)
                h_text = static_cast<char*>(p), h_bytes = want;
        }
        if(st == KACHE_OK && need > d_bytes)
        {
            if(d_text)
                kache_amd_device_free(d_text), d_text = nullptr, d_bytes = 0;
            st = kache_amd_device_alloc(want, &d_text);
            if(st == KACHE_OK)
                d_bytes = want;
        }
        if(st == KACHE_OK)
        {
            uint64_t o = 0;
            for(size_t i = a; i < b; ++i)
            {
                const auto& r = ins[i]->r;
                for(size_t j = 0; j < r.n; ++j)
                {
                    if(r.lens[j])
                        std::memcpy(h_text + o, r.seqs[j], r.lens[j]);
                    o += r.lens[j];
                    h_text[o++] = '\n';
                }
                bounds.push_back(o);
            }
            std::memset(h_text + o, '\n', KACHE_AMD_TEXT_PAD);
        }
        std::vector<uint64_t> out(b - a);
        if(st == KACHE_OK)
            st = kache_amd_copy(d_text, h_text, bytes + KACHE_AMD_TEXT_PAD, 1);
        if(st == KACHE_OK)
            st = kache_amd_insert_batches(t, user, static_cast<const char*>(d_text), bounds.data(), b - a, ins[a]->r.op, ins[a]->r.x, out.data());
        for(size_t i = a; i < b; ++i)
        {
            ins[i]->status = st;
            if(st == KACHE_OK && ins[i]->r.counts)
                *ins[i]->r.counts = out[i - a];
        }
    }

    // --- Pipelined path ---

    static uint64_t align16(uint64_t x) { return (x + 15) & ~uint64_t(15); }

    // Admission of `extra` more bytes of staging for a request larger than
    // its slot: within the server's budget, if it has one, and within the
    // device's free memory less what the table's next doubling needs.
    int admit_growth(uint64_t extra)
    {
        kache_amd_info info{};
        int st = kache_amd_get_info(t, &info);
        const uint64_t budget = hbm_budget.load();
        if(st == KACHE_OK && budget != 0 && slot_bytes[0] + slot_bytes[1] + extra > budget)
            st = KACHE_ENOMEM;
        if(st == KACHE_OK && info.device_free_bytes < info.next_doubling_bytes + extra)
            st = KACHE_ENOMEM;
        if(st == KACHE_ENOMEM)
        {
            std::lock_guard<std::mutex> lock(mu);
            stats.refused++;
        }
        return st;
    }

    // The calls of a wave, in run order (see Call).
    static void plan(Staged& w)
    {
        const auto same_call = [](const Call& c, const Req* q)
        {
            return c.kind == q->r.kind && c.kind != KACHE_AMD_REQ_UPDATE &&
                   (c.kind != KACHE_AMD_REQ_INSERT || (c.op == q->r.op && c.x == q->r.x));
        };
        const auto open = [&](const Req* q)
        {
            Call c;
            c.kind = q->r.kind, c.op = q->r.op, c.x = q->r.x;
            w.calls.push_back(std::move(c));
        };
        std::vector<Req*> rs;
        for(auto& q : w.reqs)
            rs.push_back(q.get());
        if(w.saturated)
        {
            // Ticket order already; reads by kind, then updates, then upserts.
            for(uint32_t kind : {KACHE_AMD_REQ_COUNT, KACHE_AMD_REQ_LOOKUP, KACHE_AMD_REQ_UPDATE, KACHE_AMD_REQ_INSERT})
            {
                const size_t first = w.calls.size();
                for(Req* q : rs)
                    if(q->r.kind == kind)
                    {
                        if(w.calls.size() == first || !same_call(w.calls.back(), q))
                            open(q);
                        w.calls.back().reqs.push_back(q), w.calls.back().bytes += q->bytes;
                    }
            }
            return;
        }
        std::sort(rs.begin(), rs.end(), [](const Req* a, const Req* b) { return a->bytes < b->bytes || (a->bytes == b->bytes && a->ticket < b->ticket); });
        for(Req* q : rs)
        {
            if(w.calls.empty() || !same_call(w.calls.back(), q) || q->bytes > w.calls.back().bytes)
                open(q);
            w.calls.back().reqs.push_back(q), w.calls.back().bytes += q->bytes;
        }
        for(Call& c : w.calls)
            std::sort(c.reqs.begin(), c.reqs.end(), [](const Req* a, const Req* b) { return a->ticket < b->ticket; });
    }

    // Plans a wave, packs it into slot `w.slot`'s pinned text (one 16-byte
    // aligned region per call) and copies it to the device.
    int pack(Staged& w)
    {
        plan(w);
        uint64_t o = 0;
        for(Call& c : w.calls)
            c.at = o, o = align16(o + c.bytes);
        w.bytes = o;
        const uint64_t need = o + KACHE_AMD_TEXT_PAD;
        const int sl = w.slot;
        if(need > slot_bytes[sl])
        {
            // A request larger than the slot: it runs alone, in a slot grown
            // for it, if the HBM it needs is free (admission by budget).
            const int adm = admit_growth(need - slot_bytes[sl]);
            if(adm != KACHE_OK)
                return adm;
            if(h_slot[sl])
                kache_amd_host_free(h_slot[sl]), h_slot[sl] = nullptr;
            if(d_slot[sl])
                kache_amd_device_free(d_slot[sl]), d_slot[sl] = nullptr;
            slot_bytes[sl] = 0;
            const uint64_t want = std::max<uint64_t>(need, wave_bytes + 4096 + KACHE_AMD_TEXT_PAD);
            void* hp = nullptr;
            int st = kache_amd_host_alloc(want, &hp);
            if(st != KACHE_OK)
                return st;
            h_slot[sl] = static_cast<char*>(hp);
            st = kache_amd_device_alloc(want, &d_slot[sl]);
            if(st != KACHE_OK)
                return st;
            slot_bytes[sl] = want;
        }
        char* h = h_slot[sl];
        // The sequences' copies (each followed by '\n') are collected, then
        // made by the library's packing threads; the gaps that align the
        // regions and the final pad are separators too.
        std::vector<char*> dst;
        std::vector<const char*> src;
        std::vector<uint64_t> len;
        uint64_t filled = 0;
        for(Call& c : w.calls)
        {
            uint64_t at = c.at;
            if(at > filled)
                std::memset(h + filled, '\n', at - filled);
            const bool reads = c.kind == KACHE_AMD_REQ_COUNT || c.kind == KACHE_AMD_REQ_LOOKUP;
            if(c.kind == KACHE_AMD_REQ_INSERT)
                c.off.push_back(0);
            for(Req* q : c.reqs)
            {
                for(size_t j = 0; j < q->r.n; ++j)
                {
                    if(reads)
                        c.off.push_back(at - c.at);
                    dst.push_back(h + at), src.push_back(q->r.seqs[j]), len.push_back(q->r.lens[j]);
                    at += q->r.lens[j] + 1;
                }
                if(c.kind == KACHE_AMD_REQ_INSERT)
                    c.off.push_back(at - c.at);
            }
            if(reads)
                c.off.push_back(at - c.at);
            filled = at;
        }
        if(need > filled)
            std::memset(h + filled, '\n', need - filled);
        kamd_copy_spans(dst.data(), src.data(), len.data(), dst.size(), true);
        return kache_amd_copy(d_slot[sl], h, need, 1);
    }

    void stage_loop()
    {
        while(true)
        {
            auto w = std::make_unique<Staged>();
            {
                std::unique_lock<std::mutex> lock(mu);
                poll_wait(lock, cv_in, [&]{ return stopping || !pending.empty(); });
                if(pending.empty())
                {
                    stager_done = true;
                    cv_staged.notify_all();
                    return;
                }
                // A lone request reaching an idle server is not staged: the
                // executor runs it at once by the plain calls (whose own
                // pipeline copies it), one copy and one hand-off fewer.  One
                // larger than a slot is staged, through admission.
                if(pending.size() == 1 && staged.empty() && !exec_busy &&
                   pending.front()->bytes + KACHE_AMD_TEXT_PAD <= std::min(slot_bytes[0], slot_bytes[1]))
                {
                    w->direct = true;
                    w->reqs = take_wave(w->id);
                    stats.direct_waves++;
                    staged.push_back(std::move(w));
                    cv_staged.notify_all();
                    continue;
                }
                // A wave is admitted when a slot is free: requests arriving
                // meanwhile wait, and join it.
                poll_wait(lock, cv_slot, [&]{ return !slot_busy[0] || !slot_busy[1]; });
                w->slot = slot_busy[0] ? 1 : 0;
                slot_busy[w->slot] = true;
                w->reqs = take_wave(w->id);
                w->saturated = !pending.empty();
            }
            int st = KACHE_ENOMEM;
            try
            {
                st = pack(*w);
            }
            catch(...)
            {
                st = KACHE_ENOMEM;
            }
            if(st != KACHE_OK)
            {
                for(auto& q : w->reqs)
                    q->status = st;
                std::vector<Req*> all;
                for(auto& q : w->reqs)
                    all.push_back(q.get());
                finish(all);
                std::lock_guard<std::mutex> lock(mu);
                slot_busy[w->slot] = false;
                cv_slot.notify_all();
                continue;
            }
            std::lock_guard<std::mutex> lock(mu);
            staged.push_back(std::move(w));
            cv_staged.notify_all();
        }
    }

    void exec_loop()
    {
        while(true)
        {
            std::unique_ptr<Staged> w;
            {
                std::unique_lock<std::mutex> lock(mu);
                poll_wait(lock, cv_staged, [&]{ return !staged.empty() || stager_done; });
                if(staged.empty())
                    return;
                w = std::move(staged.front());
                staged.pop_front();
                exec_busy = true;
            }
            try
            {
                if(w->direct)
                {
                    const double t0 = now_s();
                    for(auto& q : w->reqs)
                        q->t_start = t0, q->wave = w->id, q->phase = phase_of(q->r.kind);
                    run(w->reqs);
                }
                else
                    execute(*w);
            }
            catch(...)
            {
                fail(w->reqs);
            }
            std::lock_guard<std::mutex> lock(mu);
            exec_busy = false;
            if(!w->direct)
            {
                slot_busy[w->slot] = false;
                cv_slot.notify_all();
            }
        }
    }

    void execute(Staged& w)
    {
        const double t0 = now_s();
        for(size_t k = 0; k < w.calls.size(); ++k)
            for(Req* q : w.calls[k].reqs)
                q->t_start = t0, q->wave = w.id, q->phase = static_cast<uint32_t>(k);
        const char* d = static_cast<const char*>(d_slot[w.slot]);
        for(const Call& c : w.calls)
        {
            const char* text = d + c.at;
            if(c.kind == KACHE_AMD_REQ_COUNT)
            {
                const size_t n = c.off.size() - 1;
                std::vector<uint64_t> out(std::max<size_t>(n, 1));
                const int st = kache_amd_count_seqs_device(t, text, c.off.back(), c.off.data(), n, out.data());
                size_t o = 0;
                for(Req* q : c.reqs)
                {
                    q->status = st;
                    if(st == KACHE_OK && q->r.n)
                        std::memcpy(q->r.counts, out.data() + o, q->r.n * sizeof(uint64_t));
                    o += q->r.n;
                }
            }
            else if(c.kind == KACHE_AMD_REQ_LOOKUP)
            {
                std::vector<uint64_t> first;
                std::vector<uint8_t*> found, fwd;
                std::vector<void*> values;
                uint64_t o = 0;
                for(Req* q : c.reqs)
                {
                    first.push_back(o), o += q->r.n;
                    found.push_back(q->r.found), values.push_back(q->r.values), fwd.push_back(q->r.fwd);
                }
                const int st = kache_amd_lookup_groups_device(t, text, c.off.back(), c.off.data(), c.off.size() - 1, first.data(), c.reqs.size(),
                                                              found.data(), values.data(), fwd.data());
                for(Req* q : c.reqs)
                    q->status = st;
            }
            else if(c.kind == KACHE_AMD_REQ_UPDATE)
            {
                Req* q = c.reqs[0];
                q->status = kache_amd_update_device(t, text, c.bytes, q->r.op, q->r.x, q->r.counts);
            }
            else
            {
                std::vector<uint64_t> out(c.reqs.size());
                const int st = kache_amd_insert_batches(t, user, text, c.off.data(), c.reqs.size(), c.op, c.x, out.data());
                for(size_t i = 0; i < c.reqs.size(); ++i)
                {
                    c.reqs[i]->status = st;
                    if(st == KACHE_OK && c.reqs[i]->r.counts)
                        *c.reqs[i]->r.counts = out[i];
                }
            }
            finish(c.reqs);
        }
    }
};

extern "C"
{

int kache_amd_server_new(kache_table* t, uint32_t user, uint64_t wave_bytes, uint32_t max_wave_requests, kache_amd_server** out) noexcept
{
    if(t == nullptr || out == nullptr || wave_bytes == 0 || max_wave_requests == 0)
        return KACHE_EINVAL;
    kache_amd_info info{};
    const int st = kache_amd_get_info(t, &info);
    if(st != KACHE_OK)
        return st;
    try
    {
        auto* s = new kache_amd_server(t, user, wave_bytes, max_wave_requests, static_cast<uint32_t>(info.value_bytes));
        const int started = s->start();
        if(started != KACHE_OK)
        {
            delete s;
            return started;
        }
        *out = s;
        return KACHE_OK;
    }
    catch(...)
    {
        return KACHE_EINTERNAL;
    }
}

void kache_amd_server_free(kache_amd_server* s) noexcept
{
    delete s;
}

int kache_amd_server_submit(kache_amd_server* s, const kache_amd_request* r, uint64_t* ticket) noexcept
{
    if(s == nullptr || r == nullptr || ticket == nullptr || r->kind > KACHE_AMD_REQ_INSERT || (r->n > 0 && (!r->seqs || !r->lens)))
        return KACHE_EINVAL;
    if((r->kind == KACHE_AMD_REQ_COUNT && r->n > 0 && !r->counts) || (r->kind == KACHE_AMD_REQ_LOOKUP && !r->found))
        return KACHE_EINVAL;
    for(size_t i = 0; i < r->n; ++i)
        if(r->seqs[i] == nullptr && r->lens[i] > 0)
            return KACHE_EINVAL;
    try
    {
        auto q = std::make_shared<Req>();
        q->r = *r;
        q->bytes = staged_bytes(*r);
        q->t_submit = now_s();
        {
            std::lock_guard<std::mutex> lock(s->mu);
            q->ticket = s->next_ticket++;
            s->live[q->ticket] = q;
            s->pending.push_back(q);
        }
        s->cv_in.notify_one();
        *ticket = q->ticket;
        return KACHE_OK;
    }
    catch(...)
    {
        return KACHE_EINTERNAL;
    }
}

int kache_amd_server_wait(kache_amd_server* s, uint64_t ticket, kache_amd_request_result* res) noexcept
{
    if(s == nullptr)
        return KACHE_EINVAL;
    std::unique_lock<std::mutex> lock(s->mu);
    const auto it = s->live.find(ticket);
    if(it == s->live.end())
        return KACHE_EINVAL;
    auto q = it->second;
    s->cv_done.wait(lock, [&]{ return q->done; });
    s->live.erase(ticket);
    if(res)
    {
        res->status = q->status;
        res->phase = q->phase;
        res->ticket = q->ticket;
        res->wave = q->wave;
        res->submit_s = q->t_submit;
        res->start_s = q->t_start;
        res->done_s = q->t_done;
    }
    return KACHE_OK;
}

int kache_amd_server_set_hbm_budget(kache_amd_server* s, uint64_t bytes) noexcept
{
    if(s == nullptr)
        return KACHE_EINVAL;
    s->hbm_budget.store(bytes);
    return KACHE_OK;
}

int kache_amd_server_get_stats(kache_amd_server* s, kache_amd_server_stats* out) noexcept
{
    if(s == nullptr || out == nullptr)
        return KACHE_EINVAL;
    std::lock_guard<std::mutex> lock(s->mu);
    *out = s->stats;
    return KACHE_OK;
}

int kache_amd_server_replay(kache_amd_server* s, const kache_amd_request* reqs, const double* arrival_s, size_t n,
                            kache_amd_request_result* results) noexcept
{
    if(s == nullptr || (n > 0 && (reqs == nullptr || arrival_s == nullptr || results == nullptr)))
        return KACHE_EINVAL;
    for(size_t i = 1; i < n; ++i)
        if(!(arrival_s[i] >= arrival_s[i - 1]))
            return KACHE_EINVAL;
    try
    {
        std::vector<uint64_t> tickets(n);
        std::vector<int> wait_status(n, KACHE_OK);
        std::mutex m;
        std::condition_variable cv;
        size_t submitted = 0;
        bool failed = false;
        std::thread waiter([&]
        {
            for(size_t i = 0; i < n; ++i)
            {
                uint64_t ticket = 0;
                {
                    std::unique_lock<std::mutex> lock(m);
                    cv.wait(lock, [&]{ return submitted > i || failed; });
                    if(submitted <= i)
                        return;
                    ticket = tickets[i];
                }
                wait_status[i] = kache_amd_server_wait(s, ticket, results + i);
            }
        });
        const double t0 = now_s();
        int st = KACHE_OK;
        for(size_t i = 0; i < n && st == KACHE_OK; ++i)
        {
            // Sleep to within 200 us of the arrival, then spin: timer wake-ups are coarse.
            const double due = t0 + arrival_s[i];
            for(double left = due - now_s(); left > 0; left = due - now_s())
                if(left > 200e-6)
                    std::this_thread::sleep_for(std::chrono::duration<double>(left - 150e-6));
            uint64_t ticket = 0;
            st = kache_amd_server_submit(s, reqs + i, &ticket);
            std::lock_guard<std::mutex> lock(m);
            if(st == KACHE_OK)
                tickets[i] = ticket, submitted = i + 1;
            else
                failed = true;
            cv.notify_one();
        }
        waiter.join();
        if(st != KACHE_OK)
            return st;
        for(size_t i = 0; i < n; ++i)
            if(wait_status[i] != KACHE_OK)
                return wait_status[i];
        return KACHE_OK;
    }
    catch(...)
    {
        return KACHE_EINTERNAL;
    }
}

double kache_amd_clock(void) noexcept
{
    return now_s();
}

}
