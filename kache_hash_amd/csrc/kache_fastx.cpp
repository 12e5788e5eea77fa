// FASTA / FASTQ input for the k-mer table (kache_amd_insert_fastx,
// kache_amd_update_fastx in include/kache_amd.h), with the reference reader's
// format rules (counter/Fastx_Reader.hpp:17-22, counter/Fastx_Reader.cpp:36-48,
// 178-241): the first byte sets the format ('>' FASTA, '@' FASTQ; gzip input
// is refused); a FASTA sequence joins its lines up to the next header; a
// FASTQ record has four lines and a quality line as long as its sequence;
// lines end with "\n" or "\r\n"; empty sequences are skipped.
//
// The file is read in blocks and its sequences are written, each followed by
// a newline, into a host text of about BLOCK_BYTES (a longer record makes its
// block longer), which the table's double-buffered text pipeline then
// streams to the device.

#include "kache_amd.h"

#include <cerrno>
#include <cstring>
#include <fcntl.h>
#include <stdexcept>
#include <string>
#include <unistd.h>
#include <vector>

void kamd_set_last_error(const std::string& what);

namespace
{

constexpr size_t READ_BYTES = size_t(16) << 20;
constexpr size_t BLOCK_BYTES = size_t(64) << 20;

struct Malformed : std::runtime_error
{
    using std::runtime_error::runtime_error;
};

class Fastx_Text
{
    int fd = -1;
    std::string path;
    std::vector<char> in;
    size_t pos = 0, len = 0;
    uint64_t consumed = 0;      // input bytes before in[0]
    bool eof = false;
    bool fastq = false;
    bool started = false;

    // Makes in[pos..] hold a whole line (up to its '\n' or the end of the
    // input); returns its content [b, e) without the line end, and moves
    // past it.  False at the end of the input.
    bool line(const char*& b, const char*& e)
    {
        for(;;)
        {
            const char* nl = pos < len ? static_cast<const char*>(std::memchr(in.data() + pos, '\n', len - pos)) : nullptr;
            if(nl || eof)
            {
                if(pos >= len)
                    return false;
                b = in.data() + pos;
                const char* end = nl ? nl : in.data() + len;
                pos = nl ? static_cast<size_t>(nl - in.data()) + 1 : len;
                e = (end > b && end[-1] == '\r') ? end - 1 : end;
                return true;
            }
            refill();
        }
    }

    // Peeks at the next byte (-1 at the end of the input).
    int peek()
    {
        while(pos >= len && !eof)
            refill();
        return pos < len ? static_cast<unsigned char>(in[pos]) : -1;
    }

    void refill()
    {
        if(pos > 0)
        {
            std::memmove(in.data(), in.data() + pos, len - pos);
            consumed += pos, len -= pos, pos = 0;
        }
        if(in.size() < len + READ_BYTES)
            in.resize(len + READ_BYTES);
        const ssize_t r = ::read(fd, in.data() + len, READ_BYTES);
        if(r < 0)
        {
            if(errno == EINTR)
                return;
            throw std::runtime_error("cannot read the input: " + std::string(std::strerror(errno)));
        }
        if(r == 0)
            eof = true;
        len += static_cast<size_t>(r);
    }

    [[noreturn]] void malformed(const char* what) const
    {
        throw Malformed(std::string("malformed ") + (fastq ? "FASTQ" : "FASTA") + " record at byte " +
                        std::to_string(consumed + pos) + ": " + what);
    }

public:
    explicit Fastx_Text(const char* p) : path(p)
    {
        fd = ::open(p, O_RDONLY);
        if(fd < 0)
            throw std::invalid_argument("cannot open " + path + ": " + std::strerror(errno));
    }

    ~Fastx_Text()
    {
        if(fd >= 0)
            ::close(fd);
    }

    // Appends the next records' sequences to `out` (cleared first), each
    // followed by '\n', until it holds BLOCK_BYTES or the input ends; false
    // when nothing is left.
    bool next(std::string& out)
    {
        out.clear();
        if(!started)
        {
            started = true;
            const int c = peek();
            if(c < 0)
                return false;
            if(len >= 2 && static_cast<unsigned char>(in[pos]) == 0x1f && static_cast<unsigned char>(in[pos + 1]) == 0x8b)
                throw std::invalid_argument(path + " is gzip-compressed; decompress it, e.g. into a pipe");
            if(c != '>' && c != '@')
                throw std::invalid_argument(path + " is neither FASTA nor FASTQ: it starts with neither '>' nor '@'");
            fastq = (c == '@');
        }
        const char *b, *e;
        while(out.size() < BLOCK_BYTES)
        {
            const int c = peek();
            if(c < 0)
                break;
            if(c != (fastq ? '@' : '>'))
                malformed("expected a header");
            line(b, e);                                         // the header
            if(!fastq)
            {
                // Lines up to the next header, joined.
                const size_t start = out.size();
                while(peek() >= 0 && peek() != '>')
                {
                    line(b, e);
                    out.append(b, e);
                }
                if(out.size() > start)
                    out.push_back('\n');
                continue;
            }
            if(!line(b, e))
                malformed("expected a '+' line");
            const std::string seq(b, e);
            if(peek() != '+')
                malformed("expected a '+' line");
            line(b, e);
            const char *qb = nullptr, *qe = nullptr;
            if(!line(qb, qe))
                qb = qe = nullptr;
            if(static_cast<size_t>(qe - qb) != seq.size())
                malformed("its quality and sequence lengths differ");
            if(!seq.empty())
                out.append(seq).push_back('\n');
        }
        return !out.empty() || peek() >= 0;
    }
};

template <typename Call>
int run(const char* path, const Call& call) noexcept
{
    try
    {
        Fastx_Text f(path);
        std::string block;
        while(f.next(block))
        {
            if(block.empty())
                continue;
            const int st = call(block.data(), block.size());
            if(st != KACHE_OK)
                return st;
        }
        return KACHE_OK;
    }
    catch(const Malformed& e)
    {
        kamd_set_last_error(e.what());
        return KACHE_EINVAL;
    }
    catch(const std::invalid_argument& e)
    {
        kamd_set_last_error(e.what());
        return KACHE_EINVAL;
    }
    catch(const std::bad_alloc&)
    {
        kamd_set_last_error("out of memory");
        return KACHE_ENOMEM;
    }
    catch(const std::exception& e)
    {
        kamd_set_last_error(e.what());
        return KACHE_EINTERNAL;
    }
}

}

extern "C"
{

int kache_amd_read_fastx(const char* path, kache_amd_text_fn fn, void* ctx) noexcept
{
    if(path == nullptr || fn == nullptr)
        return KACHE_EINVAL;
    return run(path, [&](const char* text, uint64_t bytes){ return fn(ctx, text, bytes) == 0 ? static_cast<int>(KACHE_OK) : static_cast<int>(KACHE_EINVAL); });
}

int kache_amd_insert_fastx(kache_table* t, uint32_t user, const char* path, uint32_t op, uint64_t x, uint64_t* inserted) noexcept
{
    if(t == nullptr || path == nullptr)
        return KACHE_EINVAL;
    // Validate the call once, before reading anything.
    const int v = kache_amd_insert_text(t, user, nullptr, 0, op, x, nullptr);
    if(v != KACHE_OK)
        return v;
    uint64_t total = 0;
    const int st = run(path, [&](const char* text, uint64_t bytes)
    {
        uint64_t n = 0;
        const int s = kache_amd_insert_text(t, user, text, bytes, op, x, &n);
        total += n;
        return s;
    });
    if(inserted != nullptr)
        *inserted = total;
    return st;
}

int kache_amd_update_fastx(kache_table* t, const char* path, uint32_t op, uint64_t x, uint64_t* present) noexcept
{
    if(t == nullptr || path == nullptr)
        return KACHE_EINVAL;
    const int v = kache_amd_update_text(t, nullptr, 0, op, x, nullptr);
    if(v != KACHE_OK)
        return v;
    uint64_t total = 0;
    const int st = run(path, [&](const char* text, uint64_t bytes)
    {
        uint64_t n = 0;
        const int s = kache_amd_update_text(t, text, bytes, op, x, &n);
        total += n;
        return s;
    });
    if(present != nullptr)
        *present = total;
    return st;
}

}
