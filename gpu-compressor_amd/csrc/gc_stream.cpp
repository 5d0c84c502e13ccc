// gc_stream.cpp -- streaming NVZS / NVLZ reader: container file in, original
// bytes out, frames decoded by a pool of threads while the next ones are read
// and the finished ones written in order.
//
// decompress.rs:9-143 reads the header and size table, then for each frame
// reads it, decodes it (NVLZ with liblzma, NVZS with nvCOMP on the GPU,
// :146-365) and writes it before reading the next.  Here the same file layout
// (magic, original size, chunk size, frame count, frame sizes, frames;
// pipeline.rs:187-260, pipeline_lzma2.rs:339-400) is read by one thread,
// decoded by `threads` workers (libzstd ZSTD_decompress / liblzma raw LZMA2,
// which give the same bytes as any conforming decoder) and written by the
// caller's thread, with at most 2 x threads frames in memory.
#include <stdio.h>
#include <string.h>
#include <sys/stat.h>

#include <algorithm>
#include <new>
#include <condition_variable>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/gcompress.h"
#include "gc_runtime.h"

namespace gc {
namespace {

uint64_t le64(const uint8_t* p) {
    uint64_t v = 0;
    for (int i = 0; i < 8; i++) v |= (uint64_t)p[i] << (8 * i);
    return v;
}

struct File {
    FILE* f = nullptr;
    ~File() {
        if (f) fclose(f);
    }
};

enum SlotState { kEmpty, kRead, kDecoding, kDecoded };

struct Slot {
    SlotState state = kEmpty;
    uint64_t index = 0;
    std::vector<uint8_t> in, out;
};

}  // namespace
}  // namespace gc

using namespace gc;

extern "C" int gc_decompress_file(const char* in_path, const char* out_path, int threads, uint64_t* out_bytes) {
    if (!in_path || !out_path) return fail(GC_EINVAL, "null argument");
    if (threads <= 0) threads = (int)std::min<unsigned>(std::max(1u, std::thread::hardware_concurrency()), 8u);
    File in, out;
    in.f = fopen(in_path, "rb");
    if (!in.f) return fail(GC_EIO, std::string("cannot open ") + in_path);
    uint8_t head[28];
    if (fread(head, 1, 28, in.f) != 28) return fail(GC_EFORMAT, "container shorter than its 28-byte header");
    if (memcmp(head, "NVZS", 4) && memcmp(head, "NVLZ", 4))
        return fail(GC_EFORMAT, "Invalid file format (expected NVZS or NVLZ)");  // decompress.rs:29-31
    const uint32_t algo = memcmp(head, "NVZS", 4) ? GC_ALGO_LZMA2 : GC_ALGO_ZSTD;
    const uint64_t orig = le64(head + 4), ch = le64(head + 12), nch = le64(head + 20);
    if (ch == 0 ? orig != 0 || nch != 0 : nch != (orig + ch - 1) / ch)
        return fail(GC_EFORMAT, "entry count does not match original and chunk sizes");
    // Sizes come from the file and are bounded before anything is allocated:
    // a chunk larger than 4 GiB is no container this library or the
    // reference writes, and every entry needs 8 bytes of the size table.
    struct stat st;
    if (fstat(fileno(in.f), &st) != 0) return fail(GC_EIO, std::string("cannot stat ") + in_path);
    const uint64_t file_size = (uint64_t)st.st_size;
    if (ch > (4ull << 30)) return fail(GC_EFORMAT, "chunk size out of range (above 4 GiB)");
    if (nch > (file_size - 28) / 8) return fail(GC_EFORMAT, "truncated size table");
    std::vector<uint64_t> csize(nch);
    {
        std::vector<uint8_t> table(8 * nch);
        if (nch && fread(table.data(), 1, table.size(), in.f) != table.size())
            return fail(GC_EFORMAT, "truncated size table");
        for (uint64_t c = 0; c < nch; c++) csize[c] = le64(table.data() + 8 * c);
    }
    out.f = fopen(out_path, "wb");
    if (!out.f) return fail(GC_EIO, std::string("cannot create ") + out_path);

    // frames in memory: 2 per decoding thread, fewer for dictionary-size
    // chunks (about 1 GiB of decoded frames at most)
    const size_t window = std::max<size_t>(2, std::min<size_t>(2 * (size_t)threads, (1ull << 30) / std::max<uint64_t>(ch, 1)));
    // a frame is never larger than its decoded size plus the raw-chunk or
    // compressBound overhead; anything larger is a damaged size table
    const uint64_t max_frame = ch + ch / 16 + (64u << 10);
    std::vector<Slot> slots(window);
    std::mutex mu;
    std::condition_variable cv;
    int err = GC_OK;
    std::string err_msg;
    uint64_t next_decode = 0;
    auto set_err = [&](int code, const std::string& m) {  // with mu held
        if (err == GC_OK) err = code, err_msg = m;
        cv.notify_all();
    };
    auto dsize_of = [&](uint64_t c) { return c + 1 < nch ? ch : orig - (nch - 1) * ch; };  // decompress.rs:70-73

    std::thread reader([&] {
        for (uint64_t c = 0; c < nch; c++) {
            Slot& s = slots[c % window];
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return err != GC_OK || s.state == kEmpty; });
                if (err != GC_OK) return;
            }
            if (csize[c] > max_frame) {
                std::lock_guard<std::mutex> lk(mu);
                return set_err(GC_EFORMAT, "a frame size is out of range");
            }
            try {
                s.in.resize(csize[c]);
            } catch (const std::bad_alloc&) {
                std::lock_guard<std::mutex> lk(mu);
                return set_err(GC_ENOMEM, "no memory for a frame");
            }
            const bool ok = fread(s.in.data(), 1, csize[c], in.f) == csize[c];
            std::lock_guard<std::mutex> lk(mu);
            if (!ok) return set_err(GC_EFORMAT, "truncated payload");  // read_exact, decompress.rs:81
            s.index = c;
            s.state = kRead;
            cv.notify_all();
        }
    });
    std::vector<std::thread> workers;
    for (int t = 0; t < threads; t++)
        workers.emplace_back([&] {
            for (;;) {
                uint64_t c;
                {
                    std::unique_lock<std::mutex> lk(mu);
                    cv.wait(lk, [&] {
                        return err != GC_OK || next_decode >= nch ||
                               (slots[next_decode % window].state == kRead &&
                                slots[next_decode % window].index == next_decode);
                    });
                    if (err != GC_OK || next_decode >= nch) return;
                    c = next_decode++;
                    slots[c % window].state = kDecoding;
                }
                Slot& s = slots[c % window];
                const uint64_t d = dsize_of(c);
                try {
                    s.out.resize(d);
                } catch (const std::bad_alloc&) {
                    std::lock_guard<std::mutex> lk(mu);
                    return set_err(GC_ENOMEM, "no memory for a decoded frame");
                }
                size_t got = 0;
                const int rc = algo == GC_ALGO_LZMA2
                                   ? gc_lzma2_raw_decode(s.in.data(), s.in.size(), s.out.data(), d, &got)
                                   : gc_zstd_frame_decode(s.in.data(), s.in.size(), s.out.data(), d, &got);
                const std::string m = rc ? gc_last_error() : std::string();
                std::lock_guard<std::mutex> lk(mu);
                if (rc) return set_err(rc, m);
                if (got != d) return set_err(GC_EFORMAT, "an entry decoded to the wrong size");
                s.state = kDecoded;
                cv.notify_all();
            }
        });
    uint64_t written = 0;
    for (uint64_t c = 0; c < nch; c++) {
        Slot& s = slots[c % window];
        {
            std::unique_lock<std::mutex> lk(mu);
            cv.wait(lk, [&] { return err != GC_OK || (s.state == kDecoded && s.index == c); });
            if (err != GC_OK) break;
        }
        const bool ok = fwrite(s.out.data(), 1, s.out.size(), out.f) == s.out.size();
        std::lock_guard<std::mutex> lk(mu);
        if (!ok) {
            set_err(GC_EIO, std::string("write to ") + out_path + " failed");
            break;
        }
        written += s.out.size();
        s.state = kEmpty;
        cv.notify_all();
    }
    {
        std::lock_guard<std::mutex> lk(mu);
        if (err == GC_OK && written != orig) set_err(GC_EFORMAT, "decoded size differs from the original size");
        cv.notify_all();
    }
    reader.join();
    for (auto& w : workers) w.join();
    if (err == GC_OK && fflush(out.f) != 0) set_err(GC_EIO, std::string("write to ") + out_path + " failed");
    if (err != GC_OK) return fail(err, err_msg);
    if (out_bytes) *out_bytes = written;
    return GC_OK;
}
