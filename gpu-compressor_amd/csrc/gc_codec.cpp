// gc_codec.cpp -- consumers of the match lists and the NVZS / NVLZ containers.
//
//  * Zstd: sequences -> libzstd ZSTD_compressSequences, one frame per 64 KiB
//    sub-chunk, ZSTD_compress(level 1) when libzstd rejects the sequences
//    (compress_zstd_custom.rs:120-199).
//  * LZMA2: triage of a sub-block on "any match found" -> liblzma raw LZMA2 at
//    preset 9, else raw LZMA2 chunks and an end marker
//    (compress_lzma2_custom.rs:325-338,570-597; compress_lzma2.rs:89-133).
//  * Containers: magic[4], original size u64 LE, chunk size u64 LE, entry
//    count u64 LE, entry sizes u64 LE, payload (pipeline.rs:193-253,
//    pipeline_lzma2.rs:339-400); reader after decompress.rs:35-143, with
//    libzstd ZSTD_decompress in place of nvCOMP.
//
// libzstd and liblzma are loaded with dlopen; this image has their shared
// objects but no headers, so the few types used are declared here after the
// libraries' documented ABI (libzstd 1.4.x, liblzma 5.2.x).
#include <dlfcn.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/gcompress.h"
#include "gc_runtime.h"

using namespace gc;

namespace {

// ---- libzstd -------------------------------------------------------------
struct ZstdLib {
    void* h = nullptr;
    void* (*createCCtx)() = nullptr;
    size_t (*freeCCtx)(void*) = nullptr;
    size_t (*reset)(void*, int) = nullptr;
    size_t (*setParameter)(void*, int, int) = nullptr;
    size_t (*compressSequences)(void*, void*, size_t, const gc_zstd_sequence*, size_t, const void*, size_t) = nullptr;
    size_t (*compressBound)(size_t) = nullptr;
    size_t (*compress)(void*, size_t, const void*, size_t, int) = nullptr;
    size_t (*decompress)(void*, size_t, const void*, size_t) = nullptr;
    unsigned (*isError)(size_t) = nullptr;
    const char* (*errorName)(size_t) = nullptr;
    bool ok = false;
};

ZstdLib& zstd() {
    static ZstdLib z;
    static std::once_flag once;
    std::call_once(once, [] {
        z.h = dlopen("libzstd.so.1", RTLD_NOW);
        if (!z.h) return;
        bool all = true;
        auto get = [&](auto& f, const char* name) {
            f = reinterpret_cast<std::remove_reference_t<decltype(f)>>(dlsym(z.h, name));
            all = all && f != nullptr;
        };
        get(z.createCCtx, "ZSTD_createCCtx");
        get(z.freeCCtx, "ZSTD_freeCCtx");
        get(z.reset, "ZSTD_CCtx_reset");
        get(z.setParameter, "ZSTD_CCtx_setParameter");
        get(z.compressSequences, "ZSTD_compressSequences");
        get(z.compressBound, "ZSTD_compressBound");
        get(z.compress, "ZSTD_compress");
        get(z.decompress, "ZSTD_decompress");
        get(z.isError, "ZSTD_isError");
        get(z.errorName, "ZSTD_getErrorName");
        z.ok = all;
    });
    return z;
}

// ---- liblzma -------------------------------------------------------------
struct LzmaStream {  // lzma_stream (136 bytes on LP64)
    const uint8_t* next_in;
    size_t avail_in;
    uint64_t total_in;
    uint8_t* next_out;
    size_t avail_out;
    uint64_t total_out;
    uint8_t rest[88];
};
struct LzmaFilter {
    uint64_t id;
    void* options;
};
const uint64_t kFilterLzma2 = 0x21;
const uint64_t kVliUnknown = ~0ull;
const int kLzmaRun = 0, kLzmaFinish = 3;
const int kLzmaOk = 0, kLzmaStreamEnd = 1;

struct LzmaLib {
    void* h = nullptr;
    int (*raw_encoder)(LzmaStream*, const LzmaFilter*) = nullptr;
    int (*raw_decoder)(LzmaStream*, const LzmaFilter*) = nullptr;
    int (*code)(LzmaStream*, int) = nullptr;
    void (*end)(LzmaStream*) = nullptr;
    unsigned char (*preset)(void*, uint32_t) = nullptr;
    bool ok = false;
};

LzmaLib& lzma() {
    static LzmaLib l;
    static std::once_flag once;
    std::call_once(once, [] {
        l.h = dlopen("liblzma.so.5", RTLD_NOW);
        if (!l.h) return;
        bool all = true;
        auto get = [&](auto& f, const char* name) {
            f = reinterpret_cast<std::remove_reference_t<decltype(f)>>(dlsym(l.h, name));
            all = all && f != nullptr;
        };
        get(l.raw_encoder, "lzma_raw_encoder");
        get(l.raw_decoder, "lzma_raw_decoder");
        get(l.code, "lzma_code");
        get(l.end, "lzma_end");
        get(l.preset, "lzma_lzma_preset");
        l.ok = all;
    });
    return l;
}

// lzma_options_lzma is 112 bytes in liblzma 5.x; dict_size is its first field.
struct LzmaOptions {
    uint32_t dict_size;
    uint8_t rest[252];
};

const uint32_t kMaxSeq = GC_SUB_BLOCK / 3;

// Uninitialised host buffer for packed sequences (worst case: every slot).
struct SeqBuf {
    gc_zstd_sequence* p;
    size_t cap;
    explicit SeqBuf(size_t n) : p((gc_zstd_sequence*)malloc((n ? n : 1) * sizeof(gc_zstd_sequence))), cap(n) {}
    ~SeqBuf() { free(p); }
    SeqBuf(const SeqBuf&) = delete;
    SeqBuf& operator=(const SeqBuf&) = delete;
};

uint64_t n_sub_of(uint64_t n) { return (n + GC_SUB_BLOCK - 1) / GC_SUB_BLOCK; }

void put_u64(uint8_t* p, uint64_t v) {
    for (int i = 0; i < 8; i++) p[i] = (uint8_t)(v >> (8 * i));
}
uint64_t get_u64(const uint8_t* p) {
    uint64_t v = 0;
    for (int i = 0; i < 8; i++) v |= (uint64_t)p[i] << (8 * i);
    return v;
}

// Raw LZMA2 chunks of at most 64 KiB, then the end marker
// (compress_lzma2_custom.rs:325-338 plus the 0x00 pushed by its callers).
size_t raw_lzma2(const uint8_t* src, size_t n, uint8_t* out) {
    size_t o = 0;
    bool first = true;
    for (size_t off = 0; off < n;) {
        size_t len = std::min<size_t>(n - off, 65536);
        out[o++] = first ? 0x01 : 0x02;
        uint16_t us = (uint16_t)(len - 1);
        out[o++] = (uint8_t)(us >> 8);
        out[o++] = (uint8_t)us;
        memcpy(out + o, src + off, len);
        o += len;
        off += len;
        first = false;
    }
    out[o++] = 0x00;
    return o;
}

size_t raw_lzma2_size(size_t n) { return n + 3 * ((n + 65535) / 65536) + 1; }

// liblzma raw LZMA2 encoder at `preset` (compress_lzma2.rs:89-133).
int lzma2_encode(const uint8_t* src, size_t n, uint32_t preset, uint8_t* out, size_t cap, size_t* out_len) {
    LzmaLib& L = lzma();
    if (!L.ok) return fail(GC_ELIB, "liblzma.so.5 could not be loaded");
    LzmaOptions opts;
    memset(&opts, 0, sizeof(opts));
    if (L.preset(&opts, std::min<uint32_t>(preset, 9))) return fail(GC_ELIB, "lzma_lzma_preset failed");
    LzmaFilter filters[2] = {{kFilterLzma2, &opts}, {kVliUnknown, nullptr}};
    LzmaStream s;
    memset(&s, 0, sizeof(s));
    if (L.raw_encoder(&s, filters) != kLzmaOk) return fail(GC_ELIB, "lzma_raw_encoder init failed");
    s.next_in = src;
    s.avail_in = n;
    s.next_out = out;
    s.avail_out = cap;
    int ret;
    do {
        ret = L.code(&s, kLzmaFinish);
    } while (ret == kLzmaOk && s.avail_out > 0);
    size_t produced = (size_t)s.total_out;
    L.end(&s);
    if (ret != kLzmaStreamEnd) return fail(ret == kLzmaOk ? GC_ESIZE : GC_ELIB, "lzma_code(FINISH) did not finish");
    *out_len = produced;
    return GC_OK;
}

// One sub-block: liblzma preset 9 when the match finder found anything,
// raw chunks otherwise or when liblzma fails (compress_lzma2_custom.rs:570-597).
int encode_sub_block(const uint8_t* sub, size_t n, bool has_matches, uint8_t* out, size_t cap, size_t* out_len) {
    if (has_matches) {
        size_t len = 0;
        int rc = lzma2_encode(sub, n, 9, out, cap, &len);
        if (rc == GC_OK) {
            *out_len = len;
            return GC_OK;
        }
        if (rc == GC_ELIB && !lzma().ok) return rc;
    }
    if (cap < raw_lzma2_size(n)) return fail(GC_ESIZE, "output buffer too small for raw LZMA2 chunks");
    *out_len = raw_lzma2(sub, n, out);
    return GC_OK;
}

size_t lzma2_sub_bound(size_t n) { return n + n / 8 + 1024; }  // compress_lzma2.rs:113

// Parallel-for over [0, count) on `threads` threads.
template <class F>
void parallel_for(size_t count, int threads, F f) {
    if (threads <= 1 || count <= 1) {
        for (size_t i = 0; i < count; i++) f(i);
        return;
    }
    std::atomic<size_t> next{0};
    std::vector<std::thread> pool;
    int t = (int)std::min<size_t>((size_t)threads, count);
    for (int k = 0; k < t; k++)
        pool.emplace_back([&] {
            for (;;) {
                size_t i = next.fetch_add(1);
                if (i >= count) break;
                f(i);
            }
        });
    for (auto& th : pool) th.join();
}

int default_threads() {
    unsigned c = std::thread::hardware_concurrency();
    return (int)std::min<unsigned>(c ? c : 1, 8);  // min(cores, 8), pipeline_lzma2.rs:226
}

}  // namespace

extern "C" {

size_t gc_zstd_custom_bound(size_t n) {
    ZstdLib& z = zstd();
    if (!z.ok) {
        set_error("libzstd.so.1 could not be loaded");
        return 0;
    }
    return (size_t)n_sub_of(n) * z.compressBound(GC_SUB_BLOCK);
}

int gc_zstd_encode_sequences(const uint8_t* src, size_t n, const gc_zstd_sequence* seqs, const uint32_t* seq_counts,
                             uint8_t* out, size_t cap, uint64_t* sizes, size_t* out_len, int threads) {
    if ((!src && n) || !seq_counts || !out || !sizes || !out_len) return fail(GC_EINVAL, "null argument");
    ZstdLib& z = zstd();
    if (!z.ok) return fail(GC_ELIB, "libzstd.so.1 could not be loaded");
    const size_t nsub = (size_t)n_sub_of(n);
    const size_t bound = z.compressBound(GC_SUB_BLOCK);
    if (cap < nsub * bound) return fail(GC_ESIZE, "output capacity below gc_zstd_custom_bound(n)");
    // sequence offsets of each sub-chunk in `seqs` (packed back to back)
    std::vector<uint64_t> first(nsub + 1, 0);
    for (size_t s = 0; s < nsub; s++) first[s + 1] = first[s] + seq_counts[s];
    std::vector<int> err(nsub, 0);
    std::vector<std::string> msg(nsub);
    // frames are produced in slots of `bound` bytes, then compacted
    if (threads < 1) threads = 1;
    std::vector<void*> cctx((size_t)threads, nullptr);
    std::atomic<int> tid{0};
    auto work = [&](size_t lo, size_t hi, void* cc) {
        for (size_t s = lo; s < hi; s++) {
            const size_t start = s * GC_SUB_BLOCK;
            const size_t len = std::min<size_t>(GC_SUB_BLOCK, n - start);
            uint8_t* dst = out + s * bound;
            z.reset(cc, 1);  // ZSTD_reset_session_only (compress_zstd_custom.rs:161)
            size_t r = z.compressSequences(cc, dst, bound, seqs + first[s], seq_counts[s], src + start, len);
            if (z.isError(r)) {
                // :173-190: fall back to a plain level-1 frame
                size_t f = z.compress(dst, bound, src + start, len, 1);
                if (z.isError(f)) {
                    err[s] = 1;
                    msg[s] = std::string("sub-chunk ") + std::to_string(s) + " compression failed: " + z.errorName(r);
                    continue;
                }
                r = f;
            }
            sizes[s] = r;
        }
    };
    const size_t per = (nsub + threads - 1) / threads;
    auto run_range = [&](int k) {
        void* cc = z.createCCtx();
        if (!cc) return;
        // The reference sets parameter 1001 to 0 (commented as block delimiters; in
        // libzstd 1.4.x it is ZSTD_c_forceAttachDict, already 0): kept for identical CCtx state.
        z.setParameter(cc, 1001, 0);
        cctx[(size_t)k] = cc;
        size_t lo = (size_t)k * per, hi = std::min(nsub, lo + per);
        if (lo < hi) work(lo, hi, cc);
        z.freeCCtx(cc);
    };
    (void)tid;
    if (threads == 1) {
        run_range(0);
        if (!cctx[0]) return fail(GC_ENOMEM, "ZSTD_createCCtx failed");
    } else {
        std::vector<std::thread> pool;
        for (int k = 0; k < threads; k++) pool.emplace_back(run_range, k);
        for (auto& t : pool) t.join();
        for (int k = 0; k < threads; k++)
            if (!cctx[(size_t)k] && (size_t)k * per < nsub) return fail(GC_ENOMEM, "ZSTD_createCCtx failed");
    }
    for (size_t s = 0; s < nsub; s++)
        if (err[s]) return fail(GC_ELIB, msg[s]);
    size_t o = 0;
    for (size_t s = 0; s < nsub; s++) {
        if (o != s * bound) memmove(out + o, out + s * bound, sizes[s]);
        o += sizes[s];
    }
    *out_len = o;
    return GC_OK;
}

int gc_compress_chunk_zstd_custom(gc_ctx* ctx, const uint8_t* chunk, size_t n, uint32_t level, uint8_t* out,
                                  size_t cap, uint64_t* sizes, size_t* out_len) {
    if (!ctx || (!chunk && n) || !out || !sizes || !out_len) return fail(GC_EINVAL, "null argument");
    if (n == 0) {
        *out_len = 0;
        return GC_OK;
    }
    const size_t nsub = (size_t)n_sub_of(n);
    std::vector<uint32_t> counts(nsub);
    SeqBuf seqs(nsub * kMaxSeq);
    if (!seqs.p) return fail(GC_ENOMEM, "sequence buffer");
    size_t total = 0;
    int rc = gc_zstd_find_sequences_packed(ctx, chunk, n, level, seqs.p, seqs.cap, counts.data(), &total);
    if (rc) return rc;
    return gc_zstd_encode_sequences(chunk, n, seqs.p, counts.data(), out, cap, sizes, out_len, 1);
}

int gc_zstd_frame_decode(const uint8_t* src, size_t n, uint8_t* out, size_t expected, size_t* out_len) {
    if ((!src && n) || (!out && expected) || !out_len) return fail(GC_EINVAL, "null argument");
    ZstdLib& z = zstd();
    if (!z.ok) return fail(GC_ELIB, "libzstd.so.1 could not be loaded");
    size_t r = z.decompress(out, expected, src, n);
    if (z.isError(r)) return fail(GC_EFORMAT, std::string("ZSTD_decompress: ") + z.errorName(r));
    *out_len = r;
    return GC_OK;
}

int gc_lzma2_raw_encode(const uint8_t* src, size_t n, uint32_t preset, uint8_t* out, size_t cap, size_t* out_len) {
    if ((!src && n) || !out || !out_len) return fail(GC_EINVAL, "null argument");
    return lzma2_encode(src, n, preset, out, cap, out_len);
}

int gc_lzma2_raw_decode(const uint8_t* src, size_t n, uint8_t* out, size_t expected, size_t* out_len) {
    if ((!src && n) || (!out && expected) || !out_len) return fail(GC_EINVAL, "null argument");
    LzmaLib& L = lzma();
    if (!L.ok) return fail(GC_ELIB, "liblzma.so.5 could not be loaded");
    LzmaOptions opts;
    memset(&opts, 0, sizeof(opts));
    opts.dict_size = 64u << 20;  // compress_lzma2.rs:142
    LzmaFilter filters[2] = {{kFilterLzma2, &opts}, {kVliUnknown, nullptr}};
    LzmaStream s;
    memset(&s, 0, sizeof(s));
    if (L.raw_decoder(&s, filters) != kLzmaOk) return fail(GC_ELIB, "lzma_raw_decoder init failed");
    s.next_in = src;
    s.avail_in = n;
    s.next_out = out;
    s.avail_out = expected;
    int ret = L.code(&s, kLzmaFinish);
    size_t got = expected - s.avail_out;
    const size_t left = s.avail_in;
    L.end(&s);
    // The entry must be one whole stream (its end marker read) that decodes
    // into `expected` bytes: a stream still going when the output is full, or
    // bytes after the end marker, is a damaged entry, not a short read.
    if (ret != kLzmaStreamEnd) return fail(GC_EFORMAT, "LZMA2 decode failed or the entry decodes past its size");
    if (left) return fail(GC_EFORMAT, "LZMA2 entry has bytes after its end marker");
    *out_len = got;
    return GC_OK;
}

int gc_encode_single_sub_block(const uint8_t* sub, size_t n, const uint32_t* match_counts, uint8_t* out, size_t cap,
                               size_t* out_len) {
    if ((!sub && n) || (!match_counts && n) || !out || !out_len) return fail(GC_EINVAL, "null argument");
    bool any = false;
    for (size_t i = 0; i < n && !any; i++) any = match_counts[i] > 0;
    return encode_sub_block(sub, n, any, out, cap, out_len);
}

size_t gc_container_bound(size_t n, uint32_t algo) {
    const size_t nsub = (size_t)n_sub_of(n);
    size_t head = 28 + 8 * nsub;
    if (algo == GC_ALGO_ZSTD) {
        ZstdLib& z = zstd();
        if (!z.ok) return 0;
        return head + nsub * z.compressBound(GC_SUB_BLOCK);
    }
    return head + nsub * lzma2_sub_bound(GC_SUB_BLOCK);
}

int gc_nvzs_compress(gc_ctx* ctx, const uint8_t* src, size_t n, uint32_t level, int threads, uint8_t* out,
                     size_t cap, size_t* out_len) {
    if (!ctx || (!src && n) || !out || !out_len) return fail(GC_EINVAL, "null argument");
    if (level == 0) return fail(GC_EINVAL, "NVZS level 0 is the nvCOMP path, which this library does not build");
    const size_t nsub = (size_t)n_sub_of(n);
    const size_t head = 28 + 8 * nsub;
    if (cap < gc_container_bound(n, GC_ALGO_ZSTD)) return fail(GC_ESIZE, "output capacity below gc_container_bound");
    memcpy(out, "NVZS", 4);  // pipeline.rs:196-199
    put_u64(out + 4, n);
    put_u64(out + 12, GC_SUB_BLOCK);
    put_u64(out + 20, nsub);
    size_t o = head;
    std::vector<uint64_t> sizes(nsub, 0);
    // 8 MiB pipeline chunks (constants.rs:1); sub-chunks are tiled from the
    // input start, so the chunking does not change any frame.
    const size_t pchunk = 8u << 20;
    std::vector<uint32_t> counts(pchunk / GC_SUB_BLOCK);
    SeqBuf seqs((pchunk / GC_SUB_BLOCK) * kMaxSeq);
    if (!seqs.p) return fail(GC_ENOMEM, "sequence buffer");
    for (size_t c0 = 0; c0 < n; c0 += pchunk) {
        size_t len = std::min(pchunk, n - c0);
        size_t total = 0;
        int rc = gc_zstd_find_sequences_packed(ctx, src + c0, len, level, seqs.p, seqs.cap, counts.data(), &total);
        if (rc) return rc;
        size_t wrote = 0;
        rc = gc_zstd_encode_sequences(src + c0, len, seqs.p, counts.data(), out + o, cap - o,
                                      sizes.data() + c0 / GC_SUB_BLOCK, &wrote, threads);
        if (rc) return rc;
        o += wrote;
    }
    for (size_t s = 0; s < nsub; s++) put_u64(out + 28 + 8 * s, sizes[s]);
    *out_len = o;
    return GC_OK;
}

int gc_nvlz_compress(gc_ctx* ctx, const uint8_t* src, size_t n, uint32_t level, uint32_t dict_mb, int threads,
                     uint8_t* out, size_t cap, size_t* out_len) {
    if ((!src && n) || !out || !out_len) return fail(GC_EINVAL, "null argument");
    if (threads <= 0) threads = default_threads();
    if (cap < gc_container_bound(n, GC_ALGO_LZMA2)) return fail(GC_ESIZE, "output capacity below gc_container_bound");
    const size_t nsub = (size_t)n_sub_of(n);
    const size_t head = 28 + 8 * nsub;
    memcpy(out, "NVLZ", 4);  // pipeline_lzma2.rs:344-351
    put_u64(out + 4, n);
    put_u64(out + 12, GC_SUB_BLOCK);
    put_u64(out + 20, nsub);
    std::vector<uint64_t> sizes(nsub, 0);
    size_t o = head;
    if (level == 0) {
        return fail(GC_EINVAL, "level 0 is gc_nvlz_compress_l0 (liblzma on whole dictionary chunks)");
    }
    if (!ctx) return fail(GC_EINVAL, "null context");
    if (dict_mb == 0) dict_mb = 128;
    const size_t pchunk = (size_t)dict_mb << 20;  // pipeline_lzma2.rs:40
    std::vector<uint8_t> any_match;
    std::vector<std::vector<uint8_t>> enc;
    for (size_t c0 = 0; c0 < n; c0 += pchunk) {
        size_t len = std::min(pchunk, n - c0);
        const size_t ns = (size_t)n_sub_of(len);
        // the encoder only asks whether a sub-block has any match (compress_lzma2_custom.rs:577): one
        // byte per sub-block comes back instead of the match lists
        any_match.resize(ns);
        int rc = gc_lzma2_triage(ctx, src + c0, len, any_match.data());
        if (rc) return rc;
        enc.assign(ns, {});
        std::vector<int> errs(ns, 0);
        parallel_for(ns, threads, [&](size_t s) {
            const size_t st = s * GC_SUB_BLOCK, l = std::min<size_t>(GC_SUB_BLOCK, len - st);
            const bool any = any_match[s] != 0;
            enc[s].resize(lzma2_sub_bound(l));
            size_t w = 0;
            errs[s] = encode_sub_block(src + c0 + st, l, any, enc[s].data(), enc[s].size(), &w);
            enc[s].resize(w);
        });
        for (size_t s = 0; s < ns; s++) {
            if (errs[s]) return errs[s];
            if (o + enc[s].size() > cap) return fail(GC_ESIZE, "output capacity exceeded");
            memcpy(out + o, enc[s].data(), enc[s].size());
            o += enc[s].size();
            sizes[c0 / GC_SUB_BLOCK + s] = enc[s].size();
        }
    }
    for (size_t s = 0; s < nsub; s++) put_u64(out + 28 + 8 * s, sizes[s]);
    *out_len = o;
    return GC_OK;
}

size_t gc_nvlz_l0_bound(size_t n, uint32_t dict_mb) {
    if (dict_mb == 0) dict_mb = 128;
    size_t ch = (size_t)dict_mb << 20;
    size_t nch = (n + ch - 1) / ch;
    return 28 + 8 * nch + n + n / 8 + 1024 * (nch + 1);
}

int gc_nvlz_compress_l0(const uint8_t* src, size_t n, uint32_t dict_mb, uint32_t preset, int threads, uint8_t* out,
                        size_t cap, size_t* out_len) {
    if ((!src && n) || !out || !out_len) return fail(GC_EINVAL, "null argument");
    if (threads <= 0) threads = default_threads();
    if (dict_mb == 0) dict_mb = 128;
    const size_t ch = (size_t)dict_mb << 20;
    const size_t nch = (n + ch - 1) / ch;
    if (cap < gc_nvlz_l0_bound(n, dict_mb)) return fail(GC_ESIZE, "output capacity below gc_nvlz_l0_bound");
    memcpy(out, "NVLZ", 4);
    put_u64(out + 4, n);
    put_u64(out + 12, ch);  // level 0: entries are whole pipeline chunks (pipeline_lzma2.rs:344)
    put_u64(out + 20, nch);
    std::vector<std::vector<uint8_t>> enc(nch);
    std::vector<int> errs(nch, 0);
    parallel_for(nch, threads, [&](size_t c) {
        size_t st = c * ch, l = std::min(ch, n - st);
        enc[c].resize(l + l / 8 + 1024);
        size_t w = 0;
        errs[c] = lzma2_encode(src + st, l, preset, enc[c].data(), enc[c].size(), &w);
        enc[c].resize(w);
    });
    size_t o = 28 + 8 * nch;
    for (size_t c = 0; c < nch; c++) {
        if (errs[c]) return errs[c];
        memcpy(out + o, enc[c].data(), enc[c].size());
        o += enc[c].size();
        put_u64(out + 28 + 8 * c, enc[c].size());
    }
    *out_len = o;
    return GC_OK;
}

int gc_container_info(const uint8_t* in, size_t n, uint32_t* algo, uint64_t* original_size, uint64_t* chunk_size,
                      uint64_t* num_chunks) {
    if (!in || n < 28) return fail(GC_EFORMAT, "container shorter than its 28-byte header");
    uint32_t a;
    if (!memcmp(in, "NVZS", 4))
        a = GC_ALGO_ZSTD;
    else if (!memcmp(in, "NVLZ", 4))
        a = GC_ALGO_LZMA2;
    else
        return fail(GC_EFORMAT, "Invalid file format (expected NVZS or NVLZ)");  // decompress.rs:29-31
    uint64_t orig = get_u64(in + 4), ch = get_u64(in + 12), nch = get_u64(in + 20);
    if (ch == 0 ? orig != 0 || nch != 0 : nch != (orig + ch - 1) / ch)
        return fail(GC_EFORMAT, "entry count does not match original and chunk sizes");
    if (nch > (n - 28) / 8) return fail(GC_EFORMAT, "truncated size table");
    if (algo) *algo = a;
    if (original_size) *original_size = orig;
    if (chunk_size) *chunk_size = ch;
    if (num_chunks) *num_chunks = nch;
    return GC_OK;
}

int gc_decompress(const uint8_t* in, size_t n, uint8_t* out, size_t cap, size_t* out_len) {
    if (!out_len) return fail(GC_EINVAL, "null argument");
    uint32_t algo;
    uint64_t orig, ch, nch;
    int rc = gc_container_info(in, n, &algo, &orig, &ch, &nch);
    if (rc) return rc;
    if (cap < orig) return fail(GC_ESIZE, "output capacity below the original size");
    // every entry's place in the payload and in the output first (decompress.rs:65-73), then the entries
    // decoded by min(cores, 8) threads (they are independent)
    std::vector<uint64_t> pos(nch), dsz(nch);
    size_t p = 28 + 8 * nch;
    for (uint64_t c = 0; c < nch; c++) {
        const uint64_t csize = get_u64(in + 28 + 8 * c);
        if (csize > n - p) return fail(GC_EFORMAT, "truncated payload");  // read_exact, decompress.rs:81
        pos[c] = p;
        dsz[c] = (c + 1 < nch) ? ch : orig - (nch - 1) * ch;  // decompress.rs:70-73
        p += csize;
    }
    std::vector<int> errs(nch, 0);
    std::vector<std::string> msgs(nch);
    parallel_for(nch, default_threads(), [&](size_t c) {
        const uint64_t csize = (c + 1 < nch ? pos[c + 1] : p) - pos[c];
        size_t got = 0;
        uint8_t* dst = out + (size_t)c * ch;
        int r = algo == GC_ALGO_LZMA2 ? gc_lzma2_raw_decode(in + pos[c], csize, dst, dsz[c], &got)
                                      : gc_zstd_frame_decode(in + pos[c], csize, dst, dsz[c], &got);
        if (!r && got != dsz[c]) {
            r = GC_EFORMAT;
            set_error("an entry decoded to the wrong size");
        }
        if (r) {
            errs[c] = r;
            msgs[c] = gc_last_error();  // (the message is per thread)
        }
    });
    for (uint64_t c = 0; c < nch; c++)
        if (errs[c]) return fail(errs[c], msgs[c]);
    *out_len = (size_t)orig;
    return GC_OK;
}

}  // extern "C"
