// gc_lzma_enc.cpp -- LZMA2 encoding from the GPU's match lists.
//
// The encoder that compress_lzma2_custom.rs:344-465 (encode_lzma2_block) and
// :601-664 (cpu_encode_from_matches) sketch: one LZMA2 stream per 64 KiB
// sub-block, coded from the match finder's per-position matches instead of a
// second search on the CPU, each stream checked by decoding it and replaced by
// raw chunks if it does not round-trip (:640-647).  Written from the LZMA
// format as a decoder reads it (range coder, 12-state machine, literal coder
// with the matched-literal mode after a match, repeat distances 0..3 and the
// short repeat, distance slots with their reverse-coded low bits and the
// 4-bit align field, two length coders); each coding step below is the mirror
// of the corresponding decoding step, so the decoder's reading of the stream
// fixes it.  Properties lc = 3, lp = 0, pb = 2 (compress_lzma2_custom.rs:108-110).
//
// The parse is greedy with one lazy step: at each position the GPU's match
// (one per position: distance << 16 | length, lzma2_match_find.cu:54-194)
// competes with the four repeat distances, whose lengths are measured here.
#include <string.h>

#include <algorithm>
#include <atomic>
#include <thread>
#include <vector>

#include "../../include/gcompress.h"
#include "gc_runtime.h"

namespace gc {
namespace {

constexpr uint32_t kLc = 3, kPb = 2;
constexpr uint32_t kPosStates = 1u << kPb;
constexpr uint32_t kMatchMin = 2, kMatchMax = 273;
constexpr uint32_t kEndPosModel = 14, kFullDistances = 128, kAlignBits = 4;

using Prob = uint16_t;
constexpr Prob kProbInit = 1024;  // half of 1 << 11

// Range encoder (carry kept in the 33rd bit of `low`, pending 0xFF bytes
// counted in `cache_size`).
struct RangeEnc {
    uint64_t low = 0;
    uint32_t range = 0xFFFFFFFFu;
    uint8_t cache = 0;
    uint64_t cache_size = 1;
    uint8_t* out;
    size_t cap, n = 0;
    bool full = false;

    RangeEnc(uint8_t* o, size_t c) : out(o), cap(c) {}
    void put(uint8_t b) {
        if (n < cap)
            out[n++] = b;
        else
            full = true;
    }
    void shift_low() {
        if ((uint32_t)low < 0xFF000000u || (low >> 32) != 0) {
            const uint8_t carry = (uint8_t)(low >> 32);
            uint8_t b = cache;
            do {
                put((uint8_t)(b + carry));
                b = 0xFF;
            } while (--cache_size != 0);
            cache = (uint8_t)(low >> 24);
        }
        cache_size++;
        low = (low & 0x00FFFFFFu) << 8;
    }
    void normalize() {
        while (range < (1u << 24)) {
            range <<= 8;
            shift_low();
        }
    }
    void bit(Prob& p, uint32_t b) {
        const uint32_t bound = (range >> 11) * p;
        if (b == 0) {
            range = bound;
            p = (Prob)(p + ((2048 - p) >> 5));
        } else {
            low += bound;
            range -= bound;
            p = (Prob)(p - (p >> 5));
        }
        normalize();
    }
    void direct(uint32_t v, uint32_t nbits) {
        while (nbits--) {
            range >>= 1;
            if ((v >> nbits) & 1) low += range;
            normalize();
        }
    }
    // most significant bit first, node index from 1
    void tree(Prob* probs, uint32_t nbits, uint32_t v) {
        uint32_t m = 1;
        while (nbits--) {
            const uint32_t b = (v >> nbits) & 1;
            bit(probs[m], b);
            m = (m << 1) | b;
        }
    }
    // least significant bit first
    void tree_reverse(Prob* probs, uint32_t nbits, uint32_t v) {
        uint32_t m = 1;
        for (uint32_t i = 0; i < nbits; i++) {
            const uint32_t b = (v >> i) & 1;
            bit(probs[m], b);
            m = (m << 1) | b;
        }
    }
    void flush() {
        for (int i = 0; i < 5; i++) shift_low();
    }
};

struct LenModel {
    Prob choice, choice2, low[kPosStates][8], mid[kPosStates][8], high[256];
};

struct Model {
    Prob is_match[12][kPosStates], is_rep[12], is_rep_g0[12], is_rep_g1[12], is_rep_g2[12];
    Prob is_rep0_long[12][kPosStates];
    Prob lit[1u << kLc][0x300];
    Prob pos_slot[4][64];
    Prob pos_special[1 + kFullDistances - kEndPosModel];
    Prob align[1u << kAlignBits];
    LenModel len, rep_len;
    Model() {
        Prob* p = (Prob*)this;
        for (size_t i = 0; i < sizeof(Model) / sizeof(Prob); i++) p[i] = kProbInit;
    }
};

uint32_t after_literal(uint32_t s) { return s < 4 ? 0 : (s < 10 ? s - 3 : s - 6); }
uint32_t after_match(uint32_t s) { return s < 7 ? 7 : 10; }
uint32_t after_rep(uint32_t s) { return s < 7 ? 8 : 11; }
uint32_t after_short_rep(uint32_t s) { return s < 7 ? 9 : 11; }

struct Encoder {
    RangeEnc rc;
    Model m;
    const uint8_t* d;
    uint32_t n;
    uint32_t state = 0;
    uint32_t rep[4] = {1, 1, 1, 1};  // distances (the decoder's rep0 + 1)

    Encoder(const uint8_t* data, uint32_t size, uint8_t* out, size_t cap) : rc(out, cap), d(data), n(size) {}

    void length(LenModel& lm, uint32_t len, uint32_t ps) {
        const uint32_t l = len - kMatchMin;
        if (l < 8) {
            rc.bit(lm.choice, 0);
            rc.tree(lm.low[ps], 3, l);
        } else if (l < 16) {
            rc.bit(lm.choice, 1);
            rc.bit(lm.choice2, 0);
            rc.tree(lm.mid[ps], 3, l - 8);
        } else {
            rc.bit(lm.choice, 1);
            rc.bit(lm.choice2, 1);
            rc.tree(lm.high, 8, l - 16);
        }
    }

    void literal(uint32_t p) {
        const uint32_t ps = p & (kPosStates - 1);
        rc.bit(m.is_match[state][ps], 0);
        const uint32_t prev = p ? d[p - 1] : 0;
        Prob* probs = m.lit[prev >> (8 - kLc)];
        const uint32_t byte = d[p];
        uint32_t sym = 1;
        if (state >= 7) {
            // matched literal: the byte at the last distance steers the
            // probabilities until the first bit where the two differ
            const uint32_t mbyte = d[p - rep[0]];
            bool same = true;
            for (int i = 7; i >= 0; i--) {
                const uint32_t b = (byte >> i) & 1;
                if (same) {
                    const uint32_t mb = (mbyte >> i) & 1;
                    rc.bit(probs[((1 + mb) << 8) + sym], b);
                    same = mb == b;
                } else {
                    rc.bit(probs[sym], b);
                }
                sym = (sym << 1) | b;
            }
        } else {
            rc.tree(probs, 8, byte);
        }
        state = after_literal(state);
    }

    void match(uint32_t p, uint32_t dist, uint32_t len) {
        const uint32_t ps = p & (kPosStates - 1);
        rc.bit(m.is_match[state][ps], 1);
        rc.bit(m.is_rep[state], 0);
        length(m.len, len, ps);
        const uint32_t dd = dist - 1;
        uint32_t slot;
        if (dd < 4) {
            slot = dd;
        } else {
            const uint32_t top = 31u - (uint32_t)__builtin_clz(dd);
            slot = 2 * top + ((dd >> (top - 1)) & 1);
        }
        rc.tree(m.pos_slot[std::min<uint32_t>(len - kMatchMin, 3)], 6, slot);
        if (slot >= 4) {
            const uint32_t nbits = (slot >> 1) - 1;
            const uint32_t base = (2 | (slot & 1)) << nbits;
            const uint32_t reduced = dd - base;
            if (slot < kEndPosModel) {
                rc.tree_reverse(m.pos_special + base - slot, nbits, reduced);
            } else {
                rc.direct(reduced >> kAlignBits, nbits - kAlignBits);
                rc.tree_reverse(m.align, kAlignBits, reduced & ((1u << kAlignBits) - 1));
            }
        }
        rep[3] = rep[2];
        rep[2] = rep[1];
        rep[1] = rep[0];
        rep[0] = dist;
        state = after_match(state);
    }

    // repeat distance i; len 1 with i == 0 is the short repeat
    void rep_match(uint32_t p, uint32_t i, uint32_t len) {
        const uint32_t ps = p & (kPosStates - 1);
        rc.bit(m.is_match[state][ps], 1);
        rc.bit(m.is_rep[state], 1);
        if (i == 0) {
            rc.bit(m.is_rep_g0[state], 0);
            rc.bit(m.is_rep0_long[state][ps], len == 1 ? 0 : 1);
        } else {
            rc.bit(m.is_rep_g0[state], 1);
            if (i == 1) {
                rc.bit(m.is_rep_g1[state], 0);
            } else {
                rc.bit(m.is_rep_g1[state], 1);
                rc.bit(m.is_rep_g2[state], i == 3 ? 1 : 0);
            }
            const uint32_t dist = rep[i];
            for (uint32_t k = i; k > 0; k--) rep[k] = rep[k - 1];
            rep[0] = dist;
        }
        if (len == 1) {
            state = after_short_rep(state);
        } else {
            length(m.rep_len, len, ps);
            state = after_rep(state);
        }
    }

    // bytes in common at p and p - dist, at most `limit`
    uint32_t common(uint32_t p, uint32_t dist, uint32_t limit) const {
        if (dist > p) return 0;
        const uint8_t* a = d + p;
        const uint8_t* b = a - dist;
        uint32_t l = 0;
        while (l < limit && a[l] == b[l]) l++;
        return l;
    }
};

// the GPU's match at p (0 length when there is none), clamped to the block
// and to the format's longest match
inline void gpu_match(const uint32_t* packed, uint32_t p, uint32_t n, uint32_t& dist, uint32_t& len) {
    const uint32_t w = packed[p];
    dist = w >> 16;
    len = std::min<uint32_t>(std::min<uint32_t>(w & 0xFFFFu, kMatchMax), n - p);
    if (len < kMatchMin || dist == 0 || dist > p) len = 0;
}

// Raw LZMA2 chunks (compress_lzma2_custom.rs:325-338) and the end marker.
size_t raw_chunks(const uint8_t* src, size_t n, uint8_t* out) {
    size_t o = 0;
    for (size_t off = 0; off < n;) {
        const size_t len = std::min<size_t>(n - off, 65536);
        out[o++] = off == 0 ? 0x01 : 0x02;
        out[o++] = (uint8_t)((len - 1) >> 8);
        out[o++] = (uint8_t)(len - 1);
        memcpy(out + o, src + off, len);
        o += len;
        off += len;
    }
    out[o++] = 0x00;
    return o;
}

// One sub-block as one LZMA chunk: header (5 B + props), range-coded data,
// end marker.  Returns 0 when the coded data would not fit a chunk (64 KiB)
// or `cap`.
size_t encode_chunk(const uint8_t* d, uint32_t n, const uint32_t* packed, uint8_t* out, size_t cap) {
    if (n == 0 || cap < 8) return 0;
    const size_t room = std::min<size_t>(cap - 7, 65536);
    Encoder e(d, n, out + 6, room);
    constexpr uint32_t kNice = 64;
    uint32_t p = 0;
    while (p < n && !e.rc.full) {
        const uint32_t avail = std::min(n - p, kMatchMax);
        uint32_t mdist, mlen;
        gpu_match(packed, p, n, mdist, mlen);
        // the repeat distances
        uint32_t rlen = 0, ridx = 0;
        if (avail >= kMatchMin)
            for (uint32_t i = 0; i < 4; i++) {
                const uint32_t l = e.common(p, e.rep[i], avail);
                if (l > rlen) rlen = l, ridx = i;
            }
        // a match at a repeat distance is coded as that repeat
        for (uint32_t i = 0; i < 4 && mlen; i++)
            if (e.rep[i] == mdist && mlen >= rlen) rlen = mlen, ridx = i, mlen = 0;
        bool take_rep = false, take_match = false;
        if (rlen >= kMatchMin &&
            (rlen >= kNice || rlen + 1 >= mlen || (rlen + 2 >= mlen && mdist >= (1u << 9)) ||
             (rlen + 3 >= mlen && mdist >= (1u << 15))))
            take_rep = true;
        else if (mlen >= kMatchMin && !(mlen == 2 && mdist >= 128))
            take_match = true;
        if (take_match && mlen < kNice && p + 1 < n) {
            // one lazy step: a clearly longer match one byte later wins
            uint32_t ndist, nlen;
            gpu_match(packed, p + 1, n, ndist, nlen);
            if (nlen >= mlen + 2 || (nlen == mlen + 1 && ndist < mdist)) take_match = false;
        }
        if (take_rep) {
            e.rep_match(p, ridx, rlen);
            p += rlen;
        } else if (take_match) {
            e.match(p, mdist, mlen);
            p += mlen;
        } else if (p >= e.rep[0] && d[p] == d[p - e.rep[0]] && e.state >= 7) {
            // after a match the literal would be coded against this same byte
            e.rep_match(p, 0, 1);
            p += 1;
        } else {
            e.literal(p);
            p += 1;
        }
    }
    e.rc.flush();
    if (e.rc.full || e.rc.n > 65536) return 0;
    const uint32_t c = (uint32_t)e.rc.n;
    out[0] = (uint8_t)(0xE0 | ((n - 1) >> 16));  // LZMA, dictionary + state reset, new properties
    out[1] = (uint8_t)((n - 1) >> 8);
    out[2] = (uint8_t)(n - 1);
    out[3] = (uint8_t)((c - 1) >> 8);
    out[4] = (uint8_t)(c - 1);
    out[5] = (uint8_t)((kPb * 5 + 0) * 9 + kLc);  // lc 3, lp 0, pb 2
    out[6 + c] = 0x00;
    return 7 + (size_t)c;
}

size_t raw_bound(size_t n) { return n + 3 * ((n + 65535) / 65536) + 1; }

int encode_one(const uint8_t* sub, size_t n, const uint32_t* packed, uint8_t* out, size_t cap, size_t* out_len,
               bool verify, std::vector<uint8_t>* scratch) {
    size_t w = encode_chunk(sub, (uint32_t)n, packed, out, cap);
    if (w && w < raw_bound(n) && verify) {
        // the safety net of compress_lzma2_custom.rs:640-647: decode it back
        scratch->resize(n);
        size_t got = 0;
        if (gc_lzma2_raw_decode(out, w, scratch->data(), n, &got) != GC_OK || got != n ||
            memcmp(scratch->data(), sub, n) != 0)
            w = 0;
    }
    if (!w || w >= raw_bound(n)) {
        if (cap < raw_bound(n)) return fail(GC_ESIZE, "output buffer too small for raw LZMA2 chunks");
        w = raw_chunks(sub, n, out);
    }
    *out_len = w;
    return GC_OK;
}

}  // namespace
}  // namespace gc

using namespace gc;

extern "C" {

size_t gc_lzma2_encode_bound(size_t n) { return raw_bound(n) + 8; }

int gc_lzma2_encode_matches(const uint8_t* sub, size_t n, const uint32_t* packed, int verify, uint8_t* out,
                            size_t cap, size_t* out_len) {
    if ((!sub && n) || (!packed && n) || !out || !out_len) return fail(GC_EINVAL, "null argument");
    if (n > GC_SUB_BLOCK) return fail(GC_EINVAL, "one sub-block of at most 65536 bytes");
    if (n == 0) {
        if (cap < 1) return fail(GC_ESIZE, "output buffer too small");
        out[0] = 0x00;
        *out_len = 1;
        return GC_OK;
    }
    std::vector<uint8_t> scratch;
    return encode_one(sub, n, packed, out, cap, out_len, verify != 0, &scratch);
}

int gc_cpu_encode_from_matches(const uint8_t* src, size_t n, const uint32_t* packed, int threads, uint8_t* out,
                               size_t cap, uint64_t* sizes, size_t* out_len) {
    if ((!src && n) || (!packed && n) || !out || !sizes || !out_len) return fail(GC_EINVAL, "null argument");
    const size_t nsub = (n + GC_SUB_BLOCK - 1) / GC_SUB_BLOCK;
    if (threads <= 0) threads = (int)std::min<unsigned>(std::max(1u, std::thread::hardware_concurrency()), 8u);
    std::vector<std::vector<uint8_t>> enc(nsub);
    std::vector<int> errs(nsub, 0);
    std::atomic<size_t> next{0};
    auto work = [&] {
        std::vector<uint8_t> scratch;
        for (;;) {
            const size_t s = next.fetch_add(1);
            if (s >= nsub) break;
            const size_t st = s * GC_SUB_BLOCK, l = std::min<size_t>(GC_SUB_BLOCK, n - st);
            enc[s].resize(raw_bound(l) + 8);
            size_t w = 0;
            errs[s] = encode_one(src + st, l, packed + st, enc[s].data(), enc[s].size(), &w, true, &scratch);
            enc[s].resize(w);
        }
    };
    std::vector<std::thread> pool;
    const int t = (int)std::min<size_t>((size_t)threads, std::max<size_t>(nsub, 1));
    for (int k = 1; k < t; k++) pool.emplace_back(work);
    work();
    for (auto& th : pool) th.join();
    size_t o = 0;
    for (size_t s = 0; s < nsub; s++) {
        if (errs[s]) return errs[s];
        if (o + enc[s].size() > cap) return fail(GC_ESIZE, "output capacity exceeded");
        memcpy(out + o, enc[s].data(), enc[s].size());
        o += enc[s].size();
        sizes[s] = enc[s].size();
    }
    *out_len = o;
    return GC_OK;
}

int gc_nvlz_compress_matches(gc_ctx* ctx, const uint8_t* src, size_t n, uint32_t dict_mb, int threads, uint8_t* out,
                             size_t cap, size_t* out_len) {
    if (!ctx || (!src && n) || !out || !out_len) return fail(GC_EINVAL, "null argument");
    if (cap < gc_container_bound(n, GC_ALGO_LZMA2)) return fail(GC_ESIZE, "output capacity below gc_container_bound");
    const size_t nsub = (n + GC_SUB_BLOCK - 1) / GC_SUB_BLOCK;
    const size_t head = 28 + 8 * nsub;
    memcpy(out, "NVLZ", 4);  // pipeline_lzma2.rs:344-351
    auto put64 = [](uint8_t* p, uint64_t v) {
        for (int i = 0; i < 8; i++) p[i] = (uint8_t)(v >> (8 * i));
    };
    put64(out + 4, n);
    put64(out + 12, GC_SUB_BLOCK);
    put64(out + 20, nsub);
    if (dict_mb == 0) dict_mb = 128;
    const size_t pchunk = (size_t)dict_mb << 20;  // pipeline_lzma2.rs:40
    std::vector<uint32_t> packed;
    std::vector<uint64_t> sizes(nsub, 0);
    size_t o = head;
    for (size_t c0 = 0; c0 < n; c0 += pchunk) {
        const size_t len = std::min(pchunk, n - c0);
        packed.resize(len);
        int rc = gc_lzma2_find_matches_packed(ctx, src + c0, len, packed.data());
        if (rc) return rc;
        size_t w = 0;
        rc = gc_cpu_encode_from_matches(src + c0, len, packed.data(), threads, out + o, cap - o,
                                        sizes.data() + c0 / GC_SUB_BLOCK, &w);
        if (rc) return rc;
        o += w;
    }
    for (size_t s = 0; s < nsub; s++) put64(out + 28 + 8 * s, sizes[s]);
    *out_len = o;
    return GC_OK;
}

}  // extern "C"
