// gc_corpus.cpp -- deterministic synthetic corpora for the match-finding workloads.
//
// Four kinds of content, each written in whole segments:
//   0 uniform random bytes
//   1 word text: words of 3-12 symbols over a 64-symbol alphabet, from a
//     4096-word vocabulary drawn Zipf(s = 1.1) by rank, joined by spaces
//   2 back-references: copies of earlier output at a log-uniform distance in
//     [1 B, 32 MiB] (clamped to the history), geometric length of mean 32,
//     each copied byte replaced by a random one with probability 2%
//   3 structured binary: 16-byte little-endian u32 records (counter, slowly
//     drifting value, low-entropy field, fixed tag); after each record a zero
//     run of geometric length (mean 4 KiB) starts with probability 1/256
// The generator is splitmix64-seeded xoshiro256**; the same arguments give the
// same bytes on every machine.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/gcompress.h"
#include "gc_runtime.h"

namespace {

struct Rng {
    uint64_t s[4];
    static uint64_t splitmix(uint64_t& x) {
        uint64_t z = (x += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    explicit Rng(uint64_t seed) {
        for (int i = 0; i < 4; i++) s[i] = splitmix(seed);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        uint64_t r = rotl(s[1] * 5, 7) * 9;
        uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0, 1)
    uint64_t below(uint64_t n) { return n ? next() % n : 0; }
    // number of trials up to and including the first success, success prob p
    uint64_t geometric(double p) {
        double u = uniform();
        return 1 + (uint64_t)floor(log1p(-u) / log1p(-p));
    }
};

const char kAlphabet[] = "abcdefghijklmnopqrstuvwxyzABCDEFGHIJKLMNOPQRSTUVWXYZ0123456789-_";

struct Gen {
    uint8_t* out;
    uint64_t n;
    uint64_t pos = 0;
    Rng rng;
    // text
    std::vector<std::string> vocab;
    std::vector<double> zipf_cdf;
    // structured
    uint32_t counter = 0, drift = 1u << 20;
    uint32_t tag;

    Gen(uint8_t* o, uint64_t size, uint64_t seed) : out(o), n(size), rng(seed) {
        Rng v(seed ^ 0x766F636162756C61ull);
        vocab.resize(4096);
        for (auto& w : vocab) {
            size_t len = 3 + v.below(10);
            w.resize(len);
            for (auto& c : w) c = kAlphabet[v.below(64)];
        }
        zipf_cdf.resize(4096);
        double acc = 0;
        for (int r = 0; r < 4096; r++) {
            acc += pow((double)(r + 1), -1.1);
            zipf_cdf[r] = acc;
        }
        for (auto& c : zipf_cdf) c /= acc;
        tag = 0xC0DE0000u | (uint32_t)(seed & 0xFFFF);
    }

    void random_bytes(uint64_t end) {
        while (pos < end) {
            uint64_t r = rng.next();
            for (int i = 0; i < 8 && pos < end; i++) out[pos++] = (uint8_t)(r >> (8 * i));
        }
    }

    void text(uint64_t end) {
        while (pos < end) {
            double u = rng.uniform();
            size_t r = std::lower_bound(zipf_cdf.begin(), zipf_cdf.end(), u) - zipf_cdf.begin();
            if (r >= vocab.size()) r = vocab.size() - 1;
            const std::string& w = vocab[r];
            for (size_t i = 0; i < w.size() && pos < end; i++) out[pos++] = (uint8_t)w[i];
            if (pos < end) out[pos++] = ' ';
        }
    }

    void backrefs(uint64_t end) {
        const double max_log = log(32.0 * 1024 * 1024);
        while (pos < end) {
            if (pos == 0) {
                out[pos++] = (uint8_t)rng.next();
                continue;
            }
            uint64_t d = (uint64_t)floor(exp(rng.uniform() * max_log));
            if (d < 1) d = 1;
            if (d > pos) d = pos;
            uint64_t len = rng.geometric(1.0 / 32.0);
            for (uint64_t i = 0; i < len && pos < end; i++) {
                uint8_t b = out[pos - d];
                if (rng.uniform() < 0.02) b = (uint8_t)rng.next();
                out[pos++] = b;
            }
        }
    }

    void put32(uint32_t v, uint64_t end) {
        for (int i = 0; i < 4 && pos < end; i++) out[pos++] = (uint8_t)(v >> (8 * i));
    }

    void structured(uint64_t end) {
        while (pos < end) {
            counter++;
            uint64_t r = rng.next();
            if ((r & 0xFF) < 26) drift += ((r >> 8) & 1) ? 1u : (uint32_t)-1;  // ~10% of records move by 1
            static const uint32_t low[4] = {0, 1, 2, 0xFFFFFFFFu};
            uint32_t q = (uint32_t)((r >> 16) & 0xFF);
            uint32_t le = low[q < 160 ? 0 : q < 220 ? 1 : q < 250 ? 2 : 3];
            put32(counter, end);
            put32(drift, end);
            put32(le, end);
            put32(tag, end);
            if (((r >> 32) & 0xFF) == 0) {  // probability 1/256
                uint64_t z = rng.geometric(1.0 / 4096.0);
                uint64_t stop = std::min(end, pos + z);
                memset(out + pos, 0, stop - pos);
                pos = stop;
            }
        }
    }

    void segment(int kind, uint64_t end) {
        switch (kind) {
            case 0: random_bytes(end); break;
            case 1: text(end); break;
            case 2: backrefs(end); break;
            default: structured(end); break;
        }
    }
};

}  // namespace

extern "C" int gc_corpus_generate(uint8_t* out, uint64_t n, uint64_t seed, const double* weights, uint64_t segment,
                                  uint32_t exact_shares, uint64_t gap_every, uint64_t gap_len) {
    if ((!out && n) || !weights || segment == 0) return gc::fail(GC_EINVAL, "gc_corpus_generate: bad argument");
    double w[4], sum = 0;
    for (int i = 0; i < 4; i++) {
        w[i] = weights[i] < 0 ? 0 : weights[i];
        sum += w[i];
    }
    if (sum <= 0) return gc::fail(GC_EINVAL, "gc_corpus_generate: weights sum to zero");
    for (double& x : w) x /= sum;

    Gen g(out, n, seed);
    // Kind schedule.  exact_shares: groups of 20 segments hold round(20*w)
    // segments of each kind in a shuffled order; otherwise each segment's
    // kind is drawn independently with probabilities w.
    std::vector<int> group;
    if (exact_shares) {
        int cnt[4], tot = 0;
        for (int i = 0; i < 4; i++) {
            cnt[i] = (int)floor(w[i] * 20 + 0.5);
            tot += cnt[i];
        }
        while (tot > 20) {  // keep the group at 20 after rounding
            int k = (int)(std::max_element(cnt, cnt + 4) - cnt);
            cnt[k]--;
            tot--;
        }
        while (tot < 20) {
            int k = (int)(std::max_element(w, w + 4) - w);
            cnt[k]++;
            tot++;
        }
        for (int k = 0; k < 4; k++)
            for (int j = 0; j < cnt[k]; j++) group.push_back(k);
    }
    std::vector<int> order;
    uint64_t seg_index = 0;
    uint64_t content = 0;  // bytes of segment content since the last gap
    while (g.pos < n) {
        if (gap_every && gap_len && content >= gap_every) {
            uint64_t stop = std::min(n, g.pos + gap_len);
            memset(out + g.pos, 0, stop - g.pos);
            g.pos = stop;
            content = 0;
            continue;
        }
        int kind;
        if (exact_shares) {
            if (seg_index % group.size() == 0) {
                order = group;
                for (size_t i = order.size(); i > 1; i--) std::swap(order[i - 1], order[g.rng.below(i)]);
            }
            kind = order[seg_index % group.size()];
        } else {
            double u = g.rng.uniform(), acc = 0;
            kind = 3;
            for (int k = 0; k < 4; k++) {
                acc += w[k];
                if (u < acc) {
                    kind = k;
                    break;
                }
            }
        }
        uint64_t len = segment;
        if (gap_every && gap_len && content + len > gap_every) len = gap_every - content;
        uint64_t end = std::min(n, g.pos + len);
        content += end - g.pos;
        g.segment(kind, end);
        seg_index++;
    }
    return GC_OK;
}
