/*
 * gc_oracle.c -- CPU oracle for the two match finders.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke()
 * entry point and the CPU-baseline leg of bench.py may load it; the product
 * library (gpu-compressor_amd/) never links or calls it.
 *
 * It restates, serially and in plain C, what the reference's two device
 * kernels compute for each 64 KiB sub-block:
 *
 *   zstd_match_find   zstd_compress.cu:179-355   (orc_zstd_match_find)
 *   lzma2_match_find  lzma2_match_find.cu:54-194 (orc_lzma2_match_find)
 *
 * with the reference's output layouts:
 *   Zstd : Sequence{u32 literal_length, u32 match_length-3, u32 offset} at
 *          seq[s*max_seq + i]; literals at lit[s*chunk + j];
 *          seq_counts[s], lit_counts[s]        (zstd_compress.cu:38-42,192-193)
 *   LZMA2: LzmaMatch{u32 distance, u32 length} at m[(s*sub + p)*max + k];
 *          counts[s*sub + p]                    (lzma2_match_find.cu:26-29,84-85)
 *
 * Pinning (what checks this restatement, beyond the GPU comparing with it):
 * hand-derived known answers and closed forms (all-zero, periodic, hash
 * values, repeat-offset order, lazy rule, sequence cap, RLE fallback),
 * brute-force candidate search, reconstruction of the input from the
 * sequences, maximality of every LZMA2 match, and round trips through
 * stock libzstd / liblzma.  See tests/test_oracle.py.
 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_HASH_LOG 14
#define ORC_HASH_SIZE (1u << ORC_HASH_LOG)
#define ORC_ZSTD_MIN_MATCH 3u
#define ORC_ZSTD_MAX_MATCH 131074u /* zstd_compress.cu:29 */
#define ORC_LZMA_MAX_MATCH 273u    /* lzma2_match_find.cu:22 */

/* Little-endian 4-byte read; zstd_compress.cu:32-35, lzma2_match_find.cu:32-35. */
static inline uint32_t orc_le32(const uint8_t* p) {
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}

/* Multiplicative 14-bit hash of 4 bytes; zstd_compress.cu:228, lzma2_match_find.cu:92. */
uint32_t orc_hash4(const uint8_t* p) {
    return (uint32_t)(orc_le32(p) * 2654435761u) >> (32 - ORC_HASH_LOG);
}

/* Number of equal bytes of a[] and b[] from 0, stopping at limit. */
static inline uint32_t orc_common(const uint8_t* a, const uint8_t* b, uint32_t start, uint32_t limit) {
    uint32_t k = start;
    while (k < limit && a[k] == b[k]) k++;
    return k;
}

/* ------------------------------------------------------------------------ */
/* Zstd: zstd_compress.cu:179-355, one sub-block.                            */
/* ------------------------------------------------------------------------ */
static void orc_zstd_block(const uint8_t* src, uint32_t n, uint32_t depth, uint32_t max_seq,
                           uint32_t* seq /* 3*max_seq */, uint8_t* lit /* may be NULL */,
                           uint32_t* seq_count, uint32_t* lit_count) {
    uint16_t table[ORC_HASH_SIZE];
    memset(table, 0, sizeof(table)); /* empty slot reads as position 0 (:198-200) */
    uint32_t rep[3] = {1, 4, 8};     /* :225 */
    uint32_t pos = 0, nseq = 0, nlit = 0, lit_start = 0;

    while (pos + 3 < n) { /* :230 */
        uint32_t h = orc_hash4(src + pos);
        uint32_t cand = table[h];
        table[h] = (uint16_t)(pos & 0xFFFF);
        uint32_t off = pos - cand; /* u32 arithmetic, as :238 */
        uint32_t len = 0;

        /* hash candidate: 3-byte verify then extend (:241-257) */
        if (off > 0 && off <= pos && cand < n) {
            if (src[cand] == src[pos] && src[cand + 1] == src[pos + 1] && src[cand + 2] == src[pos + 2]) {
                len = 3;
                while (pos + len < n && cand + len < n && src[cand + len] == src[pos + len] &&
                       len < ORC_ZSTD_MAX_MATCH)
                    len++;
            }
        }
        /* repeat offsets in order; the first that verifies wins (:259-276) */
        for (int r = 0; r < 3 && len < ORC_ZSTD_MIN_MATCH; r++) {
            if (rep[r] > pos) continue;
            uint32_t rp = pos - rep[r];
            if (src[rp] != src[pos] || src[rp + 1] != src[pos + 1] || src[rp + 2] != src[pos + 2]) continue;
            uint32_t rl = 3;
            while (pos + rl < n && rp + rl < n && src[rp + rl] == src[pos + rl] && rl < ORC_ZSTD_MAX_MATCH) rl++;
            len = rl;
            off = rep[r];
            break;
        }
        /* one lazy step at pos+1 when depth > 1 (:278-307) */
        if (len >= ORC_ZSTD_MIN_MATCH && depth > 1 && pos + 4 < n) {
            uint32_t h2 = orc_hash4(src + pos + 1);
            uint32_t c2 = table[h2];
            uint32_t off2 = (pos + 1) - c2;
            if (off2 > 0 && off2 <= pos + 1 && c2 < n && src[c2] == src[pos + 1] && src[c2 + 1] == src[pos + 2] &&
                src[c2 + 2] == src[pos + 3]) {
                uint32_t l2 = 3;
                while (pos + 1 + l2 < n && c2 + l2 < n && src[c2 + l2] == src[pos + 1 + l2] &&
                       l2 < ORC_ZSTD_MAX_MATCH)
                    l2++;
                if (l2 > len + 1) {
                    pos++;
                    len = l2;
                    off = off2;
                    table[h2] = (uint16_t)(pos & 0xFFFF);
                }
            }
        }
        if (len >= ORC_ZSTD_MIN_MATCH && nseq < max_seq) { /* :309 */
            if (lit) memcpy(lit + nlit, src + lit_start, pos - lit_start);
            nlit += pos - lit_start;
            seq[3 * nseq + 0] = pos - lit_start;
            seq[3 * nseq + 1] = len - ORC_ZSTD_MIN_MATCH;
            seq[3 * nseq + 2] = off;
            nseq++;
            rep[2] = rep[1];
            rep[1] = rep[0];
            rep[0] = off;
            for (uint32_t i = 1; i < len && pos + i + 3 < n; i++) /* :328-331 */
                table[orc_hash4(src + pos + i)] = (uint16_t)((pos + i) & 0xFFFF);
            pos += len;
            lit_start = pos;
        } else {
            pos++;
        }
    }
    if (lit && n > lit_start) memcpy(lit + nlit, src + lit_start, n - lit_start); /* :341-344 */
    if (n > lit_start) nlit += n - lit_start;
    *seq_count = nseq;
    *lit_count = nlit;
}

/* ------------------------------------------------------------------------ */
/* LZMA2: lzma2_match_find.cu:54-194, one sub-block.                         */
/* ------------------------------------------------------------------------ */
static void orc_lzma2_block(const uint8_t* src, uint32_t n, uint32_t sub, uint32_t max_matches, uint32_t depth,
                            uint32_t* m /* 2*max_matches per position */, uint32_t* counts) {
    uint16_t table[ORC_HASH_SIZE];
    memset(table, 0xFF, sizeof(table)); /* 0xFFFF = empty (:73) */
    for (uint32_t pos = 0; pos < n; pos++) {
        uint32_t cnt = 0;
        uint32_t* slot = m + (uint64_t)pos * max_matches * 2;
        if (pos + 1 < n) {
            uint32_t best = 1;
            if (pos + 3 < n) {
                uint32_t h = orc_hash4(src + pos);
                uint32_t cur = table[h];
                table[h] = (uint16_t)(pos & 0xFFFF);
                /* one probe; the chain walk always breaks after it (:114-161) */
                if (cur != 0xFFFF && depth > 0 && cur < pos) {
                    uint32_t dist = pos - cur;
                    if (dist > 0 && cur + 1 < n) {
                        uint32_t maxlen = n - pos;
                        if (maxlen > ORC_LZMA_MAX_MATCH) maxlen = ORC_LZMA_MAX_MATCH;
                        if (n - cur < maxlen) maxlen = n - cur;
                        uint32_t len = 0;
                        if (pos + 3 < n && cur + 3 < n && orc_le32(src + cur) == orc_le32(src + pos))
                            len = orc_common(src + cur, src + pos, 4, maxlen);
                        else if (src[cur] == src[pos] && src[cur + 1] == src[pos + 1])
                            len = orc_common(src + cur, src + pos, 2, maxlen);
                        if (len >= 2 && len > best) {
                            slot[2 * cnt + 0] = dist;
                            slot[2 * cnt + 1] = len;
                            cnt++;
                            best = len;
                        }
                    }
                }
            }
            /* distance-1 run, only when the probe produced nothing (:165-181) */
            if (cnt == 0 && pos > 0 && src[pos] == src[pos - 1] && src[pos + 1] == src[pos]) {
                uint32_t maxlen = n - pos;
                if (maxlen > ORC_LZMA_MAX_MATCH) maxlen = ORC_LZMA_MAX_MATCH;
                uint32_t len = 2;
                while (len < maxlen && src[pos + len] == src[pos]) len++;
                slot[0] = 1;
                slot[1] = len;
                cnt = 1;
            }
        }
        counts[pos] = cnt;
    }
    for (uint32_t pos = n; pos < sub; pos++) counts[pos] = 0; /* :188-190 */
}

/* ------------------------------------------------------------------------ */
/* Whole-buffer drivers, optionally over several threads (sub-blocks are     */
/* independent, as the reference grid of one block per sub-block).           */
/* ------------------------------------------------------------------------ */
typedef struct {
    int algo; /* 0 zstd, 1 lzma2 */
    const uint8_t* in;
    uint64_t total;
    uint32_t sub, depth, max_n;
    uint32_t* out;
    uint8_t* lit;
    uint32_t* cnt_a;
    uint32_t* cnt_b;
    uint64_t nsub;
    uint64_t next; /* shared work counter */
    pthread_mutex_t mu;
} orc_job;

static void orc_one(orc_job* j, uint64_t s) {
    uint64_t start = s * j->sub;
    uint32_t n = (uint32_t)((j->total - start) < j->sub ? (j->total - start) : j->sub);
    if (j->algo == 0) {
        orc_zstd_block(j->in + start, n, j->depth, j->max_n, j->out + s * (uint64_t)j->max_n * 3,
                       j->lit ? j->lit + start : NULL, &j->cnt_a[s], &j->cnt_b[s]);
    } else {
        orc_lzma2_block(j->in + start, n, j->sub, j->max_n, j->depth, j->out + s * (uint64_t)j->sub * j->max_n * 2,
                        j->cnt_a + s * (uint64_t)j->sub);
    }
}

static void* orc_worker(void* arg) {
    orc_job* j = (orc_job*)arg;
    for (;;) {
        pthread_mutex_lock(&j->mu);
        uint64_t s = j->next++;
        pthread_mutex_unlock(&j->mu);
        if (s >= j->nsub) break;
        orc_one(j, s);
    }
    return NULL;
}

static int orc_run(orc_job* j, int threads) {
    j->nsub = (j->total + j->sub - 1) / j->sub;
    j->next = 0;
    if (threads <= 1) {
        for (uint64_t s = 0; s < j->nsub; s++) orc_one(j, s);
        return 0;
    }
    pthread_mutex_init(&j->mu, NULL);
    pthread_t* t = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)threads);
    if (!t) return -1;
    for (int i = 0; i < threads; i++) pthread_create(&t[i], NULL, orc_worker, j);
    for (int i = 0; i < threads; i++) pthread_join(t[i], NULL);
    free(t);
    pthread_mutex_destroy(&j->mu);
    return 0;
}

/*
 * Zstd oracle over a whole buffer.  seqs: n_sub*max_seq*3 u32 (zeroed by the
 * caller, as alloc_zeros at compress_zstd_custom.rs:82); literals: n_sub*chunk
 * bytes or NULL; seq_counts/lit_counts: n_sub u32.
 */
int orc_zstd_match_find(const uint8_t* in, uint64_t total, uint32_t chunk, uint32_t depth, uint32_t max_seq,
                        uint32_t* seqs, uint8_t* literals, uint32_t* seq_counts, uint32_t* lit_counts, int threads) {
    if (chunk == 0 || chunk > 65536) return -1;
    orc_job j;
    memset(&j, 0, sizeof(j));
    j.algo = 0;
    j.in = in;
    j.total = total;
    j.sub = chunk;
    j.depth = depth;
    j.max_n = max_seq;
    j.out = seqs;
    j.lit = literals;
    j.cnt_a = seq_counts;
    j.cnt_b = lit_counts;
    return orc_run(&j, threads);
}

/*
 * LZMA2 oracle over a whole buffer.  matches: n_sub*sub*max_matches*2 u32,
 * interleaved {distance,length} and zeroed by the caller (alloc_zeros at
 * compress_lzma2_custom.rs:511); counts: n_sub*sub u32.
 */
int orc_lzma2_match_find(const uint8_t* in, uint64_t total, uint32_t sub, uint32_t max_matches, uint32_t depth,
                         uint32_t* matches, uint32_t* counts, int threads) {
    if (sub == 0 || sub > 65536 || max_matches == 0) return -1;
    orc_job j;
    memset(&j, 0, sizeof(j));
    j.algo = 1;
    j.in = in;
    j.total = total;
    j.sub = sub;
    j.depth = depth;
    j.max_n = max_matches;
    j.out = matches;
    j.cnt_a = counts;
    return orc_run(&j, threads);
}

/* ------------------------------------------------------------------------ */
/* Zstd frame of one sub-block from its sequences, with the format's         */
/* predefined FSE distributions and raw literals (RFC 8878 3.1.1.3.2.2,      */
/* 4.1.1).  Not a restatement of the reference (its GPU encoder,             */
/* zstd_compress.cu:756-1038, is dead code): the checker of the GPU encoder, */
/* itself checked by decoding its frames with libzstd.                       */
/* ------------------------------------------------------------------------ */

/* Literals_Length and Match_Length codes: baseline and extra bits. */
static const uint32_t orc_ll_base[36] = {0,  1,  2,   3,   4,   5,   6,    7,    8,    9,     10,    11,
                                         12, 13, 14,  15,  16,  18,  20,   22,   24,   28,    32,    40,
                                         48, 64, 128, 256, 512, 1024, 2048, 4096, 8192, 16384, 32768, 65536};
static const uint8_t orc_ll_bits[36] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,  0,  1,  1,
                                        1, 1, 2, 2, 3, 3, 4, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16};
static const uint32_t orc_ml_base[53] = {3,  4,  5,  6,  7,  8,   9,   10,  11,  12,   13,   14,   15,   16,
                                         17, 18, 19, 20, 21, 22,  23,  24,  25,  26,   27,   28,   29,   30,
                                         31, 32, 33, 34, 35, 37,  39,  41,  43,  47,   51,   59,   67,   83,
                                         99, 131, 259, 515, 1027, 2051, 4099, 8195, 16387, 32771, 65539};
static const uint8_t orc_ml_bits[53] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,
                                        0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1,
                                        2, 2, 3, 3, 4, 4, 5, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16};
/* Predefined distributions (-1: a "less than 1" probability, one cell). */
static const int16_t orc_ll_norm[36] = {4, 3, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 1, 1, 1, 2, 2,
                                        2, 2, 2, 2, 2, 2, 2, 3, 2, 1, 1, 1, 1, 1, -1, -1, -1, -1};
static const int16_t orc_ml_norm[53] = {1, 4, 3, 2, 2, 2, 2, 2, 2, 1, 1, 1, 1, 1, 1, 1, 1, 1,
                                        1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1,
                                        1, 1, 1, 1, 1, 1, 1, 1, 1, 1, -1, -1, -1, -1, -1, -1, -1};
static const int16_t orc_of_norm[29] = {1, 1, 1, 1, 1, 1, 2, 2, 2, 1, 1, 1, 1, 1, 1,
                                        1, 1, 1, 1, 1, 1, 1, 1, 1, -1, -1, -1, -1, -1};

/* An FSE table as the decoder builds it: per cell its symbol, bits to read
 * and baseline of the next state. */
typedef struct {
    uint32_t log, size;
    uint8_t sym[64], nb[64];
    uint16_t base[64];
} orc_fse;

static void orc_fse_build(orc_fse* t, const int16_t* norm, uint32_t nsym, uint32_t log) {
    const uint32_t size = 1u << log, mask = size - 1;
    uint32_t high = size - 1, next[64];
    t->log = log;
    t->size = size;
    for (uint32_t s = 0; s < nsym; s++) {
        if (norm[s] == -1) {
            t->sym[high--] = (uint8_t)s;
            next[s] = 1;
        } else {
            next[s] = (uint32_t)norm[s];
        }
    }
    const uint32_t step = (size >> 1) + (size >> 3) + 3;
    uint32_t pos = 0;
    for (uint32_t s = 0; s < nsym; s++)
        for (int k = 0; k < norm[s]; k++) {
            t->sym[pos] = (uint8_t)s;
            do pos = (pos + step) & mask;
            while (pos > high);
        }
    for (uint32_t c = 0; c < size; c++) {
        const uint32_t x = next[t->sym[c]]++;
        uint32_t hb = 31;
        while (!(x >> hb)) hb--;
        t->nb[c] = (uint8_t)(log - hb);
        t->base[c] = (uint16_t)((x << t->nb[c]) - size);
    }
}

/* The cell of symbol s whose next-state range holds `target`. */
static uint32_t orc_fse_cell(const orc_fse* t, uint32_t s, uint32_t target) {
    for (uint32_t c = 0; c < t->size; c++)
        if (t->sym[c] == s && t->base[c] <= target && target < (uint32_t)t->base[c] + (1u << t->nb[c])) return c;
    return 0xFFFFFFFFu;
}

typedef struct {
    uint8_t* p;
    size_t n, cap;
    uint64_t acc;
    uint32_t bits;
    int full;
} orc_bits;

static void orc_put_bits(orc_bits* b, uint32_t v, uint32_t nb) {
    b->acc |= (uint64_t)v << b->bits;
    b->bits += nb;
    while (b->bits >= 8) {
        if (b->n < b->cap)
            b->p[b->n++] = (uint8_t)b->acc;
        else
            b->full = 1;
        b->acc >>= 8;
        b->bits -= 8;
    }
}

static uint32_t orc_code(const uint32_t* base, uint32_t ncode, uint32_t v) {
    uint32_t c = 0;
    while (c + 1 < ncode && base[c + 1] <= v) c++;
    return c;
}

static uint32_t orc_highbit(uint32_t v) {
    uint32_t h = 0;
    while (v >> (h + 1)) h++;
    return h;
}

/*
 * One frame for src[0..n) (n <= 65536) from nseq ZSTD_Sequence rows
 * {offset, litLength, matchLength, rep} (offsets never coded as repeats:
 * Offset_Value = offset + 3).  Returns the frame size, or 0 if it does not
 * fit `cap`.
 */
size_t orc_zstd_frame_predef(const uint8_t* src, uint32_t n, const uint32_t* seqs, uint32_t nseq, uint8_t* out,
                             size_t cap) {
    orc_fse ll, ml, of;
    orc_fse_build(&ll, orc_ll_norm, 36, 6);
    orc_fse_build(&ml, orc_ml_norm, 53, 6);
    orc_fse_build(&of, orc_of_norm, 29, 5);
    if (cap < 16) return 0;
    size_t o = 0;
    out[o++] = 0x28, out[o++] = 0xB5, out[o++] = 0x2F, out[o++] = 0xFD;
    if (n >= 256) {
        out[o++] = 0x60; /* 2-byte content size, single segment */
        out[o++] = (uint8_t)(n - 256), out[o++] = (uint8_t)((n - 256) >> 8);
    } else {
        out[o++] = 0x20;
        out[o++] = (uint8_t)n;
    }
    const size_t bh = o; /* block header */
    o += 3;
    const size_t start = o;
    /* literals: every byte outside the matches, in order */
    uint32_t nlit = 0, p = 0;
    for (uint32_t i = 0; i < nseq; i++) {
        nlit += seqs[4 * i + 1];
        p += seqs[4 * i + 1] + seqs[4 * i + 2];
    }
    nlit += n - p;
    int compressed = 1;
    if (o + 3 + nlit + 3 > cap) compressed = 0;
    if (compressed) {
        if (nlit < 32) {
            out[o++] = (uint8_t)(nlit << 3);
        } else if (nlit < 4096) {
            out[o++] = (uint8_t)((nlit << 4) | 4), out[o++] = (uint8_t)(nlit >> 4);
        } else {
            out[o++] = (uint8_t)((nlit << 4) | 12), out[o++] = (uint8_t)(nlit >> 4), out[o++] = (uint8_t)(nlit >> 12);
        }
        p = 0;
        for (uint32_t i = 0; i < nseq; i++) {
            memcpy(out + o, src + p, seqs[4 * i + 1]);
            o += seqs[4 * i + 1];
            p += seqs[4 * i + 1] + seqs[4 * i + 2];
        }
        memcpy(out + o, src + p, n - p);
        o += n - p;
        /* sequences section */
        if (nseq < 128) {
            out[o++] = (uint8_t)nseq;
        } else {
            out[o++] = (uint8_t)((nseq >> 8) + 0x80), out[o++] = (uint8_t)nseq;
        }
        if (nseq) {
            out[o++] = 0x00; /* predefined modes */
            orc_bits b = {out + o, 0, cap - o, 0, 0, 0};
            uint32_t sl = 0, sm = 0, so = 0;
            for (int32_t i = (int32_t)nseq - 1; i >= 0; i--) {
                const uint32_t lit = seqs[4 * i + 1], mlen = seqs[4 * i + 2], ofv = seqs[4 * i] + 3;
                const uint32_t lc = orc_code(orc_ll_base, 36, lit), mc = orc_code(orc_ml_base, 53, mlen);
                const uint32_t oc = orc_highbit(ofv);
                if (i == (int32_t)nseq - 1) {
                    /* the last sequence's states: any cell of its symbols (the one reached from 0) */
                    sl = orc_fse_cell(&ll, lc, 0), sm = orc_fse_cell(&ml, mc, 0), so = orc_fse_cell(&of, oc, 0);
                } else {
                    const uint32_t cl = orc_fse_cell(&ll, lc, sl), cm = orc_fse_cell(&ml, mc, sm),
                                   co = orc_fse_cell(&of, oc, so);
                    /* state updates, read as LL, ML, OF: written in reverse */
                    orc_put_bits(&b, so - of.base[co], of.nb[co]);
                    orc_put_bits(&b, sm - ml.base[cm], ml.nb[cm]);
                    orc_put_bits(&b, sl - ll.base[cl], ll.nb[cl]);
                    sl = cl, sm = cm, so = co;
                }
                /* extra bits, read as OF, ML, LL: written in reverse */
                orc_put_bits(&b, lit - orc_ll_base[lc], orc_ll_bits[lc]);
                orc_put_bits(&b, mlen - orc_ml_base[mc], orc_ml_bits[mc]);
                orc_put_bits(&b, ofv - (1u << oc), oc);
            }
            /* initial states, read as LL, OF, ML */
            orc_put_bits(&b, sm, ml.log);
            orc_put_bits(&b, so, of.log);
            orc_put_bits(&b, sl, ll.log);
            orc_put_bits(&b, 1, 1); /* end mark */
            if (b.bits) orc_put_bits(&b, 0, 8 - b.bits);
            if (b.full) compressed = 0;
            o += b.n;
        }
        if (o - start >= n) compressed = 0;
    }
    if (!compressed) { /* raw block */
        if (start + n > cap) return 0;
        memcpy(out + start, src, n);
        o = start + n;
        const uint32_t h = 1u | (0u << 1) | (n << 3);
        out[bh] = (uint8_t)h, out[bh + 1] = (uint8_t)(h >> 8), out[bh + 2] = (uint8_t)(h >> 16);
        return o;
    }
    const uint32_t h = 1u | (2u << 1) | ((uint32_t)(o - start) << 3);
    out[bh] = (uint8_t)h, out[bh + 1] = (uint8_t)(h >> 8), out[bh + 2] = (uint8_t)(h >> 16);
    return o;
}
