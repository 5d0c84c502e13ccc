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
/* Zstd frame of one sub-block from its sequences (RFC 8878): the checker   */
/* of the GPU frame coder (gc_zstd_encode_kernel), itself checked by        */
/* decoding its frames with libzstd.  What the coder chooses follows the    */
/* capability of the reference's GPU coder (zstd_compress.cu:529-740,      */
/* 835-903: per dimension an RLE or an FSE_Compressed table from the        */
/* sub-block's own code histogram), widened:                                */
/*  - per dimension the cheapest of predefined, RLE and FSE_Compressed       */
/*    (estimated bits, tables with an accuracy log from the sequence count); */
/*  - offsets as repeat codes where the format's offset history allows;     */
/*  - literals Huffman-coded (4 streams, weights written directly) or RLE   */
/*    when that is smaller than raw.                                        */
/* The format's rules are kept as the format states them; the choices (log, */
/* normalisation, code lengths) are this coder's own and are the same in    */
/* the GPU coder, byte for byte.                                            */
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

typedef struct {
    uint8_t* p;
    size_t n, cap;
    uint64_t acc;
    uint32_t bits;
    int full;
} orc_bits;

static void orc_put_bits(orc_bits* b, uint32_t v, uint32_t nb) {
    if (nb < 32) v &= (1u << nb) - 1;
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

/* the partial byte, zero-padded */
static void orc_bits_pad(orc_bits* b) {
    if (b->bits) orc_put_bits(b, 0, 8 - b->bits);
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

/* 256 x log2(x) for x >= 1, linear between powers of two (the coder's cost
 * estimates; the GPU coder computes the same integers) */
static uint32_t orc_log256(uint32_t x) {
    const uint32_t h = orc_highbit(x);
    return (h << 8) + (((x << 8) >> h) - 256u);
}

/* FSE encoding table (RFC 8878 4.1.1: the decoder's cells, spread as the
 * decoder spreads them, seen from the encoder): st[] the next states of
 * each symbol's cells in cell order (state = table size + cell), per symbol
 * dnb (bits out: (state + dnb) >> 16) and dfs (first next-state index - its
 * normalized count).  Symbols in [0, nsym). */
typedef struct {
    uint32_t log;
    uint16_t st[512];
    uint32_t dnb[64];
    int32_t dfs[64];
} orc_ctab;

static void orc_ctab_build(orc_ctab* t, const int16_t* nc, uint32_t nsym, uint32_t log) {
    const uint32_t size = 1u << log, mask = size - 1;
    uint8_t cell[512];
    uint32_t high = size - 1, next[65];
    for (uint32_t s = 0; s < nsym; s++)
        if (nc[s] == -1) cell[high--] = (uint8_t)s;
    const uint32_t step = (size >> 1) + (size >> 3) + 3;
    uint32_t pos = 0;
    for (uint32_t s = 0; s < nsym; s++)
        for (int k = 0; k < nc[s]; k++) {
            cell[pos] = (uint8_t)s;
            do pos = (pos + step) & mask;
            while (pos > high);
        }
    next[0] = 0;
    for (uint32_t s = 0; s < nsym; s++) next[s + 1] = next[s] + (nc[s] == -1 ? 1u : (nc[s] > 0 ? (uint32_t)nc[s] : 0u));
    for (uint32_t u = 0; u < size; u++) t->st[next[cell[u]]++] = (uint16_t)(size + u);
    int32_t total = 0;
    for (uint32_t s = 0; s < nsym; s++) {
        if (nc[s] == 0) {
            t->dnb[s] = ((log + 1) << 16) - size, t->dfs[s] = 0;
        } else if (nc[s] == -1 || nc[s] == 1) {
            t->dnb[s] = (log << 16) - size, t->dfs[s] = total - 1, total += 1;
        } else {
            const uint32_t mb = log - orc_highbit((uint32_t)nc[s] - 1);
            t->dnb[s] = (mb << 16) - ((uint32_t)nc[s] << mb), t->dfs[s] = total - nc[s], total += nc[s];
        }
    }
    t->log = log;
}

/* the state of a first-coded symbol (no bits out) */
static uint32_t orc_fse_init(const orc_ctab* t, uint32_t s) {
    const uint32_t nb = (t->dnb[s] + (1u << 15)) >> 16;
    const uint32_t v = (nb << 16) - t->dnb[s];
    return t->st[(int32_t)(v >> nb) + t->dfs[s]];
}

static void orc_fse_put(orc_bits* b, const orc_ctab* t, uint32_t* state, uint32_t s) {
    const uint32_t nb = (*state + t->dnb[s]) >> 16;
    orc_put_bits(b, *state, nb);
    *state = t->st[(int32_t)(*state >> nb) + t->dfs[s]];
}

/* Normalized counts summing to 1 << log: each present symbol's share
 * rounded (at least 1), then the symbol with the largest count (the first
 * such) takes the difference one cell at a time (a decrement only from a
 * symbol with more than one cell). */
static void orc_normalize(const uint32_t* hist, uint32_t nsym, uint32_t total, uint32_t log, int16_t* nc) {
    const uint32_t size = 1u << log;
    int32_t dist = 0;
    for (uint32_t s = 0; s < nsym; s++) {
        if (!hist[s]) {
            nc[s] = 0;
            continue;
        }
        uint32_t v = (uint32_t)(((uint64_t)hist[s] * size + total / 2) / total);
        if (!v) v = 1;
        nc[s] = (int16_t)v;
        dist += (int32_t)v;
    }
    while (dist != (int32_t)size) {
        const int up = dist < (int32_t)size;
        uint32_t best = 0, bc = 0;
        for (uint32_t s = 0; s < nsym; s++)
            if (hist[s] > bc && (up || nc[s] > 1)) bc = hist[s], best = s;
        nc[best] += up ? 1 : -1;
        dist += up ? 1 : -1;
    }
}

/* Accuracy log of a table for nseq symbols of which npresent differ. */
static uint32_t orc_fse_log(uint32_t nseq, uint32_t npresent, uint32_t maxlog) {
    uint32_t log = orc_highbit(nseq);
    log = log > 7 ? log - 2 : 5;
    const uint32_t need = orc_highbit(npresent - 1) + 2;  /* at least twice the cells of the symbols */
    if (log < need) log = need;
    return log < maxlog ? log : maxlog;
}

/* FSE table description (RFC 8878 4.1.1) of nc[0, nsym), nc[nsym - 1] != 0. */
static void orc_ncount(orc_bits* b, const int16_t* nc, uint32_t nsym, uint32_t log) {
    orc_put_bits(b, log - 5, 4);
    int32_t remaining = (1 << log) + 1, threshold = 1 << log;
    uint32_t nb = log + 1, s = 0;
    int prev0 = 0;
    while (s < nsym && remaining > 1) {
        if (prev0) {  /* zeros after a zero: 2-bit repeat flags, 3 = three more and go on */
            uint32_t z = 0;
            while (s < nsym && nc[s] == 0) s++, z++;
            while (z >= 3) orc_put_bits(b, 3, 2), z -= 3;
            orc_put_bits(b, z, 2);
        }
        const int32_t count = nc[s++];
        const int32_t max = (2 * threshold - 1) - remaining;
        remaining -= count < 0 ? -count : count;
        int32_t v = count + 1;
        if (v >= threshold) v += max;
        orc_put_bits(b, (uint32_t)v, nb - (v < max ? 1u : 0u));
        prev0 = count == 0;
        while (remaining < threshold) nb--, threshold >>= 1;
    }
    orc_bits_pad(b);
}

/* Huffman code lengths (at most 11 bits) of the present bytes; returns the
 * longest.  Two-queue Huffman over the symbols sorted by (count, symbol),
 * a leaf before an internal node of equal weight; depths over 11 are cut
 * to 11 and the code made complete again by moving one symbol at a time:
 * while it is over-full, from the longest length under 11 one bit longer;
 * while it is not full, from the longest length one bit shorter.  The
 * lengths then go to the symbols in that sorted order, the longest to the
 * least frequent. */
static uint32_t orc_huff_lengths(const uint32_t* hist, uint8_t* len) {
    uint16_t srt[256];
    uint32_t n = 0;
    for (uint32_t s = 0; s < 256; s++) {
        len[s] = 0;
        if (hist[s]) srt[n++] = (uint16_t)s;
    }
    for (uint32_t i = 1; i < n; i++) { /* insertion sort by (count, symbol) */
        const uint16_t v = srt[i];
        uint32_t j = i;
        while (j > 0 && hist[srt[j - 1]] > hist[v]) srt[j] = srt[j - 1], j--;
        srt[j] = v;
    }
    uint32_t nw[256], lp[256] = {0}, np[256], dn[256] = {0};
    uint32_t li = 0, ni = 0, nn = 0;
    for (uint32_t k = 0; k + 1 < n; k++) {
        uint32_t w = 0;
        for (int pick = 0; pick < 2; pick++) {
            if (li < n && (ni >= nn || hist[srt[li]] <= nw[ni]))
                w += hist[srt[li]], lp[li++] = nn;
            else
                w += nw[ni], np[ni++] = nn;
        }
        nw[nn++] = w;
    }
    dn[nn - 1] = 0;
    for (int32_t j = (int32_t)nn - 2; j >= 0; j--) dn[j] = dn[np[j]] + 1;
    uint32_t blc[12] = {0};
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t d = dn[lp[i]] + 1;
        blc[d < 11 ? d : 11]++;
    }
    uint32_t k2 = 0;
    for (uint32_t l = 1; l <= 11; l++) k2 += blc[l] << (11 - l);
    while (k2 > 2048) {
        uint32_t l = 10;
        while (blc[l] == 0) l--;
        blc[l]--, blc[l + 1]++, k2 -= 1u << (10 - l);
    }
    while (k2 < 2048) {
        uint32_t l = 11;
        while (blc[l] == 0) l--;
        blc[l]--, blc[l - 1]++, k2 += 1u << (11 - l);
    }
    uint32_t maxlen = 11;
    while (blc[maxlen] == 0) maxlen--;
    uint32_t i = 0;
    for (uint32_t l = maxlen; l >= 1; l--)
        for (uint32_t c = 0; c < blc[l]; c++) len[srt[i++]] = (uint8_t)l;
    return maxlen;
}

/* Literals section (RFC 8878 3.1.1.3.1) for lit[0, nlit): Huffman (4
 * streams, weights written directly) when the bytes are all below 129 and
 * that is smaller than raw, RLE when all bytes are one, else raw.  Returns
 * its size, 0 when it does not fit cap. */
static size_t orc_literals(const uint8_t* lit, uint32_t nlit, uint8_t* out, size_t cap) {
    const uint32_t rh = nlit < 32 ? 1u : (nlit < 4096 ? 2u : 3u);
    uint32_t hist[256] = {0}, nz = 0, maxs = 0;
    for (uint32_t i = 0; i < nlit; i++) hist[lit[i]]++;
    for (uint32_t s = 0; s < 256; s++)
        if (hist[s]) nz++, maxs = s;
    uint32_t type = 0;  /* raw */
    if (nlit > 0 && nz == 1) type = 1;
    if (nlit >= 64 && nz >= 2 && maxs <= 128) {
        uint8_t len[256];
        const uint32_t maxlen = orc_huff_lengths(hist, len);
        uint32_t w[256], cntw[13] = {0}, start[13], code[256];
        for (uint32_t s = 0; s < 256; s++) {
            w[s] = len[s] ? maxlen + 1 - len[s] : 0;
            cntw[w[s]]++;
        }
        uint32_t pos = 0;
        for (uint32_t k = 1; k <= maxlen; k++) start[k] = pos, pos += cntw[k] << (k - 1);
        for (uint32_t s = 0; s < 256; s++)
            if (w[s]) code[s] = start[w[s]] >> (w[s] - 1), start[w[s]] += 1u << (w[s] - 1);
        /* the streams into a scratch copy */
        uint8_t* buf = (uint8_t*)malloc(nlit * 2 + 64);
        const uint32_t tree = 1 + (maxs + 1) / 2, seg = (nlit + 3) / 4;
        size_t o = tree + 6, ssz[4];
        buf[0] = (uint8_t)(127 + maxs);
        for (uint32_t s = 0; s < maxs; s += 2) buf[1 + s / 2] = (uint8_t)((w[s] << 4) | (s + 1 < maxs ? w[s + 1] : 0u));
        for (uint32_t k = 0; k < 4; k++) {
            const uint32_t a = k * seg, e = k == 3 ? nlit : (k + 1) * seg;
            orc_bits b = {buf + o, 0, nlit * 2 + 64 - o, 0, 0, 0};
            for (uint32_t i = e; i-- > a;) orc_put_bits(&b, code[lit[i]], len[lit[i]]);
            orc_put_bits(&b, 1, 1);
            orc_bits_pad(&b);
            ssz[k] = b.n;
            o += b.n;
        }
        for (uint32_t k = 0; k < 3; k++) buf[tree + 2 * k] = (uint8_t)ssz[k], buf[tree + 2 * k + 1] = (uint8_t)(ssz[k] >> 8);
        const size_t cs = o;
        const uint32_t hh = (nlit <= 1023 && cs <= 1023) ? 3u : ((nlit <= 16383 && cs <= 16383) ? 4u : 5u);
        if (hh + cs < rh + nlit && hh + cs <= cap) {
            uint64_t h;
            if (hh == 3)
                h = 2u | (1u << 2) | ((uint64_t)nlit << 4) | ((uint64_t)cs << 14);
            else if (hh == 4)
                h = 2u | (2u << 2) | ((uint64_t)nlit << 4) | ((uint64_t)cs << 18);
            else
                h = 2u | (3u << 2) | ((uint64_t)nlit << 4) | ((uint64_t)cs << 22);
            for (uint32_t k = 0; k < hh; k++) out[k] = (uint8_t)(h >> (8 * k));
            memcpy(out + hh, buf, cs);
            free(buf);
            return hh + cs;
        }
        free(buf);
    }
    const uint32_t body = type == 1 ? 1u : nlit;
    if (rh + body > cap) return 0;
    if (rh == 1)
        out[0] = (uint8_t)(type | (nlit << 3));
    else if (rh == 2)
        out[0] = (uint8_t)(type | 4u | (nlit << 4)), out[1] = (uint8_t)(nlit >> 4);
    else
        out[0] = (uint8_t)(type | 12u | (nlit << 4)), out[1] = (uint8_t)(nlit >> 4), out[2] = (uint8_t)(nlit >> 12);
    if (type == 1)
        out[rh] = lit[0];
    else
        memcpy(out + rh, lit, nlit);
    return rh + body;
}

/* Per-dimension choice: 0 predefined, 1 RLE, 2 FSE_Compressed (nc, log). */
static uint32_t orc_choose(const uint32_t* hist, uint32_t nsym, uint32_t nseq, const int16_t* pnorm, uint32_t plog,
                           uint32_t maxlog, int16_t* nc, uint32_t* log, uint32_t* rle_sym) {
    uint32_t npresent = 0, maxs = 0;
    for (uint32_t s = 0; s < nsym; s++)
        if (hist[s]) npresent++, maxs = s;
    if (npresent == 1) {
        *rle_sym = maxs;
        return 1;
    }
    uint64_t cp = 0, cc = 0;
    for (uint32_t s = 0; s <= maxs; s++)
        if (hist[s]) cp += (uint64_t)hist[s] * ((plog << 8) - orc_log256(pnorm[s] < 0 ? 1u : (uint32_t)pnorm[s]));
    *log = orc_fse_log(nseq, npresent, maxlog);
    orc_normalize(hist, maxs + 1, nseq, *log, nc);
    uint8_t tmp[128];
    orc_bits b = {tmp, 0, sizeof tmp, 0, 0, 0};
    orc_ncount(&b, nc, maxs + 1, *log);
    cc = (uint64_t)b.n * 8 * 256;
    for (uint32_t s = 0; s <= maxs; s++)
        if (hist[s]) cc += (uint64_t)hist[s] * ((*log << 8) - orc_log256((uint32_t)nc[s]));
    for (uint32_t s = maxs + 1; s < nsym; s++) nc[s] = 0;
    return cc < cp ? 2u : 0u;
}

/*
 * One frame for src[0..n) (n <= 65536) from nseq ZSTD_Sequence rows
 * {offset, litLength, matchLength, rep}.  Returns the frame size, or 0 if
 * it does not fit `cap` (at least n + 16 always fits: a raw block).
 */
size_t orc_zstd_frame(const uint8_t* src, uint32_t n, const uint32_t* seqs, uint32_t nseq, uint8_t* out, size_t cap) {
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
    uint8_t* lit = (uint8_t*)malloc(n + 1);
    uint32_t nlit = 0, p = 0;
    for (uint32_t i = 0; i < nseq; i++) {
        memcpy(lit + nlit, src + p, seqs[4 * i + 1]);
        nlit += seqs[4 * i + 1];
        p += seqs[4 * i + 1] + seqs[4 * i + 2];
    }
    memcpy(lit + nlit, src + p, n - p);
    nlit += n - p;
    const size_t limit = start + n; /* a compressed block ends before this */
    int compressed = nlit + 2 < n;
    if (compressed) {
        const size_t ls = orc_literals(lit, nlit, out + o, limit > o ? limit - o : 0);
        if (!ls) compressed = 0;
        o += ls;
    }
    free(lit);
    if (compressed && o + 4 > limit) compressed = 0;
    if (compressed) {
        if (nseq < 128) {
            out[o++] = (uint8_t)nseq;
        } else {
            out[o++] = (uint8_t)((nseq >> 8) + 0x80), out[o++] = (uint8_t)nseq;
        }
    }
    if (compressed && nseq) {
        /* codes, with offsets as repeat codes where the history allows */
        uint8_t* lc = (uint8_t*)malloc(3 * nseq);
        uint8_t *mc = lc + nseq, *oc = lc + 2 * nseq;
        uint32_t* ov = (uint32_t*)malloc(4 * nseq);
        uint32_t hl[36] = {0}, hm[53] = {0}, ho[32] = {0}, r0 = 1, r1 = 4, r2 = 8;
        for (uint32_t i = 0; i < nseq; i++) {
            const uint32_t off = seqs[4 * i], ll = seqs[4 * i + 1];
            uint32_t v;
            if (ll > 0) {
                if (off == r0) v = 1;
                else if (off == r1) v = 2, r1 = r0, r0 = off;
                else if (off == r2) v = 3, r2 = r1, r1 = r0, r0 = off;
                else v = off + 3, r2 = r1, r1 = r0, r0 = off;
            } else {
                if (off == r1) v = 1, r1 = r0, r0 = off;
                else if (off == r2) v = 2, r2 = r1, r1 = r0, r0 = off;
                else if (off == r0 - 1) v = 3, r2 = r1, r1 = r0, r0 = off;
                else v = off + 3, r2 = r1, r1 = r0, r0 = off;
            }
            ov[i] = v;
            lc[i] = (uint8_t)orc_code(orc_ll_base, 36, ll);
            mc[i] = (uint8_t)orc_code(orc_ml_base, 53, seqs[4 * i + 2]);
            oc[i] = (uint8_t)orc_highbit(v);
            hl[lc[i]]++, hm[mc[i]]++, ho[oc[i]]++;
        }
        int16_t nl[36], nm[53], no[32];
        uint32_t gl = 6, gm = 6, go = 5, sl_rle = 0, sm_rle = 0, so_rle = 0;
        const uint32_t ml_ = orc_choose(hm, 53, nseq, orc_ml_norm, 6, 9, nm, &gm, &sm_rle);
        const uint32_t ll_ = orc_choose(hl, 36, nseq, orc_ll_norm, 6, 9, nl, &gl, &sl_rle);
        const uint32_t of_ = orc_choose(ho, 29, nseq, orc_of_norm, 5, 8, no, &go, &so_rle);
        orc_ctab tl, tm, to;
        if (ll_ == 2) orc_ctab_build(&tl, nl, 36, gl); else if (ll_ == 0) orc_ctab_build(&tl, orc_ll_norm, 36, 6);
        if (ml_ == 2) orc_ctab_build(&tm, nm, 53, gm); else if (ml_ == 0) orc_ctab_build(&tm, orc_ml_norm, 53, 6);
        if (of_ == 2) orc_ctab_build(&to, no, 29, go); else if (of_ == 0) orc_ctab_build(&to, orc_of_norm, 29, 5);
        orc_bits b = {out + o, 0, limit > o ? limit - o : 0, 0, 0, 0};
        orc_put_bits(&b, (ll_ << 6) | (of_ << 4) | (ml_ << 2), 8);
        /* table descriptions, LL, OF, ML */
        if (ll_ == 1) orc_put_bits(&b, sl_rle, 8);
        if (ll_ == 2) orc_ncount(&b, nl, 36, gl);
        if (of_ == 1) orc_put_bits(&b, so_rle, 8);
        if (of_ == 2) orc_ncount(&b, no, 29, go);
        if (ml_ == 1) orc_put_bits(&b, sm_rle, 8);
        if (ml_ == 2) orc_ncount(&b, nm, 53, gm);
        uint32_t sl = 0, sm = 0, so = 0;
        for (int32_t i = (int32_t)nseq - 1; i >= 0; i--) {
            const uint32_t ll = seqs[4 * i + 1], mlen = seqs[4 * i + 2];
            if (i == (int32_t)nseq - 1) {
                if (ll_ != 1) sl = orc_fse_init(&tl, lc[i]);
                if (ml_ != 1) sm = orc_fse_init(&tm, mc[i]);
                if (of_ != 1) so = orc_fse_init(&to, oc[i]);
            } else {
                if (of_ != 1) orc_fse_put(&b, &to, &so, oc[i]);
                if (ml_ != 1) orc_fse_put(&b, &tm, &sm, mc[i]);
                if (ll_ != 1) orc_fse_put(&b, &tl, &sl, lc[i]);
            }
            orc_put_bits(&b, ll - orc_ll_base[lc[i]], orc_ll_bits[lc[i]]);
            orc_put_bits(&b, mlen - orc_ml_base[mc[i]], orc_ml_bits[mc[i]]);
            orc_put_bits(&b, ov[i] - (1u << oc[i]), oc[i]);
        }
        if (ml_ != 1) orc_put_bits(&b, sm, tm.log);
        if (of_ != 1) orc_put_bits(&b, so, to.log);
        if (ll_ != 1) orc_put_bits(&b, sl, tl.log);
        orc_put_bits(&b, 1, 1); /* end mark */
        orc_bits_pad(&b);
        if (b.full) compressed = 0;
        o += b.n;
        free(lc);
        free(ov);
    }
    if (compressed && o >= limit) compressed = 0;
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
