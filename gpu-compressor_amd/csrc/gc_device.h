// gc_device.h -- records shared by the host library and the device kernels.
#pragma once
#include <stdint.h>

#include "../../include/gcompress.h"

// Threads per workgroup of the candidate kernel: one table wavefront and
// (GC_CAND_THREADS / 64 - 1) analysis wavefronts of 64 positions each.
#ifndef GC_CAND_THREADS
#define GC_CAND_THREADS 512
#endif
// The wide form of the candidate kernel (launches of few sub-blocks).
#define GC_CAND_WIDE_THREADS 1024

// Segmented Zstd parse: wavefronts (segments) per sub-block, and per
// workgroup two record buffers of GC_SEG_REC_CAP 8-byte records per segment
// (a segment is at most 65536 / GC_SEG_WAVES = 4096 bytes, so it holds at
// most 1366 sequences of 3 bytes or more).
#ifndef GC_SEG_WAVES
#define GC_SEG_WAVES 16u
#define GC_SEG_REC_CAP 1376u
#endif
#define GC_SEG_SCRATCH ((uint64_t)GC_SEG_WAVES * 2u * GC_SEG_REC_CAP * 8u)

// Sequence output formats of the Zstd parse kernel.
#define GC_SEQ_ZSTD 0u  // ZSTD_Sequence {offset, litLength, matchLength, rep}
#define GC_SEQ_REF 1u   // reference device record {literal_length, match_length - 3, offset}

// One sub-block of work, for every kernel.  Offsets are in bytes (src) or
// positions (cpos, opos): position p of this sub-block has its candidate at
// cand16[cpos + p] and its LZMA2 packed match / Zstd info word at
// out[opos + p].
typedef struct {
    uint64_t src_off;   // sub-block start in the input arena
    uint64_t avail;     // readable bytes from src_off to the end of the arena
    uint64_t cpos;      // first position in the candidate array
    uint64_t opos;      // first position in the per-position output
    uint64_t seq_pos;   // Zstd: first sequence slot (sub-block index * max_seq)
    uint64_t lit_pos;   // Zstd: first literal byte (sub-block index * chunk size)
    uint32_t count_pos; // Zstd: index into seq_counts / lit_counts
    uint32_t n;         // sub-block length, 1..65536
    uint32_t depth;     // Zstd: > 1 enables the lazy step (zstd_compress.cu:278); LZMA2: 0 = no probe
    uint32_t out_span;  // LZMA2 reference layout: positions whose counts are written (the full sub-block)
} gc_item;

// Device work queue of the persistent kernels: one counter set per kernel
// (phase).  A workgroup (the parse: a wavefront) takes item next[ph]++ until
// the value reaches n[ph]; the last one to leave checks that n[ph] items were
// taken and sets next, taken and done back to 0, so the queue is ready for
// the next launch without a reset kernel or memset.  The phase's items are
// items[base[ph] .. base[ph] + n[ph]) of the launch's item array, so a
// captured launch serves any split of a new item list.  err (never cleared by
// the device): bit 0 a loop passed its trip bound, bit 1 a phase did not take
// exactly n[ph] items.
#define GC_Q_CAND 0u
#define GC_Q_LZMA2 1u
#define GC_Q_INFO 2u
#define GC_Q_PARSE 3u
typedef struct {
    uint32_t next[4];
    uint32_t done[4];
    uint32_t taken[4];
    uint32_t n[4];
    uint32_t base[4];
    uint32_t err;
    uint32_t pad[11];
} gc_queue;
