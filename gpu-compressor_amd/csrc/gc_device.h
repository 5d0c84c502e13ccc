// gc_device.h -- records shared by the host library and the device kernels.
#pragma once
#include <stdint.h>

#include "../../include/gcompress.h"

// Sequence output formats of the Zstd parse kernel.
#define GC_SEQ_ZSTD 0u  // ZSTD_Sequence {offset, litLength, matchLength, rep}
#define GC_SEQ_REF 1u   // reference device record {literal_length, match_length - 3, offset}

// One LZMA2 sub-block of work.  Offsets are in bytes (src) or positions
// (out): the packed output of position p of this sub-block is at
// packed[out_pos + p]; the reference-layout counts span out_span positions
// (the full sub-block size, so that a short last sub-block zeroes its tail).
typedef struct {
    uint64_t src_off;  // sub-block start in the input arena
    uint64_t avail;    // readable bytes from src_off to the end of the arena
    uint64_t out_pos;  // first output position
    uint32_t n;        // sub-block length, 1..65536
    uint32_t out_span; // positions covered in the reference-layout counts
    uint32_t depth;    // probe depth (0 = no hash probe), lzma2_match_find.cu:63
    uint32_t pad;
} gc_lzma2_item;

// One Zstd sub-block of work.
typedef struct {
    uint64_t src_off;     // sub-block start in the input arena
    uint64_t avail;       // readable bytes from src_off to the end of the arena
    uint64_t scratch_pos; // per-position candidate words
    uint64_t seq_pos;     // first sequence slot (sub-block index * max_seq)
    uint64_t lit_pos;     // first literal byte (sub-block index * chunk size)
    uint32_t count_pos;   // index into seq_counts / lit_counts
    uint32_t n;           // sub-block length, 1..65536
    uint32_t depth;       // search depth (> 1 enables the lazy step), zstd_compress.cu:188
    uint32_t pad;
} gc_zstd_item;
