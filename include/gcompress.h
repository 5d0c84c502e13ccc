/*
 * gcompress.h -- C ABI of the MI355X match-finding library (libgcompress.so).
 *
 * Drop-in for the match-finding path of gpu-compressor: the two device
 * kernels, their two Rust host functions, the libzstd / liblzma consumers of
 * their output and the NVZS / NVLZ containers.  Every entry point names the
 * reference interface it replaces (file:line in the reference tree).
 *
 * Conventions
 *   - The caller owns every buffer.  Host pointers are plain memory; "_dev"
 *     entry points take device pointers (4-byte aligned inputs) and run on
 *     the context's stream, returning once the work is queued unless stated.
 *   - Return 0 on success or a negative GC_E* code; gc_last_error() gives a
 *     message for the calling thread.
 *   - A context is used by one host thread at a time.
 *   - All arithmetic is integer; results are bit-identical to the
 *     reference kernels for sub-block sizes 1..65536.
 */
#ifndef GCOMPRESS_H
#define GCOMPRESS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- records ---------------------------------------------------------- */

/* libzstd ZSTD_Sequence (1.4.x); compress_zstd_custom.rs:8-14 (ZstdSequence). */
typedef struct {
    uint32_t offset;
    uint32_t litLength;
    uint32_t matchLength;
    uint32_t rep;
} gc_zstd_sequence;

/* Device sequence record of zstd_compress.cu:38-42 (match_length stored - 3). */
typedef struct {
    uint32_t literal_length;
    uint32_t match_length_minus3;
    uint32_t offset;
} gc_ref_sequence;

/* lzma2_match_find.cu:26-29 (LzmaMatch). */
typedef struct {
    uint32_t distance;
    uint32_t length;
} gc_lzma_match;

/* ---- codes and constants ---------------------------------------------- */

#define GC_OK 0
#define GC_EINVAL (-1)  /* bad argument */
#define GC_EHIP (-2)    /* HIP runtime error (device, load, launch, copy) */
#define GC_ENOMEM (-3)  /* allocation failed */
#define GC_ELIB (-4)    /* libzstd / liblzma missing or failed */
#define GC_EFORMAT (-5) /* malformed NVZS / NVLZ container */
#define GC_ESIZE (-6)   /* output buffer too small */
#define GC_EBUDGET (-7) /* request larger than the admission budget */

/* constants.rs:5-6,10-13 */
#define GC_SUB_BLOCK 65536u
#define GC_ZSTD_DEPTH_LAZY 16u
#define GC_LZMA2_MAX_MATCHES 8u
#define GC_LZMA2_DEPTH 32u

/* seq_format values for gc_zstd_match_find_dev */
#define GC_SEQ_FORMAT_ZSTD 0u /* gc_zstd_sequence (16 B) */
#define GC_SEQ_FORMAT_REF 1u  /* gc_ref_sequence (12 B) */

typedef struct gc_ctx gc_ctx;

/* ---- context ------------------------------------------------------------ */

/* Device + stream + code object loaded once with hipModuleLoadData.
 * Replaces the per-call cudaSetDevice / CudaDevice::new / load_ptx of
 * compress_zstd_custom.rs:63-77 and compress_lzma2_custom.rs:492-506.
 * NULL on failure (see gc_last_error). */
gc_ctx* gc_open(int device);
void gc_close(gc_ctx* ctx);
const char* gc_last_error(void);
/* cuda.rs:16-45 (detect_gpus): number of visible devices, or < 0. */
int gc_device_count(void);

/* Device memory helpers, so that a caller without its own GPU runtime can
 * stage data for the _dev entry points.  Copies are synchronous. */
void* gc_dev_alloc(gc_ctx* ctx, size_t bytes);
int gc_dev_free(gc_ctx* ctx, void* p);
int gc_htod(gc_ctx* ctx, void* dst, const void* src, size_t bytes);
int gc_dtoh(gc_ctx* ctx, void* dst, const void* src, size_t bytes);
int gc_dev_zero(gc_ctx* ctx, void* p, size_t bytes);
int gc_sync(gc_ctx* ctx);
/* The context's HIP stream (hipStream_t), for callers that order their own
 * work against it. */
void* gc_stream(gc_ctx* ctx);

/* ---- kernel-level entry points (device pointers) ---------------------- */

/* zstd_match_find (zstd_compress.cu:179-190): sequences of every sub-chunk
 * of d_input[0, total).  d_sequences holds num_chunks * max_seq records in
 * seq_format; d_literals (num_chunks * chunk_size bytes) and d_lit_counts
 * may be NULL.  Slots past seq_counts[s] are left untouched. */
int gc_zstd_match_find_dev(gc_ctx* ctx, const uint8_t* d_input, void* d_sequences, uint32_t seq_format,
                           uint8_t* d_literals, uint32_t* d_seq_counts, uint32_t* d_lit_counts, uint32_t chunk_size,
                           uint32_t num_chunks, uint64_t total_input_size, uint32_t search_depth,
                           uint32_t max_sequences_per_chunk);

/* lzma2_match_find (lzma2_match_find.cu:54-63): writes slot 0 of
 * d_matches[(s*sub + p)*max_matches] where a match exists and every
 * d_match_counts[s*sub + p]; other slots are left untouched (the reference
 * caller zeroes them, compress_lzma2_custom.rs:510-512). */
int gc_lzma2_match_find_dev(gc_ctx* ctx, const uint8_t* d_input, gc_lzma_match* d_matches, uint32_t* d_match_counts,
                            uint32_t sub_block_size, uint32_t num_sub_blocks, uint64_t total_input_size,
                            uint32_t max_matches_per_pos, uint32_t search_depth);

/* Compact form of lzma2_match_find: one u32 per input position,
 * (distance << 16) | length, 0 where there is no match.  Same matches as
 * above; 4 bytes per position instead of 68. */
int gc_lzma2_match_find_packed_dev(gc_ctx* ctx, const uint8_t* d_input, uint32_t* d_packed, uint32_t sub_block_size,
                                   uint64_t total_input_size, uint32_t search_depth);

/* ---- host-level entry points (host pointers, synchronous) ------------- */

/* gpu_find_matches (compress_lzma2_custom.rs:483-559) with GpuMatchData's
 * layout (:472-479): n_sub = ceil(n / 65536); match_distances and
 * match_lengths hold n_sub*65536*8 u32 and match_counts n_sub*65536 u32.
 * All three are fully written (unused slots 0). */
int gc_gpu_find_matches(gc_ctx* ctx, const uint8_t* chunk, size_t n, uint32_t* match_distances,
                        uint32_t* match_lengths, uint32_t* match_counts);

/* Same matches, compact: packed[n] as in gc_lzma2_match_find_packed_dev. */
int gc_lzma2_find_matches_packed(gc_ctx* ctx, const uint8_t* chunk, size_t n, uint32_t* packed);

/* The sequences of compress_chunk_zstd_custom (compress_zstd_custom.rs:
 * 79-153) before entropy coding: seqs holds n_sub * 21845 records
 * (sub-chunk s at s*21845), seq_counts n_sub.  level: 1 -> depth 16,
 * 2 -> depth 64, otherwise 1 (:57-61). */
int gc_zstd_find_sequences(gc_ctx* ctx, const uint8_t* chunk, size_t n, uint32_t level, gc_zstd_sequence* seqs,
                           uint32_t* seq_counts);

#ifdef __cplusplus
}
#endif
#endif /* GCOMPRESS_H */
