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
#define GC_EIO (-8)     /* file could not be opened, read or written */

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
/* No counterpart in the reference (one parse kernel there): the parse form
 * of the per-call Zstd entry points.  A call of at most max_sub_blocks
 * sub-chunks splits each sub-chunk's parse over 16 wavefronts that meet in
 * fix-up rounds (the segmented parse, the same sequences); a larger call
 * fills the GPU with one serial walk per wavefront.  0: never segmented.
 * Default GC_SEG_DEFAULT_MAX_SUB. */
#define GC_SEG_DEFAULT_MAX_SUB 2048u
int gc_set_segmented_parse(gc_ctx* ctx, uint32_t max_sub_blocks);

/* ---- kernel-level entry points (device pointers) ---------------------- */

/* zstd_match_find (zstd_compress.cu:179-190): sequences of every sub-chunk
 * of d_input[0, total).  d_sequences holds num_chunks * max_seq records in
 * seq_format; d_literals (num_chunks * chunk_size bytes) and d_lit_counts
 * may be NULL.  Slots past seq_counts[s] are left untouched. */
int gc_zstd_match_find_dev(gc_ctx* ctx, const uint8_t* d_input, void* d_sequences, uint32_t seq_format,
                           uint8_t* d_literals, uint32_t* d_seq_counts, uint32_t* d_lit_counts, uint32_t chunk_size,
                           uint32_t num_chunks, uint64_t total_input_size, uint32_t search_depth,
                           uint32_t max_sequences_per_chunk);

/* No counterpart in the reference: the same as gc_zstd_match_find_dev, for
 * callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  It allocates nothing, copies
 * nothing and does not synchronise: its scratch is one fixed device array per
 * GPU, held until the process ends, so num_chunks may be at most
 * GC_NOALLOC_MAX_SUB_BLOCKS and num_chunks * round_up(chunk_size, 8) at most
 * GC_NOALLOC_MAX_POSITIONS (8 MiB of input in 64 KiB sub-chunks); larger
 * calls fail with GC_EINVAL.  Calls that overlap on one GPU (other streams or
 * threads) share that array and race; calls ordered on one stream do not. */
#define GC_NOALLOC_MAX_SUB_BLOCKS 4096u
#define GC_NOALLOC_MAX_POSITIONS (8u << 20)
int gc_zstd_match_find_dev_noalloc(gc_ctx* ctx, const uint8_t* d_input, void* d_sequences, uint32_t seq_format,
                                   uint8_t* d_literals, uint32_t* d_seq_counts, uint32_t* d_lit_counts,
                                   uint32_t chunk_size, uint32_t num_chunks, uint64_t total_input_size,
                                   uint32_t search_depth, uint32_t max_sequences_per_chunk);

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
/* Device-side triage: any_match[s] = 1 if some position of 64 KiB sub-block s
 * has an LZMA2 match, else 0 (what encode_single_sub_block asks of the match
 * lists, compress_lzma2_custom.rs:577).  No lengths are measured and one byte
 * per sub-block comes back instead of the lists. */
int gc_lzma2_triage(gc_ctx* ctx, const uint8_t* chunk, size_t n, uint8_t* any_match);
int gc_lzma2_triage_dev(gc_ctx* ctx, const uint8_t* d_input, uint64_t n, uint8_t* d_any_match);

/* The sequences of compress_chunk_zstd_custom (compress_zstd_custom.rs:
 * 79-153) before entropy coding: seqs holds n_sub * 21845 records
 * (sub-chunk s at s*21845), seq_counts n_sub.  level: 1 -> depth 16,
 * 2 -> depth 64, otherwise 1 (:57-61). */
int gc_zstd_find_sequences(gc_ctx* ctx, const uint8_t* chunk, size_t n, uint32_t level, gc_zstd_sequence* seqs,
                           uint32_t* seq_counts);

/* Sequences of every 64 KiB sub-chunk packed back to back (sub-chunk s
 * starts at sum(seq_counts[0..s))): what ZSTD_compressSequences consumes.
 * total receives the number of sequences; with seqs == NULL only the counts
 * are returned; GC_ESIZE if seq_capacity < total. */
int gc_zstd_find_sequences_packed(gc_ctx* ctx, const uint8_t* chunk, size_t n, uint32_t level, gc_zstd_sequence* seqs,
                                  size_t seq_capacity, uint32_t* seq_counts, size_t* total);

/* ---- consumers and containers (CPU; libzstd / liblzma) ----------------- */

/* Frames produced by gc_compress_chunk_zstd_custom for an n-byte chunk fit in
 * this many bytes (n_sub * ZSTD_compressBound(65536)); 0 if libzstd is missing. */
size_t gc_zstd_custom_bound(size_t n);

/* compress_zstd_custom.rs:120-199: one frame per 64 KiB sub-chunk from the
 * packed sequences (ZSTD_compressSequences, default CCtx level, reset per
 * sub-chunk; ZSTD_compress level 1 when libzstd rejects the sequences).
 * Frames are written back to back; sizes[s] receives each frame's size.
 * threads = 1 is the reference's shape; more threads give the same bytes. */
int gc_zstd_encode_sequences(const uint8_t* src, size_t n, const gc_zstd_sequence* seqs, const uint32_t* seq_counts,
                             uint8_t* out, size_t cap, uint64_t* sizes, size_t* out_len, int threads);

/* compress_chunk_zstd_custom (compress_zstd_custom.rs:47-203): GPU match
 * finding + libzstd; (Vec<Vec<u8>>, Vec<usize>) becomes frames back to back
 * in out plus sizes[n_sub]. */
int gc_compress_chunk_zstd_custom(gc_ctx* ctx, const uint8_t* chunk, size_t n, uint32_t level, uint8_t* out,
                                  size_t cap, uint64_t* sizes, size_t* out_len);

/* One Zstd frame -> exactly `expected` bytes (replaces the nvCOMP decode of
 * decompress.rs:146-365 with libzstd ZSTD_decompress). */
int gc_zstd_frame_decode(const uint8_t* src, size_t n, uint8_t* out, size_t expected, size_t* out_len);

/* compress_lzma2.rs:89-133: liblzma raw LZMA2 encoder at preset (<= 9). */
int gc_lzma2_raw_encode(const uint8_t* src, size_t n, uint32_t preset, uint8_t* out, size_t cap, size_t* out_len);

/* compress_lzma2.rs:135-172: raw LZMA2 decoder, 64 MiB dictionary. */
int gc_lzma2_raw_decode(const uint8_t* src, size_t n, uint8_t* out, size_t expected, size_t* out_len);

/* encode_single_sub_block (compress_lzma2_custom.rs:570-597): liblzma
 * preset 9 if any match_counts[i] > 0, else raw LZMA2 chunks; both end with
 * the 0x00 end marker.  cap >= n + n/8 + 1024 always suffices. */
int gc_encode_single_sub_block(const uint8_t* sub, size_t n, const uint32_t* match_counts, uint8_t* out, size_t cap,
                               size_t* out_len);

/* LZMA2 coded from the match lists (compress_lzma2_custom.rs:344-465
 * encode_lzma2_block, :601-664 cpu_encode_from_matches), with the
 * matched-literal, special-distance and repeat-match coding of the LZMA
 * format: one stream (one chunk + the 0x00 end marker) per sub-block of at
 * most 65536 bytes, from its packed matches (gc_lzma2_find_matches_packed
 * layout: distance << 16 | length per position, 0 = none).  verify != 0
 * decodes the stream with liblzma and falls back to raw chunks if it does
 * not round-trip (:640-647); a stream that does not fit one chunk or is not
 * smaller than the raw chunks is replaced by them as well.
 * cap >= gc_lzma2_encode_bound(n) always suffices. */
size_t gc_lzma2_encode_bound(size_t n);
int gc_lzma2_encode_matches(const uint8_t* sub, size_t n, const uint32_t* packed, int verify, uint8_t* out,
                            size_t cap, size_t* out_len);
/* A whole chunk: one verified stream per 64 KiB sub-block, concatenated,
 * sizes[s] = bytes of stream s; threads <= 0: min(cores, 8).  cap >= the sum
 * of gc_lzma2_encode_bound over the sub-blocks suffices. */
int gc_cpu_encode_from_matches(const uint8_t* src, size_t n, const uint32_t* packed, int threads, uint8_t* out,
                               size_t cap, uint64_t* sizes, size_t* out_len);
/* NVLZ container (chunk size 65536) whose sub-blocks are coded from the GPU's
 * matches by gc_cpu_encode_from_matches instead of liblzma preset 9:
 * dictionary-size chunks (dict_mb, 0 = 128) match-found on the GPU.  No
 * counterpart in the reference's pipeline (its combined API,
 * compress_lzma2_custom.rs:667-673, is not called).  Decoded by gc_decompress. */
int gc_nvlz_compress_matches(gc_ctx* ctx, const uint8_t* src, size_t n, uint32_t dict_mb, int threads, uint8_t* out,
                             size_t cap, size_t* out_len);

/* The coder of gc_lzma2_encode_matches on the GPU (the capability of
 * compress_lzma2_custom.rs:601-664 cpu_encode_from_matches, moved to the
 * device): one stream per 64 KiB sub-block of d_input (total_input_size
 * bytes) from d_packed (gc_lzma2_match_find_packed_dev layout, sub-block size
 * 65536).  Stream s is written to d_output + s * GC_LZMA2_GPU_SLOT +
 * GC_LZMA2_GPU_HEAD, its length to d_output_sizes[s]; d_output holds
 * ceil(total_input_size / 65536) slots.  Each stream has the bytes
 * gc_lzma2_encode_matches writes for the sub-block with verify = 0 (no
 * decode-back check: the coder is pinned by its tests instead).  Device
 * pointers, asynchronous on the context's stream. */
#define GC_LZMA2_GPU_SLOT 65552u
#define GC_LZMA2_GPU_HEAD 2u
int gc_lzma2_encode_dev(gc_ctx* ctx, const uint8_t* d_input, const uint32_t* d_packed, uint8_t* d_output,
                        uint32_t* d_output_sizes, uint64_t total_input_size);
/* Host chunk -> match finding and LZMA2 coding on the GPU: one stream per
 * 64 KiB sub-block, packed in order into out, sizes[s]; cap >= n + 4 per
 * sub-block suffices.  The GPU counterpart of gc_cpu_encode_from_matches on
 * gc_lzma2_find_matches_packed's matches.  Coded in two halves, one launch
 * each on the context's two streams, so that copies to and from host memory
 * overlap the coding. */
int gc_lzma2_streams_gpu(gc_ctx* ctx, const uint8_t* chunk, size_t n, uint8_t* out, size_t cap, uint64_t* sizes,
                         size_t* out_len);
/* NVLZ container of GPU-coded streams, byte-equal to gc_nvlz_compress_matches'
 * for the same input (dict_mb is accepted for that interface and changes
 * nothing: no match crosses a sub-block); cap >= gc_container_bound(n,
 * GC_ALGO_LZMA2).  Decoded by gc_decompress.  Coded in pieces of up to 2 GiB,
 * each as gc_lzma2_streams_gpu codes a chunk. */
int gc_nvlz_compress_gpu(gc_ctx* ctx, const uint8_t* src, size_t n, uint32_t dict_mb, uint8_t* out, size_t cap,
                         size_t* out_len);

/* zstd_encode_block (zstd_compress.cu:756-1038, not called by the reference's
 * pipeline): one Zstd frame per sub-chunk from the match finder's sequences
 * (seq_format as gc_zstd_match_find_dev) and literals (d_literals at
 * s * chunk_size, d_lit_counts), written at d_output + s * max_frame_size,
 * its size in d_output_sizes[s].  The sequences are coded with the format's
 * predefined FSE distributions and the literals raw; a sub-chunk whose frame
 * would not be smaller than its bytes gets a raw block (d_input).  Device
 * pointers, asynchronous on the context's stream; max_frame_size >=
 * chunk_size + 16.  Frames are decodable by any Zstd decoder. */
int gc_zstd_encode_block_dev(gc_ctx* ctx, const uint8_t* d_input, const void* d_sequences, uint32_t seq_format,
                             const uint8_t* d_literals, const uint32_t* d_seq_counts, const uint32_t* d_lit_counts,
                             uint8_t* d_output, uint32_t* d_output_sizes, uint32_t chunk_size, uint32_t num_chunks,
                             uint64_t total_input_size, uint32_t max_sequences_per_chunk, uint32_t max_frame_size);
/* Host chunk -> match finding and frame coding on the GPU: one frame per
 * 64 KiB sub-block, packed in order into out, sizes[s]; cap >= n + 16 per
 * sub-block suffices.  The GPU counterpart of gc_compress_chunk_zstd_custom
 * (libzstd's entropy coding on the CPU): faster, larger.  Coded in pieces of
 * 128 MiB, one piece's copies to and from host memory overlapping the next
 * piece's kernels (the context's second stream). */
int gc_zstd_frames_gpu(gc_ctx* ctx, const uint8_t* chunk, size_t n, uint32_t level, uint8_t* out, size_t cap,
                       uint64_t* sizes, size_t* out_len);
/* NVZS container with GPU-coded frames (decoded by gc_decompress), coded as
 * gc_zstd_frames_gpu codes them. */
size_t gc_nvzs_gpu_bound(size_t n);
int gc_nvzs_compress_gpu(gc_ctx* ctx, const uint8_t* src, size_t n, uint32_t level, uint8_t* out, size_t cap,
                         size_t* out_len);

/* Largest NVZS (algo GC_ALGO_ZSTD) / NVLZ level >= 1 container for n bytes. */
size_t gc_container_bound(size_t n, uint32_t algo);

/* pipeline.rs:19-282 at level > 0: NVZS container of the whole input
 * (chunk size 65536, one frame per sub-chunk). */
int gc_nvzs_compress(gc_ctx* ctx, const uint8_t* src, size_t n, uint32_t level, int threads, uint8_t* out,
                     size_t cap, size_t* out_len);

/* pipeline_lzma2.rs:150-292 at level > 0: dictionary-size chunks (dict_mb,
 * 0 = 128) match-found on the GPU, sub-blocks encoded by a pool of threads
 * (<= 0: min(cores, 8)), NVLZ with chunk size 65536. */
int gc_nvlz_compress(gc_ctx* ctx, const uint8_t* src, size_t n, uint32_t level, uint32_t dict_mb, int threads,
                     uint8_t* out, size_t cap, size_t* out_len);

/* pipeline_lzma2.rs:293-333, level 0 (CPU only, the reference's baseline):
 * liblzma on whole dictionary-size chunks at `preset`. */
size_t gc_nvlz_l0_bound(size_t n, uint32_t dict_mb);
int gc_nvlz_compress_l0(const uint8_t* src, size_t n, uint32_t dict_mb, uint32_t preset, int threads, uint8_t* out,
                        size_t cap, size_t* out_len);

/* decompress.rs:9-143 as a pipeline: the container file at in_path decoded
 * into out_path, frames read by one thread, decoded by `threads` (<= 0:
 * min(cores, 8)) and written in order, at most 2 x threads frames in memory.
 * *out_bytes (optional) = bytes written.  GC_EIO for files that cannot be
 * opened or written, GC_EFORMAT for a malformed or truncated container. */
int gc_decompress_file(const char* in_path, const char* out_path, int threads, uint64_t* out_bytes);

/* decompress.rs:35-143: header of an NVZS / NVLZ container. */
int gc_container_info(const uint8_t* in, size_t n, uint32_t* algo, uint64_t* original_size, uint64_t* chunk_size,
                      uint64_t* num_chunks);

/* decompress.rs:9-379: whole container -> original bytes. */
int gc_decompress(const uint8_t* in, size_t n, uint8_t* out, size_t cap, size_t* out_len);

/* ---- batched form ----------------------------------------------------- */
/* No single reference counterpart: the reference calls the two host
 * functions once per chunk (pipeline.rs:145, pipeline_lzma2.rs:190).  A batch
 * runs every 64 KiB sub-block of many inputs in one launch per kernel (work
 * items assigned grid-stride to the workgroups).  Per input, results equal
 * gc_zstd_find_sequences / gc_lzma2_find_matches_packed on that input alone. */

#define GC_ALGO_ZSTD 0u
#define GC_ALGO_LZMA2 1u

typedef struct {
    uint64_t offset; /* byte offset of the input in the device arena (multiple of 4) */
    uint64_t size;   /* input bytes */
    uint32_t algo;   /* GC_ALGO_ZSTD or GC_ALGO_LZMA2 */
    uint32_t level;  /* Zstd: 1 -> depth 16, 2 -> depth 64, else 1; LZMA2: depth 32 */
} gc_batch_input;

/* Output-invariant launch knobs: workgroups per CU of each kernel (grid =
 * min(knob * CUs, items)), 1..64, and the Zstd verify pass's strip width. */
typedef struct {
    uint32_t cand_wg_per_cu;  /* candidate pass (LDS table), both finders */
    uint32_t parse_wg_per_cu; /* Zstd parse (4 sub-blocks per workgroup) */
    uint32_t elem_wg_per_cu;  /* per-position passes: LZMA2 lengths, Zstd verify */
    uint32_t info_strip;      /* positions per thread in the Zstd verify pass: 4 or 8 */
} gc_tune;

/* Batch execution modes.
 *   GC_BATCH_STATIC: one launch per kernel, grid = items (or the tuned
 *     workgroups per CU), items assigned grid-stride.
 *   GC_BATCH_PERSISTENT: resident workgroups of each kernel drain a device
 *     work queue; the whole pass (one stream, kernel nodes only) is captured
 *     in a hipGraph once and replayed by gc_batch_run.  Queue loops are
 *     bounded; gc_batch_check reports a loop that passed its bound or a
 *     kernel that did not take every item (and the copy functions check it).
 *   GC_BATCH_QUEUE_GRAPH: the first queue + hipGraph design (LZMA2 on a
 *     forked stream inside the capture, queues reset by a captured kernel);
 *     it hung on its first replay on MI355X with no cause found and is
 *     refused with GC_EINVAL. */
#define GC_BATCH_STATIC 0u
#define GC_BATCH_QUEUE_GRAPH 1u
#define GC_BATCH_PERSISTENT 2u

typedef struct gc_batch gc_batch;

/* Inputs must lie inside d_arena[0, arena_size); the arena stays owned by the
 * caller and must outlive the batch.  Output buffers are owned by the batch:
 * Zstd sequences (21845 slots of gc_zstd_sequence per sub-block) and
 * per-sub-block counts, LZMA2 packed matches (one u32 per position). */
gc_batch* gc_batch_create(gc_ctx* ctx, const uint8_t* d_arena, uint64_t arena_size, const gc_batch_input* inputs,
                          uint32_t n_inputs);
int gc_batch_destroy(gc_batch* b);
/* Enqueue one pass on the context stream (asynchronous). */
int gc_batch_run(gc_batch* b);
int gc_batch_set_mode(gc_batch* b, uint32_t mode);
/* Waits for the passes queued so far; GC_EHIP if a persistent pass's work
 * queue reported an error since the last check (then cleared), else 0. */
int gc_batch_check(gc_batch* b);
/* One pass with events between the kernels, synchronous: ms[0] candidate
 * kernel, ms[1] LZMA2 length kernel, ms[2] Zstd verify kernel, ms[3] Zstd
 * parse kernel. */
int gc_batch_run_timed(gc_batch* b, float* ms);
int gc_batch_set_tune(gc_batch* b, const gc_tune* t);
int gc_batch_get_tune(gc_batch* b, gc_tune* t);
/* Try the knob values on this batch (reps timed passes each) and keep the fastest. */
int gc_batch_autotune(gc_batch* b, uint32_t reps, gc_tune* best, float* best_ms);
/* Where input i's results live: n_sub sub-blocks; Zstd: sequences from slot
 * out_base, counts from count_base; LZMA2: packed from position out_base. */
int gc_batch_info(gc_batch* b, uint32_t i, uint32_t* n_sub, uint64_t* out_base, uint64_t* count_base);
int gc_batch_device_outputs(gc_batch* b, gc_zstd_sequence** d_seqs, uint32_t** d_counts, uint32_t** d_packed);
/* Copy input i's results to the host (synchronous).  Zstd: counts[n_sub] and
 * seqs[n_sub * 21845] (only the used slots are written; seqs may be NULL). */
int gc_batch_copy_zstd(gc_batch* b, uint32_t i, gc_zstd_sequence* seqs, uint32_t* counts);
int gc_batch_copy_lzma2(gc_batch* b, uint32_t i, uint32_t* packed);

/* ---- served form ------------------------------------------------------- */
/* No reference counterpart (the reference compresses one chunk at a time,
 * pipeline.rs:138-175).  A server thread admits requests in arrival order
 * while their device memory fits hbm_budget bytes (a host input is copied in
 * as it is admitted) and batches them continuously into the persistent
 * kernels: whenever one of its tick slots (each with its own stream, work
 * queue and staging memory) is free, every admitted request whose input is on
 * the device goes into a tick, so requests that arrive while others run start
 * beside them at the next free slot.  A request's results equal the plain
 * form's for its input alone.  Request memory (caller's `data` or
 * `d_input`) must stay valid until the request is done. */
typedef struct gc_server gc_server;

typedef struct {
    uint64_t size;
    uint32_t algo;
    uint32_t n_sub;
    uint64_t n_sequences; /* Zstd: sequences found */
    double t_submit, t_admit, t_start, t_done; /* seconds, steady clock */
} gc_request_info;

typedef struct {
    uint64_t batches, items, completed, bytes; /* batches: ticks started */
    uint64_t hbm_in_use, pending;
    double t_batches;    /* device seconds of the ticks (events around each; ticks overlap, so the sum can exceed wall time) */
    double t_fetch;      /* seconds of result copies on the copy-back stream (events around each tick's) */
    double t_host;       /* server-thread seconds spent building ticks and staging result copies */
    /* per kernel of a tick (setup, candidates, LZMA2 lengths, Zstd info, parse
     * order, Zstd parse): device seconds between the events around it, summed
     * over the ticks, and its launches */
    double t_kernel[6];
    uint64_t launches[6];
    /* finished ticks: Zstd input bytes, sub-blocks and sequences, LZMA2 input bytes */
    uint64_t zstd_bytes, zstd_sub_blocks, zstd_sequences, lzma2_bytes;
} gc_server_stats;

/* max_batch_items: sub-blocks per tick (0 = 65536); a request with more is
 * refused (GC_EBUDGET). */
gc_server* gc_server_start(gc_ctx* ctx, uint64_t hbm_budget, uint32_t max_batch_items);
/* Returns the request id (> 0), or a negative GC_E* code (GC_EBUDGET: the
 * request alone needs more than the budget). */
int64_t gc_server_submit(gc_server* s, const uint8_t* data, uint64_t n, uint32_t algo, uint32_t level);
/* As gc_server_submit, with the results copied by DMA straight into the
 * caller's memory (pinned memory from hipHostMalloc gives full PCIe rate):
 * LZMA2 `out` holds n u32 (out_bytes >= 4n, else GC_ESIZE at once); Zstd
 * `out` holds packed gc_zstd_sequence records and `seq_counts` one u32 per
 * sub-block.  A Zstd request whose sequences exceed out_bytes fails with
 * GC_ESIZE (gc_request_info.n_sequences says how many it found).  Both
 * buffers must stay valid until the request is done; gc_server_result_* do
 * not apply to such requests. */
int64_t gc_server_submit_into(gc_server* s, const uint8_t* data, uint64_t n, uint32_t algo, uint32_t level, void* out,
                              uint64_t out_bytes, uint32_t* seq_counts);
/* A request whose input is already in device memory (4-byte aligned, n
 * bytes readable): it is read in place and its results stay on the device
 * (gc_server_result_dev) until gc_server_release. */
int64_t gc_server_submit_dev(gc_server* s, const uint8_t* d_input, uint64_t n, uint32_t algo, uint32_t level);
/* count gc_server_submit_dev requests at once (request i: d_inputs[i],
 * sizes[i], algos[i], levels[i]), queued in order under one lock; ids[i]
 * receives request i's id.  All are checked first: on an error none is
 * submitted.  No reference counterpart (a serving client's batch of
 * arrivals). */
int gc_server_submit_dev_many(gc_server* s, const uint8_t* const* d_inputs, const uint64_t* sizes,
                              const uint32_t* algos, const uint32_t* levels, uint32_t count, int64_t* ids);
/* 0 done, 1 still running after timeout_ms (< 0 waits), < 0 failed. */
int gc_server_wait(gc_server* s, uint64_t id, int timeout_ms, gc_request_info* info);
/* Results of a finished request: Zstd packed sequences (cap records) and per
 * sub-block counts; LZMA2 packed matches, one u32 per input byte. */
int gc_server_result_zstd(gc_server* s, uint64_t id, gc_zstd_sequence* seqs, uint64_t cap, uint32_t* seq_counts);
int gc_server_result_lzma2(gc_server* s, uint64_t id, uint32_t* packed);
/* Results of a finished gc_server_submit_dev request, valid until it is
 * released: *d_results is a device pointer, Zstd: n_sub slots of
 * GC_SERVE_ZSTD_SLOT gc_zstd_sequence records (sub-block k's at k *
 * GC_SERVE_ZSTD_SLOT, the first seq_counts[k] used, at most 21845), LZMA2:
 * one packed u32 per input byte; *seq_counts (Zstd, host memory, n_sub
 * entries; NULL for LZMA2).  (A slot is 6 bytes per position of its
 * sub-block: while the request runs it holds the sub-block's candidates and
 * info words, so the request takes 6 bytes of the HBM budget per input byte.) */
#define GC_SERVE_ZSTD_SLOT 24576u
int gc_server_result_dev(gc_server* s, uint64_t id, void** d_results, const uint32_t** seq_counts);
int gc_server_release(gc_server* s, uint64_t id);
/* For each of ids in order: wait for it (no time limit), fill infos[i],
 * status[i] = what gc_server_wait returned, then release it (its device
 * results too).  Returns the first error, or GC_OK.  For a client that
 * reaps many completions at once. */
int gc_server_wait_release_many(gc_server* s, const uint64_t* ids, uint32_t count, gc_request_info* infos,
                                int* status);
int gc_server_get_stats(gc_server* s, gc_server_stats* out);
/* Finishes every submitted request, then stops. */
int gc_server_stop(gc_server* s);

/* ---- synthetic corpora -------------------------------------------------- */
/* Deterministic four-kind mix (uniform random, word text, back-references,
 * structured binary) in segments of `segment` bytes; weights[4] are relative
 * shares.  exact_shares: each group of 20 segments holds round(20*w) of each
 * kind, shuffled; otherwise kinds are drawn independently.  gap_every /
 * gap_len: a zero gap of gap_len bytes after every gap_every bytes of content
 * (0 = none).  No reference counterpart (workload generator). */
int gc_corpus_generate(uint8_t* out, uint64_t n, uint64_t seed, const double* weights, uint64_t segment,
                       uint32_t exact_shares, uint64_t gap_every, uint64_t gap_len);

#ifdef __cplusplus
}
#endif
#endif /* GCOMPRESS_H */
