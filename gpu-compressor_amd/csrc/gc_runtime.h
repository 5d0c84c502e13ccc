// gc_runtime.h -- internal host-side declarations of libgcompress.
#pragma once
#include <hip/hip_runtime_api.h>
#include <stddef.h>
#include <stdint.h>

#include <mutex>
#include <string>
#include <vector>

#include "gc_device.h"

namespace gc {

// Thread-local error message behind gc_last_error() (a fixed buffer; the
// const char* forms allocate nothing).
void set_error(const char* msg);
void set_error(const std::string& msg);
int fail(int code, const char* msg);
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

// Grow-only device buffer owned by a context.
struct DevBuf {
    void* p = nullptr;
    size_t size = 0;
    int ensure(size_t bytes);
    void release();
};

// Loaded kernels.
struct Kernels {
    hipModule_t module = nullptr;        // every kernel but the parse
    hipModule_t module_parse = nullptr;  // the Zstd parse kernels
    hipFunction_t cand = nullptr;
    hipFunction_t cand_wide = nullptr;  // 1024 threads (launches of few sub-blocks)
    hipFunction_t lzma2_len = nullptr;      // packed words
    hipFunction_t lzma2_len_ref = nullptr;  // reference layout
    hipFunction_t lzma2_triage = nullptr;   // triage byte only
    hipFunction_t zstd_info = nullptr;
    hipFunction_t zstd_info4 = nullptr;
    hipFunction_t zstd_parse[2][2] = {};  // [reference layout][literals]
    hipFunction_t zstd_parse_seg[2] = {};  // segmented parse [lazy step]
    hipFunction_t fill_u32 = nullptr;
    hipFunction_t gather_seqs = nullptr;
    hipFunction_t copy_u32 = nullptr;
    hipFunction_t order_items = nullptr;
    hipFunction_t zstd_encode = nullptr;
    hipFunction_t gather_bytes = nullptr;
    hipFunction_t lzma2_encode = nullptr;
    hipFunction_t lzma2_cost = nullptr;
    hipFunction_t make_items = nullptr;
    // persistent forms: items from a device work queue (gc_queue)
    hipFunction_t cand_q = nullptr;
    hipFunction_t lzma2_len_q = nullptr;
    hipFunction_t zstd_info_q = nullptr;
    hipFunction_t zstd_info4_q = nullptr;
    hipFunction_t zstd_parse_q = nullptr;
    hipFunction_t tick_setup = nullptr;
    // the served form's launches of the same persistent kernels (names of their own)
    hipFunction_t cand_sq = nullptr;
    hipFunction_t lzma2_len_sq = nullptr;
    hipFunction_t zstd_info_sq = nullptr;
    hipFunction_t zstd_parse_sq = nullptr;
};

}  // namespace gc

struct gc_ctx {
    int device = 0;
    int cu_count = 0;
    hipStream_t stream = nullptr;   // main stream (default priority)
    hipStream_t stream2 = nullptr;  // second stream for independent work (default priority)
    gc::Kernels k;
    // Scratch reused across calls.
    gc::DevBuf items;      // work items
    gc::DevBuf cand;       // per-position candidates (u16)
    gc::DevBuf info;       // zstd per-position candidate words
    gc::DevBuf input;      // staging for host-pointer entry points
    gc::DevBuf out_a, out_b, out_c, out_d, out_e, out_f;
    gc::DevBuf input_b, out_c_b;  // the second slot of input and out_c (the NVZS / NVLZ pipelines)
    gc::DevBuf items_b, cand_b, out_a_b, out_b_b, out_e_b, out_f_b;  // the second set (NVLZ pipeline)
    gc::DevBuf seg;        // record buffers of the segmented parse
    // launches of at most this many sub-blocks parse with the segmented
    // kernel (0: never); larger ones fill the GPU with one walk per wavefront
    uint32_t seg_max_sub = 0;
    std::mutex mu;
};

namespace gc {

// Launch helpers.  grid = workgroups; items are taken grid-stride.
// (the wide form for launches of at most 2 sub-blocks per CU: each
// workgroup's latency sets their time)
int launch_cand(gc_ctx* ctx, hipStream_t stream, const uint8_t* arena, const gc_item* items, uint32_t n_items,
                uint32_t grid, uint16_t* cand16);
int launch_lzma2_len(gc_ctx* ctx, hipStream_t stream, const uint8_t* arena, const gc_item* items, uint32_t n_items,
                     uint32_t grid, const uint16_t* cand16, uint32_t* packed, gc_lzma_match* ref_m, uint32_t* ref_c,
                     uint32_t max_matches,
                     uint8_t* triage = nullptr);
int launch_zstd_info(gc_ctx* ctx, hipStream_t stream, const uint8_t* arena, const gc_item* items, uint32_t n_items,
                     uint32_t grid, const uint16_t* cand16, uint32_t* info, uint32_t strip = 8,
                     uint32_t* cost = nullptr);
int launch_zstd_parse(gc_ctx* ctx, hipStream_t stream, const uint8_t* arena, const gc_item* items, uint32_t n_items,
                      uint32_t grid, const uint32_t* info, void* seq, uint32_t seq_fmt, uint8_t* lit,
                      uint32_t* seq_counts, uint32_t* lit_counts, uint32_t max_seq,
                      const uint32_t* order = nullptr);
// Segmented parse (one workgroup of GC_SEG_WAVES wavefronts per item); the
// context's record buffers are sized for the grid here.
int launch_zstd_parse_seg(gc_ctx* ctx, hipStream_t stream, const uint8_t* arena, const gc_item* items,
                          uint32_t n_items, const uint32_t* info, void* seq, uint32_t seq_fmt, uint8_t* lit,
                          uint32_t* seq_counts, uint32_t* lit_counts, uint32_t max_seq, uint32_t depth);
int launch_fill(gc_ctx* ctx, hipStream_t stream, uint32_t* p, uint32_t v, uint64_t n);
int launch_gather(gc_ctx* ctx, hipStream_t stream, const void* d_in, const uint32_t* d_counts, const uint64_t* d_offsets,
                  void* d_out, uint32_t nsub, uint32_t max_seq);
int launch_order(gc_ctx* ctx, hipStream_t stream, const uint32_t* cost, uint32_t n, uint32_t* order,
                 const gc_queue* q = nullptr);
// Persistent kernels: grid workgroups take items from the queue q (item
// counts in q->n); nothing else differs from the static launches.
// served: the served form's kernel names.
int launch_cand_q(gc_ctx* ctx, hipStream_t stream, const uint8_t* arena, const gc_item* items, gc_queue* q,
                  uint32_t grid, uint16_t* cand16, bool served = false);
int launch_lzma2_len_q(gc_ctx* ctx, hipStream_t stream, const uint8_t* arena, const gc_item* items, gc_queue* q,
                       uint32_t grid, const uint16_t* cand16, uint32_t* packed, bool served = false);
int launch_zstd_info_q(gc_ctx* ctx, hipStream_t stream, const uint8_t* arena, const gc_item* items, gc_queue* q,
                       uint32_t grid, const uint16_t* cand16, uint32_t* info, uint32_t strip, uint32_t* cost,
                       bool served = false);
// Served form: copy a tick's counts and items from pinned memory into the
// slot's queue and item array.
int launch_tick_setup(gc_ctx* ctx, hipStream_t stream, const gc_queue* hdr, const gc_item* src, gc_queue* q,
                      gc_item* items, uint32_t cap);
int launch_zstd_parse_q(gc_ctx* ctx, hipStream_t stream, const uint8_t* arena, const gc_item* items, gc_queue* q,
                        uint32_t grid, const uint32_t* info, void* seq, uint32_t* seq_counts, uint32_t max_seq,
                        const uint32_t* order, bool served = false);
// Zstd frames from sequences and literals (one wavefront per item), and the
// packing of frames written in fixed slots.
int launch_zstd_encode(gc_ctx* ctx, hipStream_t stream, const uint8_t* arena, const gc_item* items, uint32_t n_items,
                       const void* seqs, uint32_t seq_fmt, const uint8_t* lits, const uint32_t* seq_counts,
                       const uint32_t* lit_counts, uint8_t* out, uint32_t* sizes, uint32_t max_frame,
                       uint32_t max_seq, uint32_t* order);
// LZMA2 chunks coded from packed matches: one 64-thread workgroup per item,
// chunk of item count_pos s in out[s * GC_LZMA2_GPU_SLOT + GC_LZMA2_GPU_HEAD ...],
// its length in sizes[s]; the items taken costliest first: a cost per item
// (gc_lzma2_cost_kernel), then their order (gc_order_items), both in `scratch`
// (8 bytes per item), then the coder.
int launch_lzma2_encode_ordered(gc_ctx* ctx, hipStream_t stream, const uint8_t* arena, const gc_item* items,
                                uint32_t n_items, const uint32_t* packed, uint8_t* out, uint32_t* sizes,
                                uint32_t* scratch);
int launch_gather_bytes(gc_ctx* ctx, hipStream_t stream, const uint8_t* in, const uint32_t* sizes,
                        const uint64_t* offsets, uint8_t* out, uint32_t slot, uint32_t nsub);
int launch_copy_u32(gc_ctx* ctx, hipStream_t stream, const uint32_t* src, uint32_t* dst, uint64_t n);
// Items of a buffer tiled in sub-blocks from its start, written on the device.
int launch_make_items(gc_ctx* ctx, hipStream_t stream, gc_item* items, uint64_t total, uint32_t sub, uint32_t nsub,
                      uint32_t max_seq, uint32_t depth);
// Per-device scratch of the allocation-free entry point (gc_noalloc.hip):
// looked up once per device when a context opens.
int noalloc_attach(int device);
int upload(gc_ctx* ctx, DevBuf& buf, const void* src, size_t bytes);
int upload_on(gc_ctx* ctx, hipStream_t stream, DevBuf& buf, const void* src, size_t bytes);

uint32_t zstd_depth_for_level(uint32_t level);

}  // namespace gc
