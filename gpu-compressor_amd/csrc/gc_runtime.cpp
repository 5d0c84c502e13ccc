// gc_runtime.cpp -- context, embedded code object, device memory, launches.
#include "gc_runtime.h"

#include <stdio.h>
#include <string.h>

#include "../../include/gcompress.h"

// The gfx950 code objects built from kernels.hip (the parse kernels, and
// the rest) are linked into this library and loaded with hipModuleLoadData
// (no files, no JIT at run time).
#if !defined(GC_CO_PATH) || !defined(GC_PARSE_CO_PATH)
#error "GC_CO_PATH and GC_PARSE_CO_PATH must name the code objects to embed"
#endif
__asm__(".section .rodata\n"
        ".balign 4096\n"
        ".globl gc_code_object\n"
        "gc_code_object:\n"
        ".incbin \"" GC_CO_PATH "\"\n"
        ".byte 0\n"
        ".balign 4096\n"
        ".globl gc_parse_code_object\n"
        "gc_parse_code_object:\n"
        ".incbin \"" GC_PARSE_CO_PATH "\"\n"
        ".byte 0\n"
        ".previous\n");
extern "C" const unsigned char gc_code_object[];
extern "C" const unsigned char gc_parse_code_object[];

namespace gc {

// A fixed buffer: reporting an error allocates nothing (the allocation-free
// entry points stay so on their error paths too).
static thread_local char t_error[512];

void set_error(const char* msg) {
    size_t k = 0;
    for (; msg && msg[k] && k + 1 < sizeof(t_error); k++) t_error[k] = msg[k];
    t_error[k] = 0;
}

void set_error(const std::string& msg) { set_error(msg.c_str()); }

int fail(int code, const char* msg) {
    set_error(msg);
    return code;
}

int fail(int code, const std::string& msg) { return fail(code, msg.c_str()); }

int hip_fail(hipError_t e, const char* what) {
    char buf[256];
    snprintf(buf, sizeof(buf), "%s: %s", what, hipGetErrorString(e));
    set_error(buf);
    return GC_EHIP;
}

int DevBuf::ensure(size_t bytes) {
    if (bytes <= size && p) return GC_OK;
    release();
    size_t want = bytes < 256 ? 256 : bytes;
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
        p = nullptr;
        return hip_fail(e, "hipMalloc");
    }
    size = want;
    return GC_OK;
}

void DevBuf::release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    size = 0;
}

int upload_on(gc_ctx* ctx, hipStream_t stream, DevBuf& buf, const void* src, size_t bytes) {
    (void)ctx;
    int rc = buf.ensure(bytes);
    if (rc) return rc;
    hipError_t e = hipMemcpyAsync(buf.p, src, bytes, hipMemcpyHostToDevice, stream);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync(items)");
    e = hipStreamSynchronize(stream);  // src is pageable host memory owned by the caller
    if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    return GC_OK;
}

int upload(gc_ctx* ctx, DevBuf& buf, const void* src, size_t bytes) {
    return upload_on(ctx, ctx->stream, buf, src, bytes);
}

static int launch(hipStream_t stream, hipFunction_t f, uint32_t grid, uint32_t block, void** args, const char* name) {
    if (grid == 0) return GC_OK;
    hipError_t e = hipModuleLaunchKernel(f, grid, 1, 1, block, 1, 1, 0, stream, args, nullptr);
    if (e != hipSuccess) return hip_fail(e, name);
    return GC_OK;
}

int launch_cand(gc_ctx* ctx, hipStream_t stream, const uint8_t* arena, const gc_item* items, uint32_t n_items,
                uint32_t grid, uint16_t* cand16) {
    void* args[] = {&arena, &items, &n_items, &cand16};
    if (n_items <= 2u * (uint32_t)ctx->cu_count)
        return launch(stream, ctx->k.cand_wide, grid, GC_CAND_WIDE_THREADS, args, "launch gc_cand_wide_kernel");
    return launch(stream, ctx->k.cand, grid, GC_CAND_THREADS, args, "launch gc_cand_kernel");
}

int launch_lzma2_len(gc_ctx* ctx, hipStream_t stream, const uint8_t* arena, const gc_item* items, uint32_t n_items,
                     uint32_t grid, const uint16_t* cand16, uint32_t* packed, gc_lzma_match* ref_m, uint32_t* ref_c,
                     uint32_t max_matches, uint8_t* triage) {
    void* args[] = {&arena, &items, &n_items, &cand16, &packed, &ref_m, &ref_c, &max_matches, &triage};
    hipFunction_t f = ref_c ? ctx->k.lzma2_len_ref : (packed ? ctx->k.lzma2_len : ctx->k.lzma2_triage);
    return launch(stream, f, grid, 1024, args, "launch gc_lzma2_len_kernel");
}

int launch_zstd_info(gc_ctx* ctx, hipStream_t stream, const uint8_t* arena, const gc_item* items, uint32_t n_items,
                     uint32_t grid, const uint16_t* cand16, uint32_t* info, uint32_t strip, uint32_t* cost) {
    void* args[] = {&arena, &items, &n_items, &cand16, &info, &cost};
    // one 1024-thread workgroup per sub-block (the sub-block is staged in LDS)
    if (strip == 4) return launch(stream, ctx->k.zstd_info4, grid, 1024, args, "launch gc_zstd_info4_kernel");
    return launch(stream, ctx->k.zstd_info, grid, 1024, args, "launch gc_zstd_info_kernel");
}

int launch_zstd_parse(gc_ctx* ctx, hipStream_t stream, const uint8_t* arena, const gc_item* items, uint32_t n_items,
                      uint32_t grid, const uint32_t* info, void* seq, uint32_t seq_fmt, uint8_t* lit,
                      uint32_t* seq_counts, uint32_t* lit_counts, uint32_t max_seq, const uint32_t* order) {
    void* args[] = {&arena, &items, &n_items, &info, &seq, &seq_fmt, &lit, &seq_counts, &lit_counts, &max_seq, &order};
    const int ref = seq_fmt == GC_SEQ_REF ? 1 : 0, lits = lit ? 1 : 0;
    return launch(stream, ctx->k.zstd_parse[ref][lits], grid, 256, args, "launch gc_zstd_parse_kernel");
}

int launch_zstd_parse_seg(gc_ctx* ctx, hipStream_t stream, const uint8_t* arena, const gc_item* items,
                          uint32_t n_items, const uint32_t* info, void* seq, uint32_t seq_fmt, uint8_t* lit,
                          uint32_t* seq_counts, uint32_t* lit_counts, uint32_t max_seq, uint32_t depth) {
    // two workgroups per CU are resident at most (66 KiB of LDS each); the
    // grid takes items grid-stride, so the record buffers stay bounded
    const uint32_t cap = 2u * (uint32_t)ctx->cu_count;
    uint32_t grid = n_items < cap ? n_items : cap;
    if (grid == 0) return GC_OK;
    int rc = ctx->seg.ensure((size_t)grid * GC_SEG_SCRATCH);
    if (rc) return rc;
    void* scratch = ctx->seg.p;
    void* args[] = {&arena, &items, &n_items, &info, &seq, &seq_fmt, &lit, &seq_counts, &lit_counts, &max_seq,
                    &scratch};
    return launch(stream, ctx->k.zstd_parse_seg[depth > 1 ? 1 : 0], grid, GC_SEG_WAVES * 64, args,
                  "launch gc_zstd_parse_seg_kernel");
}

int launch_zstd_encode(gc_ctx* ctx, hipStream_t stream, const uint8_t* arena, const gc_item* items, uint32_t n_items,
                       const void* seqs, uint32_t seq_fmt, const uint8_t* lits, const uint32_t* seq_counts,
                       const uint32_t* lit_counts, uint8_t* out, uint32_t* sizes, uint32_t max_frame,
                       uint32_t max_seq, uint32_t* order) {
    // the items with the most sequences first (item i's count is seq_counts[i])
    if (order && n_items) {
        if (int rc = launch_order(ctx, stream, seq_counts, n_items, order, nullptr)) return rc;
    }
    const uint32_t* ord = order;
    void* args[] = {&arena, &items,      &n_items, &seqs,  &seq_fmt,   &lits,     &seq_counts,
                    &lit_counts, &out, &sizes, &max_frame, &max_seq, &ord};
    // four wavefronts (items) per workgroup; each workgroup first tabulates the FSE tables in LDS
    uint32_t grid = (n_items + 3) / 4;
    const uint32_t cap = (uint32_t)ctx->cu_count * 8;
    if (grid > cap) grid = cap;
    if (grid == 0) grid = 1;
    return launch(stream, ctx->k.zstd_encode, grid, 256, args, "launch gc_zstd_encode_kernel");
}

int launch_lzma2_encode_ordered(gc_ctx* ctx, hipStream_t stream, const uint8_t* arena, const gc_item* items,
                                uint32_t n_items, const uint32_t* packed, uint8_t* out, uint32_t* sizes,
                                uint32_t* scratch) {
    if (n_items == 0) return GC_OK;
    uint32_t* cost = scratch;
    uint32_t* order = scratch + n_items;
    void* cargs[] = {&items, &n_items, &packed, &cost};
    int rc = launch(stream, ctx->k.lzma2_cost, n_items < 4096 ? n_items : 4096, 256, cargs, "launch gc_lzma2_cost_kernel");
    if (!rc) rc = launch_order(ctx, stream, cost, n_items, order, nullptr);
    if (rc) return rc;
    const uint32_t* ord = order;
    void* args[] = {&arena, &items, &n_items, &packed, &out, &sizes, &ord};
    return launch(stream, ctx->k.lzma2_encode, n_items, 64, args, "launch gc_lzma2_encode_kernel");
}

int launch_gather_bytes(gc_ctx* ctx, hipStream_t stream, const uint8_t* in, const uint32_t* sizes,
                        const uint64_t* offsets, uint8_t* out, uint32_t slot, uint32_t nsub) {
    void* args[] = {&in, &sizes, &offsets, &out, &slot, &nsub};
    uint32_t grid = nsub < 4096 ? (nsub ? nsub : 1) : 4096;
    return launch(stream, ctx->k.gather_bytes, grid, 256, args, "launch gc_gather_bytes");
}

int launch_fill(gc_ctx* ctx, hipStream_t stream, uint32_t* p, uint32_t v, uint64_t n) {
    uint64_t blocks = (n + 255) / 256;
    uint32_t grid = (uint32_t)(blocks < 4096 ? (blocks ? blocks : 1) : 4096);
    void* args[] = {&p, &v, &n};
    return launch(stream, ctx->k.fill_u32, grid, 256, args, "launch gc_fill_u32");
}

int launch_gather(gc_ctx* ctx, hipStream_t stream, const void* d_in, const uint32_t* d_counts, const uint64_t* d_offsets,
                  void* d_out, uint32_t nsub, uint32_t max_seq) {
    // at most 64 workgroups: when d_out is host memory the copy is PCIe-bound, and the CUs stay
    // free for the match-finding kernels running beside it
    void* args[] = {&d_in, &d_counts, &d_offsets, &d_out, &max_seq, &nsub};
    return launch(stream, ctx->k.gather_seqs, nsub < 64 ? nsub : 64, 256, args, "launch gc_gather_seqs");
}

int launch_order(gc_ctx* ctx, hipStream_t stream, const uint32_t* cost, uint32_t n, uint32_t* order,
                 const gc_queue* q) {
    void* args[] = {&cost, &n, &order, &q};
    return launch(stream, ctx->k.order_items, 1, 256, args, "launch gc_order_items");
}

int launch_cand_q(gc_ctx* ctx, hipStream_t stream, const uint8_t* arena, const gc_item* items, gc_queue* q,
                  uint32_t grid, uint16_t* cand16, bool served) {
    void* args[] = {&arena, &items, &q, &cand16};
    return launch(stream, served ? ctx->k.cand_sq : ctx->k.cand_q, grid, GC_CAND_THREADS, args,
                  "launch gc_cand_q_kernel");
}

int launch_lzma2_len_q(gc_ctx* ctx, hipStream_t stream, const uint8_t* arena, const gc_item* items, gc_queue* q,
                       uint32_t grid, const uint16_t* cand16, uint32_t* packed, bool served) {
    void* args[] = {&arena, &items, &q, &cand16, &packed};
    return launch(stream, served ? ctx->k.lzma2_len_sq : ctx->k.lzma2_len_q, grid, 1024, args,
                  "launch gc_lzma2_len_q_kernel");
}

int launch_zstd_info_q(gc_ctx* ctx, hipStream_t stream, const uint8_t* arena, const gc_item* items, gc_queue* q,
                       uint32_t grid, const uint16_t* cand16, uint32_t* info, uint32_t strip, uint32_t* cost,
                       bool served) {
    void* args[] = {&arena, &items, &q, &cand16, &info, &cost};
    hipFunction_t f = served ? ctx->k.zstd_info_sq : (strip == 4 ? ctx->k.zstd_info4_q : ctx->k.zstd_info_q);
    return launch(stream, f, grid, 1024, args, "launch gc_zstd_info_q_kernel");
}

int launch_tick_setup(gc_ctx* ctx, hipStream_t stream, const gc_queue* hdr, const gc_item* src, gc_queue* q,
                      gc_item* items, uint32_t cap) {
    void* args[] = {&hdr, &src, &q, &items, &cap};
    return launch(stream, ctx->k.tick_setup, 64, 256, args, "launch gc_tick_setup");
}

int launch_zstd_parse_q(gc_ctx* ctx, hipStream_t stream, const uint8_t* arena, const gc_item* items, gc_queue* q,
                        uint32_t grid, const uint32_t* info, void* seq, uint32_t* seq_counts, uint32_t max_seq,
                        const uint32_t* order, bool served) {
    void* args[] = {&arena, &items, &q, &info, &seq, &seq_counts, &max_seq, &order};
    return launch(stream, served ? ctx->k.zstd_parse_sq : ctx->k.zstd_parse_q, grid, 256, args,
                  "launch gc_zstd_parse_q_kernel");
}

int launch_copy_u32(gc_ctx* ctx, hipStream_t stream, const uint32_t* src, uint32_t* dst, uint64_t n) {
    uint32_t vec4 = (((uintptr_t)src | (uintptr_t)dst) & 15) == 0 ? 1u : 0u;
    // 32 workgroups keep PCIe busy when dst is host memory (the writes are posted), and leave the
    // CUs to the match-finding kernels running beside them
    uint64_t blocks = (n / 4 + 255) / 256;
    uint32_t grid = (uint32_t)(blocks < 32 ? (blocks ? blocks : 1) : 32);
    void* args[] = {&src, &dst, &n, &vec4};
    return launch(stream, ctx->k.copy_u32, grid, 256, args, "launch gc_copy_u32");
}

int launch_make_items(gc_ctx* ctx, hipStream_t stream, gc_item* items, uint64_t total, uint32_t sub, uint32_t nsub,
                      uint32_t max_seq, uint32_t depth) {
    void* args[] = {&items, &total, &sub, &nsub, &max_seq, &depth};
    return launch(stream, ctx->k.make_items, (nsub + 255) / 256, 256, args, "launch gc_make_items");
}

uint32_t zstd_depth_for_level(uint32_t level) {
    // compress_zstd_custom.rs:57-61
    if (level == 1) return GC_ZSTD_DEPTH_LAZY;
    if (level == 2) return 64u;
    return 1u;
}

}  // namespace gc

using namespace gc;

extern "C" {

const char* gc_last_error(void) { return t_error; }

int gc_device_count(void) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) return hip_fail(e, "hipGetDeviceCount");
    return n;
}

gc_ctx* gc_open(int device) {
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) {
        hip_fail(e, "hipSetDevice");
        return nullptr;
    }
    gc_ctx* ctx = new gc_ctx();
    ctx->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) ctx->cu_count = prop.multiProcessorCount;
    if (ctx->cu_count <= 0) ctx->cu_count = 256;
    e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipModuleLoadData(&ctx->k.module, gc_code_object);
    if (e == hipSuccess) e = hipModuleGetFunction(&ctx->k.cand, ctx->k.module, "gc_cand_kernel");
    if (e == hipSuccess) e = hipModuleGetFunction(&ctx->k.cand_wide, ctx->k.module, "gc_cand_wide_kernel");
    if (e == hipSuccess) e = hipModuleGetFunction(&ctx->k.lzma2_len, ctx->k.module, "gc_lzma2_len_kernel");
    if (e == hipSuccess) e = hipModuleGetFunction(&ctx->k.lzma2_len_ref, ctx->k.module, "gc_lzma2_len_ref_kernel");
    if (e == hipSuccess) e = hipModuleGetFunction(&ctx->k.lzma2_triage, ctx->k.module, "gc_lzma2_triage_kernel");
    if (e == hipSuccess) e = hipModuleGetFunction(&ctx->k.zstd_info, ctx->k.module, "gc_zstd_info_kernel");
    if (e == hipSuccess) e = hipModuleGetFunction(&ctx->k.zstd_info4, ctx->k.module, "gc_zstd_info4_kernel");
    if (e == hipSuccess) e = hipModuleLoadData(&ctx->k.module_parse, gc_parse_code_object);
    const char* parse_names[2][2] = {{"gc_zstd_parse_kernel", "gc_zstd_parse_lits_kernel"},
                                     {"gc_zstd_parse_ref_kernel", "gc_zstd_parse_ref_lits_kernel"}};
    for (int r = 0; r < 2; r++)
        for (int l = 0; l < 2; l++)
            if (e == hipSuccess) e = hipModuleGetFunction(&ctx->k.zstd_parse[r][l], ctx->k.module_parse, parse_names[r][l]);
    if (e == hipSuccess)
        e = hipModuleGetFunction(&ctx->k.zstd_parse_seg[0], ctx->k.module_parse, "gc_zstd_parse_seg1_kernel");
    if (e == hipSuccess)
        e = hipModuleGetFunction(&ctx->k.zstd_parse_seg[1], ctx->k.module_parse, "gc_zstd_parse_seg_kernel");
    if (e == hipSuccess) e = hipModuleGetFunction(&ctx->k.fill_u32, ctx->k.module, "gc_fill_u32");
    if (e == hipSuccess) e = hipModuleGetFunction(&ctx->k.gather_seqs, ctx->k.module, "gc_gather_seqs");
    if (e == hipSuccess) e = hipModuleGetFunction(&ctx->k.copy_u32, ctx->k.module, "gc_copy_u32");
    if (e == hipSuccess) e = hipModuleGetFunction(&ctx->k.order_items, ctx->k.module, "gc_order_items");
    if (e == hipSuccess) e = hipModuleGetFunction(&ctx->k.zstd_encode, ctx->k.module, "gc_zstd_encode_kernel");
    if (e == hipSuccess) e = hipModuleGetFunction(&ctx->k.gather_bytes, ctx->k.module, "gc_gather_bytes");
    if (e == hipSuccess) e = hipModuleGetFunction(&ctx->k.lzma2_encode, ctx->k.module, "gc_lzma2_encode_kernel");
    if (e == hipSuccess) e = hipModuleGetFunction(&ctx->k.lzma2_cost, ctx->k.module, "gc_lzma2_cost_kernel");
    if (e == hipSuccess) e = hipModuleGetFunction(&ctx->k.make_items, ctx->k.module, "gc_make_items");
    if (e == hipSuccess) e = hipModuleGetFunction(&ctx->k.cand_q, ctx->k.module, "gc_cand_q_kernel");
    if (e == hipSuccess) e = hipModuleGetFunction(&ctx->k.lzma2_len_q, ctx->k.module, "gc_lzma2_len_q_kernel");
    if (e == hipSuccess) e = hipModuleGetFunction(&ctx->k.zstd_info_q, ctx->k.module, "gc_zstd_info_q_kernel");
    if (e == hipSuccess) e = hipModuleGetFunction(&ctx->k.zstd_info4_q, ctx->k.module, "gc_zstd_info4_q_kernel");
    if (e == hipSuccess)
        e = hipModuleGetFunction(&ctx->k.zstd_parse_q, ctx->k.module_parse, "gc_zstd_parse_q_kernel");
    if (e == hipSuccess) e = hipModuleGetFunction(&ctx->k.tick_setup, ctx->k.module, "gc_tick_setup");
    if (e == hipSuccess) e = hipModuleGetFunction(&ctx->k.cand_sq, ctx->k.module, "gc_cand_sq_kernel");
    if (e == hipSuccess) e = hipModuleGetFunction(&ctx->k.lzma2_len_sq, ctx->k.module, "gc_lzma2_len_sq_kernel");
    if (e == hipSuccess) e = hipModuleGetFunction(&ctx->k.zstd_info_sq, ctx->k.module, "gc_zstd_info_sq_kernel");
    if (e == hipSuccess)
        e = hipModuleGetFunction(&ctx->k.zstd_parse_sq, ctx->k.module_parse, "gc_zstd_parse_sq_kernel");
    if (e != hipSuccess) {
        hip_fail(e, "gc_open (stream / hipModuleLoadData / hipModuleGetFunction)");
        gc_close(ctx);
        return nullptr;
    }
    ctx->seg_max_sub = GC_SEG_DEFAULT_MAX_SUB;
    // the allocation-free entry point's scratch on this device (a failure is
    // reported by that entry point; every other entry point works without it)
    (void)noalloc_attach(device);
    return ctx;
}

int gc_set_segmented_parse(gc_ctx* ctx, uint32_t max_sub_blocks) {
    if (!ctx) return fail(GC_EINVAL, "null context");
    ctx->seg_max_sub = max_sub_blocks;
    return GC_OK;
}

void gc_close(gc_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    ctx->items.release();
    ctx->cand.release();
    ctx->info.release();
    ctx->input.release();
    ctx->out_a.release();
    ctx->out_b.release();
    ctx->out_c.release();
    ctx->out_d.release();
    ctx->out_e.release();
    ctx->out_f.release();
    ctx->input_b.release();
    ctx->out_c_b.release();
    ctx->items_b.release();
    ctx->cand_b.release();
    ctx->out_a_b.release();
    ctx->out_b_b.release();
    ctx->out_e_b.release();
    ctx->out_f_b.release();
    ctx->seg.release();
    if (ctx->k.module) (void)hipModuleUnload(ctx->k.module);
    if (ctx->k.module_parse) (void)hipModuleUnload(ctx->k.module_parse);
    if (ctx->stream2) (void)hipStreamSynchronize(ctx->stream2);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    if (ctx->stream2) (void)hipStreamDestroy(ctx->stream2);
    delete ctx;
}

void* gc_dev_alloc(gc_ctx* ctx, size_t bytes) {
    if (!ctx) return nullptr;
    void* p = nullptr;
    (void)hipSetDevice(ctx->device);
    hipError_t e = hipMalloc(&p, bytes ? bytes : 1);
    if (e != hipSuccess) {
        hip_fail(e, "hipMalloc");
        return nullptr;
    }
    return p;
}

int gc_dev_free(gc_ctx* ctx, void* p) {
    if (!ctx) return fail(GC_EINVAL, "null context");
    hipError_t e = hipFree(p);
    return e == hipSuccess ? GC_OK : hip_fail(e, "hipFree");
}

int gc_htod(gc_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx) return fail(GC_EINVAL, "null context");
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? GC_OK : hip_fail(e, "gc_htod");
}

int gc_dtoh(gc_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx) return fail(GC_EINVAL, "null context");
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? GC_OK : hip_fail(e, "gc_dtoh");
}

int gc_dev_zero(gc_ctx* ctx, void* p, size_t bytes) {
    if (!ctx) return fail(GC_EINVAL, "null context");
    hipError_t e = hipMemsetAsync(p, 0, bytes, ctx->stream);
    return e == hipSuccess ? GC_OK : hip_fail(e, "hipMemsetAsync");
}

int gc_sync(gc_ctx* ctx) {
    if (!ctx) return fail(GC_EINVAL, "null context");
    hipError_t e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? GC_OK : hip_fail(e, "hipStreamSynchronize");
}

void* gc_stream(gc_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

}  // extern "C"
