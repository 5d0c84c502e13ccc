// gc_batch.cpp -- batched form: many inputs of uneven size, every 64 KiB
// sub-block of every input in one launch per kernel.
//
// Per input the results are exactly those of the plain per-chunk form:
// sub-blocks are tiled from each input's own start and are independent.
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../include/gcompress.h"
#include "gc_runtime.h"

using namespace gc;

struct gc_batch {
    gc_ctx* ctx = nullptr;
    const uint8_t* arena = nullptr;
    uint64_t arena_size = 0;
    std::vector<gc_batch_input> inputs;
    std::vector<uint32_t> nsub;      // sub-blocks per input
    std::vector<uint64_t> out_base;  // zstd: first sequence slot; lzma2: first packed position
    std::vector<uint64_t> cnt_base;  // zstd: first count index
    uint32_t n_z = 0, n_l = 0;       // items
    uint64_t z_bytes = 0, l_bytes = 0;
    DevBuf zitems, litems, info, seq, counts, packed;
    gc_tune tune{};
    hipEvent_t ev[6] = {};
};

namespace {

const uint32_t kMaxSeq = GC_SUB_BLOCK / 3;

uint32_t grid_for(gc_batch* b, uint32_t per_cu, uint32_t items) {
    uint64_t g = (uint64_t)per_cu * (uint64_t)b->ctx->cu_count;
    if (g > items) g = items;
    return (uint32_t)g;
}

// Enqueue one full pass on the main stream.  Items of all inputs go to
// each kernel in one launch; every workgroup takes items blockIdx.x,
// blockIdx.x + gridDim.x, ... (static grid-stride assignment).
//
// The other form built here, a device work queue drained by persistent
// workgroups and replayed from a captured hipGraph with the LZMA2 kernel
// on a forked stream, hung on its first replay on MI355X and its cause has
// not been found; it is switched off (gc_batch_set_mode refuses it).
int enqueue(gc_batch* b, bool timed) {
    gc_ctx* ctx = b->ctx;
    if (timed) (void)hipEventRecord(b->ev[0], ctx->stream);
    int rc;
    if (b->n_l) {
        rc = launch_lzma2(ctx, ctx->stream, b->arena, (const gc_lzma2_item*)b->litems.p, b->n_l, nullptr,
                          grid_for(b, b->tune.lzma2_wg_per_cu, b->n_l), (uint32_t*)b->packed.p, nullptr, nullptr, 1);
        if (rc) return rc;
    }
    if (timed) (void)hipEventRecord(b->ev[1], ctx->stream);
    if (b->n_z) {
        rc = launch_zstd(ctx, ctx->stream, b->arena, (const gc_zstd_item*)b->zitems.p, b->n_z, nullptr, nullptr,
                         grid_for(b, b->tune.cand_wg_per_cu, b->n_z), 0, (uint32_t*)b->info.p, b->seq.p,
                         GC_SEQ_FORMAT_ZSTD, nullptr, (uint32_t*)b->counts.p, nullptr, kMaxSeq);
        if (rc) return rc;
    }
    if (timed) (void)hipEventRecord(b->ev[2], ctx->stream);
    if (b->n_z) {
        rc = launch_zstd(ctx, ctx->stream, b->arena, (const gc_zstd_item*)b->zitems.p, b->n_z, nullptr, nullptr, 0,
                         grid_for(b, b->tune.parse_wg_per_cu, (b->n_z + 3) / 4), (uint32_t*)b->info.p, b->seq.p,
                         GC_SEQ_FORMAT_ZSTD, nullptr, (uint32_t*)b->counts.p, nullptr, kMaxSeq);
        if (rc) return rc;
    }
    if (timed) (void)hipEventRecord(b->ev[3], ctx->stream);
    return GC_OK;
}

}  // namespace

extern "C" {

gc_batch* gc_batch_create(gc_ctx* ctx, const uint8_t* d_arena, uint64_t arena_size, const gc_batch_input* inputs,
                          uint32_t n_inputs) {
    if (!ctx || !d_arena || (!inputs && n_inputs)) {
        set_error("gc_batch_create: null argument");
        return nullptr;
    }
    if ((uintptr_t)d_arena & 3) {
        set_error("gc_batch_create: arena must be 4-byte aligned");
        return nullptr;
    }
    std::lock_guard<std::mutex> g(ctx->mu);
    (void)hipSetDevice(ctx->device);
    gc_batch* b = new gc_batch();
    b->ctx = ctx;
    b->arena = d_arena;
    b->arena_size = arena_size;
    b->inputs.assign(inputs, inputs + n_inputs);
    b->tune = gc_tune{64, 64, 64, 0};  // grid = items unless tuned down
    std::vector<gc_zstd_item> zi;
    std::vector<gc_lzma2_item> li;
    uint64_t zseq = 0, zcnt = 0, zinfo = 0, lpos = 0;
    for (uint32_t i = 0; i < n_inputs; i++) {
        const gc_batch_input& in = inputs[i];
        if (in.algo > GC_ALGO_LZMA2 || (in.offset & 3) || in.offset > arena_size || in.size > arena_size - in.offset) {
            set_error("gc_batch_create: input " + std::to_string(i) + " is out of the arena, unaligned or of unknown algo");
            delete b;
            return nullptr;
        }
        uint32_t ns = (uint32_t)((in.size + GC_SUB_BLOCK - 1) / GC_SUB_BLOCK);
        b->nsub.push_back(ns);
        if (in.algo == GC_ALGO_ZSTD) {
            b->out_base.push_back(zseq);
            b->cnt_base.push_back(zcnt);
            for (uint32_t s = 0; s < ns; s++) {
                uint64_t start = (uint64_t)s * GC_SUB_BLOCK;
                gc_zstd_item it{};
                it.src_off = in.offset + start;
                it.avail = arena_size - it.src_off;
                it.scratch_pos = zinfo + start;
                it.seq_pos = zseq + (uint64_t)s * kMaxSeq;
                it.lit_pos = 0;
                it.count_pos = (uint32_t)(zcnt + s);
                it.n = (uint32_t)std::min<uint64_t>(GC_SUB_BLOCK, in.size - start);
                it.depth = zstd_depth_for_level(in.level);
                zi.push_back(it);
            }
            zseq += (uint64_t)ns * kMaxSeq;
            zcnt += ns;
            zinfo += (uint64_t)ns * GC_SUB_BLOCK;
            b->z_bytes += in.size;
        } else {
            b->out_base.push_back(lpos);
            b->cnt_base.push_back(0);
            for (uint32_t s = 0; s < ns; s++) {
                uint64_t start = (uint64_t)s * GC_SUB_BLOCK;
                gc_lzma2_item it{};
                it.src_off = in.offset + start;
                it.avail = arena_size - it.src_off;
                it.out_pos = lpos + start;
                it.n = (uint32_t)std::min<uint64_t>(GC_SUB_BLOCK, in.size - start);
                it.out_span = GC_SUB_BLOCK;
                it.depth = GC_LZMA2_DEPTH;
                li.push_back(it);
            }
            lpos += (uint64_t)ns * GC_SUB_BLOCK;
            b->l_bytes += in.size;
        }
    }
    // Longest work first: a full sub-block before a ragged tail.
    auto zcost = [](const gc_zstd_item& a, const gc_zstd_item& c) { return a.n > c.n; };
    std::stable_sort(zi.begin(), zi.end(), zcost);
    auto lcost = [](const gc_lzma2_item& a, const gc_lzma2_item& c) { return a.n > c.n; };
    std::stable_sort(li.begin(), li.end(), lcost);
    b->n_z = (uint32_t)zi.size();
    b->n_l = (uint32_t)li.size();
    int rc = GC_OK;
    if (b->n_z) rc = upload(ctx, b->zitems, zi.data(), zi.size() * sizeof(gc_zstd_item));
    if (!rc && b->n_l) rc = upload(ctx, b->litems, li.data(), li.size() * sizeof(gc_lzma2_item));
    if (!rc && b->n_z) rc = b->info.ensure(zinfo * 4);
    if (!rc && b->n_z) rc = b->seq.ensure(zseq * sizeof(gc_zstd_sequence));
    if (!rc && b->n_z) rc = b->counts.ensure(zcnt * 4);
    if (!rc && b->n_l) rc = b->packed.ensure(lpos * 4);
    for (int k = 0; !rc && k < 6; k++) {
        hipError_t e = hipEventCreate(&b->ev[k]);
        if (e != hipSuccess) rc = hip_fail(e, "hipEventCreate");
    }
    if (rc) {
        std::string keep = gc_last_error();
        gc_batch_destroy(b);
        set_error(keep);
        return nullptr;
    }
    return b;
}

int gc_batch_destroy(gc_batch* b) {
    if (!b) return GC_OK;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    for (auto& e : b->ev)
        if (e) (void)hipEventDestroy(e);
    b->zitems.release();
    b->litems.release();
    b->info.release();
    b->seq.release();
    b->counts.release();
    b->packed.release();
    delete b;
    return GC_OK;
}

int gc_batch_set_tune(gc_batch* b, const gc_tune* t) {
    if (!b || !t) return fail(GC_EINVAL, "null argument");
    if (t->cand_wg_per_cu < 1 || t->cand_wg_per_cu > 64 || t->lzma2_wg_per_cu < 1 || t->lzma2_wg_per_cu > 64 ||
        t->parse_wg_per_cu < 1 || t->parse_wg_per_cu > 64)
        return fail(GC_EINVAL, "tune out of range (1..64 workgroups per CU)");
    b->tune = *t;
    return GC_OK;
}

int gc_batch_get_tune(gc_batch* b, gc_tune* t) {
    if (!b || !t) return fail(GC_EINVAL, "null argument");
    *t = b->tune;
    return GC_OK;
}

int gc_batch_run(gc_batch* b) {
    if (!b) return fail(GC_EINVAL, "null batch");
    std::lock_guard<std::mutex> g(b->ctx->mu);
    (void)hipSetDevice(b->ctx->device);
    return enqueue(b, false);
}

int gc_batch_set_mode(gc_batch* b, uint32_t mode) {
    if (!b) return fail(GC_EINVAL, "null batch");
    if (mode == GC_BATCH_STATIC) return GC_OK;
    if (mode == GC_BATCH_QUEUE_GRAPH)
        return fail(GC_EINVAL,
                    "the device-queue + hipGraph form is switched off: it hung on its first replay on MI355X "
                    "and the cause has not been found");
    return fail(GC_EINVAL, "unknown batch mode");
}

int gc_batch_run_timed(gc_batch* b, float* ms) {
    if (!b || !ms) return fail(GC_EINVAL, "null argument");
    std::lock_guard<std::mutex> g(b->ctx->mu);
    (void)hipSetDevice(b->ctx->device);
    int rc = enqueue(b, true);
    if (rc) return rc;
    hipError_t e = hipEventSynchronize(b->ev[3]);
    if (e != hipSuccess) return hip_fail(e, "hipEventSynchronize");
    // [0] lzma2, [1] zstd candidate pass, [2] zstd parse
    (void)hipEventElapsedTime(&ms[0], b->ev[0], b->ev[1]);
    (void)hipEventElapsedTime(&ms[1], b->ev[1], b->ev[2]);
    (void)hipEventElapsedTime(&ms[2], b->ev[2], b->ev[3]);
    return GC_OK;
}

int gc_batch_info(gc_batch* b, uint32_t i, uint32_t* n_sub, uint64_t* out_base, uint64_t* count_base) {
    if (!b || i >= b->inputs.size()) return fail(GC_EINVAL, "bad batch or input index");
    if (n_sub) *n_sub = b->nsub[i];
    if (out_base) *out_base = b->out_base[i];
    if (count_base) *count_base = b->cnt_base[i];
    return GC_OK;
}

int gc_batch_device_outputs(gc_batch* b, gc_zstd_sequence** d_seqs, uint32_t** d_counts, uint32_t** d_packed) {
    if (!b) return fail(GC_EINVAL, "null batch");
    if (d_seqs) *d_seqs = (gc_zstd_sequence*)b->seq.p;
    if (d_counts) *d_counts = (uint32_t*)b->counts.p;
    if (d_packed) *d_packed = (uint32_t*)b->packed.p;
    return GC_OK;
}

int gc_batch_copy_zstd(gc_batch* b, uint32_t i, gc_zstd_sequence* seqs, uint32_t* counts) {
    if (!b || i >= b->inputs.size() || !counts) return fail(GC_EINVAL, "bad argument");
    if (b->inputs[i].algo != GC_ALGO_ZSTD) return fail(GC_EINVAL, "input is not a Zstd input");
    std::lock_guard<std::mutex> g(b->ctx->mu);
    hipStream_t s = b->ctx->stream;
    const uint32_t ns = b->nsub[i];
    hipError_t e = hipMemcpyAsync(counts, (uint32_t*)b->counts.p + b->cnt_base[i], ns * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    for (uint32_t k = 0; seqs && e == hipSuccess && k < ns; k++)
        if (counts[k])
            e = hipMemcpyAsync(seqs + (uint64_t)k * kMaxSeq,
                               (gc_zstd_sequence*)b->seq.p + b->out_base[i] + (uint64_t)k * kMaxSeq,
                               counts[k] * sizeof(gc_zstd_sequence), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e == hipSuccess ? GC_OK : hip_fail(e, "gc_batch_copy_zstd");
}

int gc_batch_copy_lzma2(gc_batch* b, uint32_t i, uint32_t* packed) {
    if (!b || i >= b->inputs.size() || !packed) return fail(GC_EINVAL, "bad argument");
    if (b->inputs[i].algo != GC_ALGO_LZMA2) return fail(GC_EINVAL, "input is not an LZMA2 input");
    std::lock_guard<std::mutex> g(b->ctx->mu);
    hipStream_t s = b->ctx->stream;
    hipError_t e = hipMemcpyAsync(packed, (uint32_t*)b->packed.p + b->out_base[i], b->inputs[i].size * 4,
                                  hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e == hipSuccess ? GC_OK : hip_fail(e, "gc_batch_copy_lzma2");
}

int gc_batch_autotune(gc_batch* b, uint32_t reps, gc_tune* best, float* best_ms) {
    if (!b || !best) return fail(GC_EINVAL, "null argument");
    // Output-invariant knobs only: workgroups per CU of each kernel (the grid
    // is min(knob * CUs, items), items assigned grid-stride).  The hash table
    // size and every other parameter of the algorithm are fixed.
    static const uint32_t cand_opts[] = {1, 2, 4, 64};
    static const uint32_t parse_opts[] = {2, 4, 8, 16};
    static const uint32_t lz_opts[] = {1, 2, 4, 64};
    if (reps == 0) reps = 1;
    gc_tune keep = b->tune, top = b->tune;
    float top_ms[3] = {1e30f, 1e30f, 1e30f};
    auto measure = [&](gc_tune t, float out[3]) -> int {
        int rc = gc_batch_set_tune(b, &t);
        if (rc) return rc;
        float acc[3] = {0, 0, 0}, ms[3];
        for (uint32_t r = 0; r < reps + 1; r++) {
            rc = gc_batch_run_timed(b, ms);
            if (rc) return rc;
            if (r == 0) continue;  // warm-up
            for (int k = 0; k < 3; k++) acc[k] += ms[k];
        }
        for (int k = 0; k < 3; k++) out[k] = acc[k] / reps;
        return GC_OK;
    };
    // The three kernels are independent: tune each knob on its own kernel.
    for (uint32_t c : cand_opts) {
        gc_tune t = top;
        t.cand_wg_per_cu = c;
        float ms[3];
        if (int rc = measure(t, ms)) return rc;
        if (ms[1] < top_ms[1]) {
            top_ms[1] = ms[1];
            top.cand_wg_per_cu = c;
        }
    }
    for (uint32_t p : parse_opts) {
        gc_tune t = top;
        t.parse_wg_per_cu = p;
        float ms[3];
        if (int rc = measure(t, ms)) return rc;
        if (ms[2] < top_ms[2]) {
            top_ms[2] = ms[2];
            top.parse_wg_per_cu = p;
        }
    }
    for (uint32_t l : lz_opts) {
        gc_tune t = top;
        t.lzma2_wg_per_cu = l;
        float ms[3];
        if (int rc = measure(t, ms)) return rc;
        if (ms[0] < top_ms[0]) {
            top_ms[0] = ms[0];
            top.lzma2_wg_per_cu = l;
        }
    }
    (void)keep;
    int rc = gc_batch_set_tune(b, &top);
    if (rc) return rc;
    *best = top;
    if (best_ms) *best_ms = (b->n_l ? top_ms[0] : 0) + (b->n_z ? top_ms[1] + top_ms[2] : 0);
    return GC_OK;
}

}  // extern "C"
