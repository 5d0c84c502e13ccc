// gc_batch.cpp -- batched form: many inputs of uneven size, every 64 KiB
// sub-block of every input in one launch per kernel.
//
// Two ways to run a pass:
//  * static (GC_BATCH_STATIC): grid = items (or the tuned workgroups per
//    CU), workgroup b takes items b, b + G, ...;
//  * persistent (GC_BATCH_PERSISTENT): a resident grid of each kernel drains
//    a device work queue (gc_queue: one counter per kernel, the item counts
//    in device memory), and the pass -- candidates, LZMA2 lengths, Zstd info,
//    the parse order and the parse, in that order on one stream -- is
//    captured once in a hipGraph and replayed.  The queue's counters are set
//    back to zero by the last workgroup of each kernel, so the graph holds
//    kernel nodes only (no memset or reset node), and every queue loop is
//    bounded: an error bit, read by gc_batch_check, replaces any spin.
//
// Per input the results are exactly those of the plain per-chunk form:
// sub-blocks are tiled from each input's own start and are independent.
#include <stddef.h>
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
    uint32_t n_z = 0, n_l = 0;       // items: LZMA2 at [0, n_l), Zstd at [n_l, n_l + n_z)
    uint64_t z_bytes = 0, l_bytes = 0;
    DevBuf items, cand, info, seq, counts, packed;
    DevBuf cost, order;  // Zstd: per-item parse cost (verified positions) and the parse's item order
    DevBuf queue;        // persistent form: the device work queue (gc_queue)
    gc_tune tune{};
    uint32_t mode = GC_BATCH_STATIC;
    hipGraph_t graph = nullptr;  // persistent form: one pass, captured
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev[6] = {};
};

namespace {

const uint32_t kMaxSeq = GC_SUB_BLOCK / 3;

uint32_t grid_for(gc_batch* b, uint32_t per_cu, uint32_t items) {
    uint64_t g = (uint64_t)per_cu * (uint64_t)b->ctx->cu_count;
    if (g > items) g = items;
    return (uint32_t)g;
}

void drop_graph(gc_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// Persistent form: the same kernels fed from the device work queue, on the
// main stream.  Grids are resident workgroups (the tuned count per CU); a
// kernel with no items of its kind is not launched.
int enqueue_persistent(gc_batch* b, bool timed) {
    gc_ctx* ctx = b->ctx;
    hipStream_t st = ctx->stream;
    gc_queue* q = (gc_queue*)b->queue.p;
    // every kernel gets the whole item list: the queue holds each one's first item and count
    const gc_item* all = (const gc_item*)b->items.p;
    const uint32_t n_all = b->n_l + b->n_z;
    const uint16_t* cand = (const uint16_t*)b->cand.p;
    if (timed) (void)hipEventRecord(b->ev[0], st);
    int rc = GC_OK;
    if (n_all) rc = launch_cand_q(ctx, st, b->arena, all, q, grid_for(b, b->tune.cand_wg_per_cu, n_all), (uint16_t*)b->cand.p);
    if (rc) return rc;
    if (timed) (void)hipEventRecord(b->ev[1], st);
    if (b->n_l) {
        rc = launch_lzma2_len_q(ctx, st, b->arena, all, q, grid_for(b, b->tune.elem_wg_per_cu, b->n_l), cand,
                                (uint32_t*)b->packed.p);
        if (rc) return rc;
    }
    if (timed) (void)hipEventRecord(b->ev[2], st);
    if (b->n_z) {
        rc = launch_zstd_info_q(ctx, st, b->arena, all, q, grid_for(b, b->tune.elem_wg_per_cu, b->n_z), cand,
                                (uint32_t*)b->info.p, b->tune.info_strip, (uint32_t*)b->cost.p);
        if (!rc) rc = launch_order(ctx, st, (const uint32_t*)b->cost.p, b->n_z, (uint32_t*)b->order.p, q);
        if (rc) return rc;
    }
    if (timed) (void)hipEventRecord(b->ev[3], st);
    if (b->n_z) {
        rc = launch_zstd_parse_q(ctx, st, b->arena, all, q, grid_for(b, b->tune.parse_wg_per_cu, (b->n_z + 3) / 4),
                                 (const uint32_t*)b->info.p, b->seq.p, (uint32_t*)b->counts.p, kMaxSeq,
                                 (const uint32_t*)b->order.p);
        if (rc) return rc;
    }
    if (timed) (void)hipEventRecord(b->ev[4], st);
    return GC_OK;
}

// The persistent pass captured on the main stream (one stream, kernel nodes
// only), instantiated once and replayed by gc_batch_run.
int build_graph(gc_batch* b) {
    hipStream_t st = b->ctx->stream;
    hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hip_fail(e, "hipStreamBeginCapture");
    int rc = enqueue_persistent(b, false);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(st, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    b->graph = g;
    e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        drop_graph(b);
        return hip_fail(e, "hipGraphInstantiate");
    }
    return GC_OK;
}

// The queue's error word after the passes queued so far (synchronises the
// stream); cleared once read.
int check_queue(gc_batch* b) {
    if (!b->queue.p) return GC_OK;
    hipStream_t st = b->ctx->stream;
    uint32_t err = 0;
    hipError_t e = hipMemcpyAsync(&err, (uint8_t*)b->queue.p + offsetof(gc_queue, err), 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "gc_batch_check");
    if (!err) return GC_OK;
    const uint32_t zero = 0;
    e = hipMemcpyAsync((uint8_t*)b->queue.p + offsetof(gc_queue, err), &zero, 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return fail(GC_EHIP, (err & 1u) ? "device work queue: a persistent loop passed its trip bound"
                                    : "device work queue: a kernel did not take every item");
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
    hipStream_t st = ctx->stream;
    const gc_item* all = (const gc_item*)b->items.p;
    const gc_item* zit = all + b->n_l;
    const uint32_t n_all = b->n_l + b->n_z;
    const uint16_t* cand = (const uint16_t*)b->cand.p;
    if (timed) (void)hipEventRecord(b->ev[0], st);
    int rc = launch_cand(ctx, st, b->arena, all, n_all, grid_for(b, b->tune.cand_wg_per_cu, n_all),
                         (uint16_t*)b->cand.p);
    if (rc) return rc;
    if (timed) (void)hipEventRecord(b->ev[1], st);
    if (b->n_l) {
        rc = launch_lzma2_len(ctx, st, b->arena, all, b->n_l, grid_for(b, b->tune.elem_wg_per_cu, b->n_l), cand,
                              (uint32_t*)b->packed.p, nullptr, nullptr, 1);
        if (rc) return rc;
    }
    if (timed) (void)hipEventRecord(b->ev[2], st);
    if (b->n_z) {
        rc = launch_zstd_info(ctx, st, b->arena, zit, b->n_z, grid_for(b, b->tune.elem_wg_per_cu, b->n_z), cand,
                              (uint32_t*)b->info.p, b->tune.info_strip, (uint32_t*)b->cost.p);
        if (!rc) rc = launch_order(ctx, st, (const uint32_t*)b->cost.p, b->n_z, (uint32_t*)b->order.p);
        if (rc) return rc;
    }
    if (timed) (void)hipEventRecord(b->ev[3], st);
    if (b->n_z) {
        rc = launch_zstd_parse(ctx, st, b->arena, zit, b->n_z, grid_for(b, b->tune.parse_wg_per_cu, (b->n_z + 3) / 4),
                               (const uint32_t*)b->info.p, b->seq.p, GC_SEQ_FORMAT_ZSTD, nullptr,
                               (uint32_t*)b->counts.p, nullptr, kMaxSeq, (const uint32_t*)b->order.p);
        if (rc) return rc;
    }
    if (timed) (void)hipEventRecord(b->ev[4], st);
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
    b->tune = gc_tune{64, 64, 64, 8};  // grid = items unless tuned down
    std::vector<gc_item> zi, li;
    uint64_t zseq = 0, zcnt = 0, zpos = 0, lpos = 0;
    for (uint32_t i = 0; i < n_inputs; i++) {
        const gc_batch_input& in = inputs[i];
        if (in.algo > GC_ALGO_LZMA2 || (in.offset & 3) || in.offset > arena_size || in.size > arena_size - in.offset) {
            set_error("gc_batch_create: input " + std::to_string(i) + " is out of the arena, unaligned or of unknown algo");
            delete b;
            return nullptr;
        }
        const uint32_t ns = (uint32_t)((in.size + GC_SUB_BLOCK - 1) / GC_SUB_BLOCK);
        const bool z = in.algo == GC_ALGO_ZSTD;
        b->nsub.push_back(ns);
        b->out_base.push_back(z ? zseq : lpos);
        b->cnt_base.push_back(z ? zcnt : 0);
        for (uint32_t s = 0; s < ns; s++) {
            const uint64_t start = (uint64_t)s * GC_SUB_BLOCK;
            gc_item it{};
            it.src_off = in.offset + start;
            it.avail = arena_size - it.src_off;
            it.n = (uint32_t)std::min<uint64_t>(GC_SUB_BLOCK, in.size - start);
            it.out_span = GC_SUB_BLOCK;
            if (z) {
                it.opos = zpos + start;
                it.seq_pos = zseq + (uint64_t)s * kMaxSeq;
                it.count_pos = (uint32_t)(zcnt + s);
                it.depth = zstd_depth_for_level(in.level);
                zi.push_back(it);
            } else {
                it.opos = lpos + start;
                it.depth = GC_LZMA2_DEPTH;
                li.push_back(it);
            }
        }
        if (z) {
            zseq += (uint64_t)ns * kMaxSeq;
            zcnt += ns;
            zpos += (uint64_t)ns * GC_SUB_BLOCK;
            b->z_bytes += in.size;
        } else {
            lpos += (uint64_t)ns * GC_SUB_BLOCK;
            b->l_bytes += in.size;
        }
    }
    // Longest work first: full sub-blocks before ragged tails.
    auto longer = [](const gc_item& a, const gc_item& c) { return a.n > c.n; };
    std::stable_sort(zi.begin(), zi.end(), longer);
    std::stable_sort(li.begin(), li.end(), longer);
    b->n_z = (uint32_t)zi.size();
    b->n_l = (uint32_t)li.size();
    std::vector<gc_item> all(li);
    all.insert(all.end(), zi.begin(), zi.end());
    for (size_t k = 0; k < all.size(); k++) all[k].cpos = (uint64_t)k * GC_SUB_BLOCK;  // candidates: one slot per item
    int rc = GC_OK;
    if (!all.empty()) rc = upload(ctx, b->items, all.data(), all.size() * sizeof(gc_item));
    if (!rc && !all.empty()) rc = b->cand.ensure(all.size() * (size_t)GC_SUB_BLOCK * 2);
    if (!rc && b->n_z) rc = b->info.ensure(zpos * 4);
    if (!rc && b->n_z) rc = b->seq.ensure(zseq * sizeof(gc_zstd_sequence));
    if (!rc && b->n_z) rc = b->counts.ensure(zcnt * 4);
    if (!rc && b->n_z) rc = b->cost.ensure((size_t)b->n_z * 4);
    if (!rc && b->n_z) rc = b->order.ensure((size_t)b->n_z * 4);
    if (!rc && b->n_l) rc = b->packed.ensure(lpos * 4);
    if (!rc) rc = b->queue.ensure(sizeof(gc_queue));
    if (!rc) {
        // counters at zero; the item count of each kernel
        gc_queue q0{};
        q0.n[GC_Q_CAND] = (uint32_t)all.size();
        q0.n[GC_Q_LZMA2] = b->n_l;
        q0.n[GC_Q_INFO] = b->n_z;
        q0.n[GC_Q_PARSE] = b->n_z;
        q0.base[GC_Q_INFO] = q0.base[GC_Q_PARSE] = b->n_l;  // Zstd items follow the LZMA2 ones
        rc = upload(ctx, b->queue, &q0, sizeof(q0));
    }
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
    drop_graph(b);
    for (auto& e : b->ev)
        if (e) (void)hipEventDestroy(e);
    b->items.release();
    b->cand.release();
    b->info.release();
    b->seq.release();
    b->counts.release();
    b->packed.release();
    b->cost.release();
    b->order.release();
    b->queue.release();
    delete b;
    return GC_OK;
}

int gc_batch_set_tune(gc_batch* b, const gc_tune* t) {
    if (!b || !t) return fail(GC_EINVAL, "null argument");
    if (t->cand_wg_per_cu < 1 || t->cand_wg_per_cu > 64 || t->elem_wg_per_cu < 1 || t->elem_wg_per_cu > 64 ||
        t->parse_wg_per_cu < 1 || t->parse_wg_per_cu > 64)
        return fail(GC_EINVAL, "tune out of range (1..64 workgroups per CU)");
    if (t->info_strip != 4 && t->info_strip != 8) return fail(GC_EINVAL, "info_strip must be 4 or 8");
    b->tune = *t;
    drop_graph(b);  // grid sizes are part of the captured pass
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
    if (b->mode != GC_BATCH_PERSISTENT) return enqueue(b, false);
    if (!b->exec) {
        int rc = build_graph(b);
        if (rc) return rc;
    }
    hipError_t e = hipGraphLaunch(b->exec, b->ctx->stream);
    return e == hipSuccess ? GC_OK : hip_fail(e, "hipGraphLaunch");
}

int gc_batch_set_mode(gc_batch* b, uint32_t mode) {
    if (!b) return fail(GC_EINVAL, "null batch");
    if (mode == GC_BATCH_QUEUE_GRAPH)
        return fail(GC_EINVAL,
                    "the first device-queue + hipGraph form (LZMA2 on a forked stream inside the capture, "
                    "queues reset by a captured kernel) hung on its first replay on MI355X with no cause "
                    "found and stays switched off; GC_BATCH_PERSISTENT is the persistent form");
    if (mode != GC_BATCH_STATIC && mode != GC_BATCH_PERSISTENT) return fail(GC_EINVAL, "unknown batch mode");
    std::lock_guard<std::mutex> g(b->ctx->mu);
    if (mode != b->mode) {
        drop_graph(b);
        // resident grids for the persistent kernels: 4 candidate workgroups per CU (LDS), 2 of the
        // 1024-thread per-position kernels, 8 four-wavefront parse workgroups (8 waves per SIMD);
        // the static form's grid is one workgroup per item
        b->tune = mode == GC_BATCH_PERSISTENT ? gc_tune{4, 8, 2, 8} : gc_tune{64, 64, 64, 8};
        b->mode = mode;
    }
    return GC_OK;
}

int gc_batch_check(gc_batch* b) {
    if (!b) return fail(GC_EINVAL, "null batch");
    std::lock_guard<std::mutex> g(b->ctx->mu);
    (void)hipSetDevice(b->ctx->device);
    return check_queue(b);
}

int gc_batch_run_timed(gc_batch* b, float* ms) {
    if (!b || !ms) return fail(GC_EINVAL, "null argument");
    std::lock_guard<std::mutex> g(b->ctx->mu);
    (void)hipSetDevice(b->ctx->device);
    int rc = b->mode == GC_BATCH_PERSISTENT ? enqueue_persistent(b, true) : enqueue(b, true);
    if (rc) return rc;
    hipError_t e = hipEventSynchronize(b->ev[4]);
    if (e != hipSuccess) return hip_fail(e, "hipEventSynchronize");
    // [0] candidates, [1] LZMA2 lengths, [2] Zstd info, [3] Zstd parse
    for (int k = 0; k < 4; k++) (void)hipEventElapsedTime(&ms[k], b->ev[k], b->ev[k + 1]);
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
    if (int rc = check_queue(b)) return rc;
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
    if (int rc = check_queue(b)) return rc;
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
    static const uint32_t opts[] = {1, 2, 4, 8, 64};
    if (reps == 0) reps = 1;
    gc_tune top = b->tune;
    auto measure = [&](const gc_tune& t, float out[4]) -> int {
        int rc = gc_batch_set_tune(b, &t);
        if (rc) return rc;
        float acc[4] = {0, 0, 0, 0}, ms[4];
        for (uint32_t r = 0; r < reps + 1; r++) {
            rc = gc_batch_run_timed(b, ms);
            if (rc) return rc;
            if (r == 0) continue;  // warm-up
            for (int k = 0; k < 4; k++) acc[k] += ms[k];
        }
        for (int k = 0; k < 4; k++) out[k] = acc[k] / reps;
        return GC_OK;
    };
    // The kernels are independent: tune each knob on the kernels it drives.
    float best_cand = 1e30f, best_elem = 1e30f, best_elem_z = 1e30f, best_parse = 1e30f, ms[4];
    for (uint32_t v : opts) {
        gc_tune t = top;
        t.cand_wg_per_cu = v;
        if (int rc = measure(t, ms)) return rc;
        if (ms[0] < best_cand) best_cand = ms[0], top.cand_wg_per_cu = v;
    }
    for (uint32_t v : opts) {
        gc_tune t = top;
        t.elem_wg_per_cu = v;
        if (int rc = measure(t, ms)) return rc;
        if (ms[1] + ms[2] < best_elem) best_elem = ms[1] + ms[2], top.elem_wg_per_cu = v;
    }
    for (uint32_t v : {4u, 8u}) {
        gc_tune t = top;
        t.info_strip = v;
        if (int rc = measure(t, ms)) return rc;
        if (ms[2] < best_elem_z) best_elem_z = ms[2], top.info_strip = v;
    }
    for (uint32_t v : opts) {
        gc_tune t = top;
        t.parse_wg_per_cu = v;
        if (int rc = measure(t, ms)) return rc;
        if (ms[3] < best_parse) best_parse = ms[3], top.parse_wg_per_cu = v;
    }
    int rc = gc_batch_set_tune(b, &top);
    if (rc) return rc;
    *best = top;
    if (best_ms) *best_ms = best_cand + best_elem + best_parse;
    return GC_OK;
}

}  // extern "C"
