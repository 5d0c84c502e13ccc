// gc_matchfind.cpp -- plain (per-chunk) match-finding entry points.
//
// Kernel-level forms mirror the reference kernel ABI (one launch over the
// sub-blocks of one buffer); host-level forms mirror the two Rust host
// functions: copy in, find, copy out.
#include <string.h>

#include <vector>

#include "../../include/gcompress.h"
#include "gc_runtime.h"

using namespace gc;

namespace {

bool aligned4(const void* p) { return ((uintptr_t)p & 3u) == 0; }

uint64_t div_up(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

// Scratch positions per item: the sub-block size rounded up to 8, so that
// per-position passes can move whole 8-position strips.
uint64_t pitch(uint32_t sub) { return ((uint64_t)sub + 7) & ~(uint64_t)7; }

// Items for a buffer tiled in sub-blocks of `sub` bytes from its start.
// internal_out: the per-position output is scratch (Zstd info) and takes the
// padded pitch; otherwise it is the caller's array indexed by position.
void make_items(std::vector<gc_item>& v, uint64_t total, uint32_t sub, uint32_t nsub, uint32_t max_seq,
                uint32_t depth, bool internal_out) {
    v.resize(nsub);
    for (uint32_t s = 0; s < nsub; s++) {
        uint64_t start = (uint64_t)s * sub;
        gc_item& it = v[s];
        it.src_off = start;
        it.avail = total - start;
        it.cpos = (uint64_t)s * pitch(sub);
        it.opos = internal_out ? (uint64_t)s * pitch(sub) : start;
        it.seq_pos = (uint64_t)s * max_seq;
        it.lit_pos = start;
        it.count_pos = s;
        it.n = (uint32_t)((total - start) < sub ? (total - start) : sub);
        it.depth = depth;
        it.out_span = sub;
    }
}

int check_sub(uint32_t sub) {
    if (sub == 0 || sub > GC_SUB_BLOCK)
        return fail(GC_EINVAL, "sub-block size must be 1..65536 (positions are kept in 16 bits)");
    return GC_OK;
}

int run_lzma2(gc_ctx* ctx, const uint8_t* d_in, uint64_t total, uint32_t sub, uint32_t nsub, uint32_t* d_packed,
              gc_lzma_match* d_m, uint32_t* d_c, uint32_t max_matches, uint32_t depth, uint8_t* d_triage = nullptr) {
    std::vector<gc_item> items;
    make_items(items, total, sub, nsub, 0, depth, false);
    int rc = ctx->cand.ensure((size_t)nsub * pitch(sub) * 2 + 256);
    if (!rc) rc = upload(ctx, ctx->items, items.data(), items.size() * sizeof(gc_item));
    if (rc) return rc;
    const gc_item* d_items = (const gc_item*)ctx->items.p;
    rc = launch_cand(ctx, ctx->stream, d_in, d_items, nsub, nsub, (uint16_t*)ctx->cand.p);
    if (rc) return rc;
    return launch_lzma2_len(ctx, ctx->stream, d_in, d_items, nsub, nsub, (const uint16_t*)ctx->cand.p, d_packed, d_m,
                            d_c, max_matches, d_triage);
}

int run_zstd(gc_ctx* ctx, const uint8_t* d_in, uint64_t total, uint32_t chunk, uint32_t nsub, void* d_seq,
             uint32_t fmt, uint8_t* d_lit, uint32_t* d_sc, uint32_t* d_lc, uint32_t depth, uint32_t max_seq) {
    std::vector<gc_item> items;
    make_items(items, total, chunk, nsub, max_seq, depth, true);
    // the info words, then per sub-block a cost word and the parse order
    const size_t info_bytes = (size_t)nsub * pitch(chunk) * 4;
    int rc = ctx->cand.ensure((size_t)nsub * pitch(chunk) * 2 + 256);
    if (!rc) rc = ctx->info.ensure(info_bytes + (size_t)nsub * 8 + 256);
    if (!rc) rc = upload(ctx, ctx->items, items.data(), items.size() * sizeof(gc_item));
    if (rc) return rc;
    const gc_item* d_items = (const gc_item*)ctx->items.p;
    const bool seg = nsub <= ctx->seg_max_sub;  // too few sub-blocks to fill the GPU with one walk each
    uint32_t* cost = seg ? nullptr : (uint32_t*)((uint8_t*)ctx->info.p + info_bytes);
    uint32_t* order = seg ? nullptr : cost + nsub;
    rc = launch_cand(ctx, ctx->stream, d_in, d_items, nsub, nsub, (uint16_t*)ctx->cand.p);
    if (!rc)
        rc = launch_zstd_info(ctx, ctx->stream, d_in, d_items, nsub, nsub, (const uint16_t*)ctx->cand.p,
                              (uint32_t*)ctx->info.p, 8, cost);
    if (rc) return rc;
    if (seg)
        return launch_zstd_parse_seg(ctx, ctx->stream, d_in, d_items, nsub, (const uint32_t*)ctx->info.p, d_seq, fmt,
                                     d_lit, d_sc, d_lc, max_seq, depth);
    // the sub-blocks with the most walk steps start first, so that the
    // longest walks do not set the launch's end
    rc = launch_order(ctx, ctx->stream, cost, nsub, order);
    if (rc) return rc;
    return launch_zstd_parse(ctx, ctx->stream, d_in, d_items, nsub, (uint32_t)div_up(nsub, 4),
                             (const uint32_t*)ctx->info.p, d_seq, fmt, d_lit, d_sc, d_lc, max_seq, order);
}

// Device copy of a host buffer, padded so that aligned reads near the end
// stay inside the allocation.
int stage_input(gc_ctx* ctx, const uint8_t* chunk, size_t n) {
    int rc = ctx->input.ensure(n + 64);
    if (rc) return rc;
    hipError_t e = hipMemcpyAsync(ctx->input.p, chunk, n, hipMemcpyHostToDevice, ctx->stream);
    return e == hipSuccess ? GC_OK : hip_fail(e, "hipMemcpyAsync(input)");
}

// Frames on the GPU (one per 64 KiB sub-block, compress_zstd_custom.rs:47-203 as the reference's
// unused zstd_encode_block would code them), in three steps on ctx->stream.  The buffers shared by
// the steps of one call: out_a sequences, out_b / out_d sequence and literal counts, out_e one frame
// slot per sub-block, out_f the frame sizes, their offsets and the frames' order.
const uint32_t kFrameMaxSeq = GC_SUB_BLOCK / 3;  // compress_zstd_custom.rs:55
const uint32_t kFrameSlot = GC_SUB_BLOCK + 16;

int frame_buffers(gc_ctx* ctx, uint64_t nsub) {
    int rc = ctx->out_a.ensure((size_t)nsub * kFrameMaxSeq * sizeof(gc_zstd_sequence));
    if (!rc) rc = ctx->out_b.ensure(nsub * 4);
    if (!rc) rc = ctx->out_d.ensure(nsub * 4);
    if (!rc) rc = ctx->out_e.ensure((size_t)nsub * kFrameSlot);
    if (!rc) rc = ctx->out_f.ensure(nsub * 16 + 32);
    return rc;
}

uint32_t* frame_sizes(gc_ctx* ctx) { return (uint32_t*)ctx->out_f.p; }
uint64_t* frame_offsets(gc_ctx* ctx, uint64_t nsub) {
    return (uint64_t*)((uint8_t*)ctx->out_f.p + ((nsub * 4 + 15) & ~(uint64_t)15));
}

// Step 1: match finding with literals (into d_lp, nsub x kFrameSlot bytes) and frame coding of the n
// bytes at d_in.
int frames_launch(gc_ctx* ctx, const uint8_t* d_in, size_t n, uint32_t level, uint8_t* d_lp) {
    const uint64_t nsub = div_up(n, GC_SUB_BLOCK);
    int rc = run_zstd(ctx, d_in, n, GC_SUB_BLOCK, (uint32_t)nsub, ctx->out_a.p, GC_SEQ_FORMAT_ZSTD, d_lp,
                      (uint32_t*)ctx->out_b.p, (uint32_t*)ctx->out_d.p, zstd_depth_for_level(level), kFrameMaxSeq);
    if (rc) return rc;
    // the items run_zstd uploaded describe the same sub-blocks
    return launch_zstd_encode(ctx, ctx->stream, d_in, (const gc_item*)ctx->items.p, (uint32_t)nsub, ctx->out_a.p,
                              GC_SEQ_FORMAT_ZSTD, d_lp, (const uint32_t*)ctx->out_b.p, (const uint32_t*)ctx->out_d.p,
                              (uint8_t*)ctx->out_e.p, frame_sizes(ctx), kFrameSlot, kFrameMaxSeq,
                              (uint32_t*)(frame_offsets(ctx, nsub) + nsub));
}

// Step 2: waits for step 1, reads the frame sizes (into sizes[]) and packs the frames back to back into
// d_lp (the literals are no longer needed); *total = their bytes, at most cap.
int frames_pack(gc_ctx* ctx, uint64_t nsub, uint8_t* d_lp, uint64_t* sizes, size_t cap, uint64_t* total) {
    std::vector<uint32_t> fs(nsub);
    hipError_t e = hipMemcpyAsync(fs.data(), frame_sizes(ctx), nsub * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(e, "frame sizes");
    std::vector<uint64_t> offs(nsub);
    uint64_t t = 0;
    for (uint64_t s = 0; s < nsub; s++) {
        offs[s] = t;
        sizes[s] = fs[s];
        t += fs[s];
    }
    if (t > cap) return fail(GC_ESIZE, "output capacity below the frames' total size");
    uint64_t* d_offs = frame_offsets(ctx, nsub);
    e = hipMemcpyAsync(d_offs, offs.data(), nsub * 8, hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) return hip_fail(e, "frame offsets");
    int rc = launch_gather_bytes(ctx, ctx->stream, (const uint8_t*)ctx->out_e.p, frame_sizes(ctx), d_offs, d_lp,
                                 kFrameSlot, (uint32_t)nsub);
    // (offs is pageable: the copy has read it before hipMemcpyAsync returned)
    *total = t;
    return rc;
}

// LZMA2 streams on the GPU from a host buffer, in two steps on one stream; the buffers of one pass (the
// NVLZ pipeline runs two passes at once, one set each): the input, the items, the candidates, the packed
// matches, one output slot per sub-block, the stream sizes and offsets, the coder's costs and order, and
// the streams back to back.
struct LzSet {
    hipStream_t stream;
    gc::DevBuf *in, *items, *cand, *packed, *slots, *meta, *scratch, *streams;
};

LzSet lz_set(gc_ctx* ctx, int j) {
    if (j == 0)
        return {ctx->stream, &ctx->input, &ctx->items, &ctx->cand, &ctx->out_a,
                &ctx->out_e, &ctx->out_f, &ctx->out_b, &ctx->out_c};
    return {ctx->stream2, &ctx->input_b, &ctx->items_b, &ctx->cand_b, &ctx->out_a_b,
            &ctx->out_e_b, &ctx->out_f_b, &ctx->out_b_b, &ctx->out_c_b};
}

// Step 1: copy the n bytes at src in (a copy from pageable memory holds the calling thread until it is
// done), find the matches and launch the coder.
int lz_launch(gc_ctx* ctx, const LzSet& s, const uint8_t* src, size_t n) {
    const uint64_t nsub = div_up(n, GC_SUB_BLOCK);
    int rc = s.in->ensure(n + 64);
    if (!rc) rc = s.cand->ensure((size_t)nsub * pitch(GC_SUB_BLOCK) * 2 + 256);
    if (!rc) rc = s.packed->ensure(n * 4);
    if (!rc) rc = s.slots->ensure((size_t)nsub * GC_LZMA2_GPU_SLOT);
    if (!rc) rc = s.meta->ensure(nsub * 12 + 16);
    if (!rc) rc = s.scratch->ensure(nsub * 8);
    if (!rc) rc = s.streams->ensure(n + 4 * nsub + 16);
    if (rc) return rc;
    hipError_t e = hipMemcpyAsync(s.in->p, src, n, hipMemcpyHostToDevice, s.stream);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync(input)");
    std::vector<gc_item> items;
    make_items(items, n, GC_SUB_BLOCK, (uint32_t)nsub, 0, GC_LZMA2_DEPTH, false);
    rc = upload_on(ctx, s.stream, *s.items, items.data(), items.size() * sizeof(gc_item));
    if (rc) return rc;
    const gc_item* d_items = (const gc_item*)s.items->p;
    const uint8_t* d_in = (const uint8_t*)s.in->p;
    rc = launch_cand(ctx, s.stream, d_in, d_items, (uint32_t)nsub, (uint32_t)nsub, (uint16_t*)s.cand->p);
    if (!rc)
        rc = launch_lzma2_len(ctx, s.stream, d_in, d_items, (uint32_t)nsub, (uint32_t)nsub,
                              (const uint16_t*)s.cand->p, (uint32_t*)s.packed->p, nullptr, nullptr, 1);
    if (!rc)
        rc = launch_lzma2_encode_ordered(ctx, s.stream, d_in, d_items, (uint32_t)nsub, (const uint32_t*)s.packed->p,
                                         (uint8_t*)s.slots->p, (uint32_t*)s.meta->p, (uint32_t*)s.scratch->p);
    return rc;
}

// Step 2: wait for the coder, gather the streams back to back and copy them to out (at most cap bytes);
// sizes[] (one per sub-block) and *len set.
int lz_finish(gc_ctx* ctx, const LzSet& s, size_t n, uint8_t* out, size_t cap, uint64_t* sizes, size_t* len) {
    const uint64_t nsub = div_up(n, GC_SUB_BLOCK);
    uint32_t* d_sizes = (uint32_t*)s.meta->p;
    uint64_t* d_offs = (uint64_t*)((uint8_t*)s.meta->p + ((nsub * 4 + 15) & ~(uint64_t)15));
    std::vector<uint32_t> fs(nsub);
    hipError_t e = hipMemcpyAsync(fs.data(), d_sizes, nsub * 4, hipMemcpyDeviceToHost, s.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s.stream);
    if (e != hipSuccess) return hip_fail(e, "LZMA2 stream sizes");
    std::vector<uint64_t> offs(nsub);
    uint64_t t = 0;
    for (uint64_t k = 0; k < nsub; k++) {
        offs[k] = t;
        sizes[k] = fs[k];
        t += fs[k];
    }
    if (t > cap) return fail(GC_ESIZE, "output capacity below the streams' total size");
    e = hipMemcpyAsync(d_offs, offs.data(), nsub * 8, hipMemcpyHostToDevice, s.stream);
    if (e != hipSuccess) return hip_fail(e, "LZMA2 stream offsets");
    int rc = launch_gather_bytes(ctx, s.stream, (const uint8_t*)s.slots->p + GC_LZMA2_GPU_HEAD, d_sizes, d_offs,
                                 (uint8_t*)s.streams->p, GC_LZMA2_GPU_SLOT, (uint32_t)nsub);
    if (rc) return rc;
    e = hipMemcpyAsync(out, s.streams->p, t, hipMemcpyDeviceToHost, s.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s.stream);
    if (e != hipSuccess) return hip_fail(e, "LZMA2 streams");
    *len = (size_t)t;
    return GC_OK;
}

// Frames of the n bytes at src (host memory), packed back to back into out (at most cap bytes),
// sizes[] one per sub-block, *len their total.
int frames_pipelined(gc_ctx* ctx, const uint8_t* src, size_t n, uint32_t level, uint8_t* out, size_t cap,
                     uint64_t* sizes, size_t* len) {
    size_t o = 0;
    // Pieces of 128 MiB, two in flight: while the GPU codes piece k on ctx->stream, this thread copies
    // piece k + 1 in and piece k - 1's frames out on stream2 (copies from and to pageable memory hold
    // the calling thread until they are done; the kernels run meanwhile).  A slot j = k & 1 holds a
    // piece's input and its literals, then its packed frames, until they are copied out; the other
    // buffers serve one piece at a time in stream order.
    const size_t piece = (size_t)128 << 20;
    const uint64_t psub = div_up(n < piece ? n : piece, GC_SUB_BLOCK);
    std::lock_guard<std::mutex> g(ctx->mu);
    (void)hipSetDevice(ctx->device);
    gc::DevBuf* in[2] = {&ctx->input, &ctx->input_b};
    gc::DevBuf* lp[2] = {&ctx->out_c, &ctx->out_c_b};
    int rc = frame_buffers(ctx, psub);
    for (int j = 0; j < 2 && !rc; j++) {
        rc = in[j]->ensure(psub * GC_SUB_BLOCK + 64);
        if (!rc) rc = lp[j]->ensure((size_t)psub * kFrameSlot);
    }
    if (rc) return rc;
    hipEvent_t packed = nullptr;
    hipError_t e = hipEventCreateWithFlags(&packed, hipEventDisableTiming);
    if (e != hipSuccess) return hip_fail(e, "hipEventCreate");
    // a piece's frames packed in its slot j (the pack's end recorded as `packed` on ctx->stream, before
    // the next piece's kernels), out to the container
    auto copy_out = [&](int j, uint64_t t) -> int {
        hipError_t e2 = hipStreamWaitEvent(ctx->stream2, packed, 0);
        if (e2 == hipSuccess) e2 = hipMemcpyAsync(out + o, lp[j]->p, t, hipMemcpyDeviceToHost, ctx->stream2);
        if (e2 == hipSuccess) e2 = hipStreamSynchronize(ctx->stream2);
        if (e2 != hipSuccess) return hip_fail(e2, "GPU frames: frames out");
        o += t;
        return GC_OK;
    };
    size_t prev0 = 0, prev_len = 0;  // piece k - 1
    uint64_t t_prev = 0;
    int k = 0;
    for (size_t c0 = 0; c0 < n && !rc; c0 += piece, k++) {
        const int j = k & 1;
        const size_t len = n - c0 < piece ? n - c0 : piece;
        e = hipMemcpyAsync(in[j]->p, src + c0, len, hipMemcpyHostToDevice, ctx->stream2);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream2);
        if (e != hipSuccess) {
            rc = hip_fail(e, "GPU frames: input");
            break;
        }
        if (k > 0) {  // piece k - 1 coded: its frames packed before piece k's kernels take the shared buffers
            const uint64_t ps = div_up(prev_len, GC_SUB_BLOCK);
            rc = frames_pack(ctx, ps, (uint8_t*)lp[j ^ 1]->p, sizes + prev0 / GC_SUB_BLOCK, cap - o, &t_prev);
            if (!rc && (e = hipEventRecord(packed, ctx->stream)) != hipSuccess) rc = hip_fail(e, "hipEventRecord");
            if (rc) break;
        }
        rc = frames_launch(ctx, (const uint8_t*)in[j]->p, len, level, (uint8_t*)lp[j]->p);
        if (!rc && k > 0) rc = copy_out(j ^ 1, t_prev);
        prev0 = c0;
        prev_len = len;
    }
    if (!rc) {  // the last piece
        const uint64_t ps = div_up(prev_len, GC_SUB_BLOCK);
        rc = frames_pack(ctx, ps, (uint8_t*)lp[(k - 1) & 1]->p, sizes + prev0 / GC_SUB_BLOCK, cap - o, &t_prev);
        if (!rc && (e = hipEventRecord(packed, ctx->stream)) != hipSuccess) rc = hip_fail(e, "hipEventRecord");
        if (!rc) rc = copy_out((k - 1) & 1, t_prev);
    }
    // nothing of this call stays in flight when it returns, on success or not
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamSynchronize(ctx->stream2);
    (void)hipEventDestroy(packed);
    if (rc) return rc;
    *len = o;
    return GC_OK;
}

// The streams of one piece of at most 2 GiB at src, coded in two halves, one coder launch each on the
// context's two streams: the second half's copy in overlaps the first half's coding, and the first
// half's streams are copied out while the second half's coder runs.  The caller holds ctx->mu.
int lz_two_halves(gc_ctx* ctx, const uint8_t* src, size_t len, uint8_t* out, size_t cap, uint64_t* sizes,
                  size_t* written) {
    const LzSet set[2] = {lz_set(ctx, 0), lz_set(ctx, 1)};
    const size_t half = len > GC_SUB_BLOCK ? (size_t)div_up(len / 2, GC_SUB_BLOCK) * GC_SUB_BLOCK : len;
    const size_t part[2] = {half, len - half};
    size_t o = 0;
    int rc = GC_OK;
    for (int j = 0; j < 2 && !rc; j++)
        if (part[j]) rc = lz_launch(ctx, set[j], src + (j ? half : 0), part[j]);
    for (int j = 0; j < 2 && !rc; j++) {
        if (!part[j]) continue;
        size_t w = 0;
        rc = lz_finish(ctx, set[j], part[j], out + o, cap - o, sizes + (j ? half : 0) / GC_SUB_BLOCK, &w);
        o += w;
    }
    // nothing stays in flight when this returns, on success or not
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamSynchronize(ctx->stream2);
    *written = o;
    return rc;
}

}  // namespace

extern "C" {

int gc_zstd_match_find_dev(gc_ctx* ctx, const uint8_t* d_input, void* d_sequences, uint32_t seq_format,
                           uint8_t* d_literals, uint32_t* d_seq_counts, uint32_t* d_lit_counts, uint32_t chunk_size,
                           uint32_t num_chunks, uint64_t total_input_size, uint32_t search_depth,
                           uint32_t max_sequences_per_chunk) {
    if (!ctx || !d_input || !d_sequences || !d_seq_counts) return fail(GC_EINVAL, "null argument");
    if (int rc = check_sub(chunk_size)) return rc;
    if (!aligned4(d_input)) return fail(GC_EINVAL, "d_input must be 4-byte aligned");
    if (seq_format > GC_SEQ_FORMAT_REF) return fail(GC_EINVAL, "unknown seq_format");
    if ((uint64_t)num_chunks > div_up(total_input_size, chunk_size))
        return fail(GC_EINVAL, "num_chunks exceeds the sub-chunks of total_input_size");
    if (num_chunks == 0) return GC_OK;
    return run_zstd(ctx, d_input, total_input_size, chunk_size, num_chunks, d_sequences, seq_format, d_literals,
                    d_seq_counts, d_lit_counts, search_depth, max_sequences_per_chunk);
}

int gc_lzma2_match_find_dev(gc_ctx* ctx, const uint8_t* d_input, gc_lzma_match* d_matches, uint32_t* d_match_counts,
                            uint32_t sub_block_size, uint32_t num_sub_blocks, uint64_t total_input_size,
                            uint32_t max_matches_per_pos, uint32_t search_depth) {
    if (!ctx || !d_input || !d_matches || !d_match_counts) return fail(GC_EINVAL, "null argument");
    if (int rc = check_sub(sub_block_size)) return rc;
    if (!aligned4(d_input)) return fail(GC_EINVAL, "d_input must be 4-byte aligned");
    if (max_matches_per_pos == 0) return fail(GC_EINVAL, "max_matches_per_pos must be >= 1");
    if ((uint64_t)num_sub_blocks > div_up(total_input_size, sub_block_size))
        return fail(GC_EINVAL, "num_sub_blocks exceeds the sub-blocks of total_input_size");
    if (num_sub_blocks == 0) return GC_OK;
    return run_lzma2(ctx, d_input, total_input_size, sub_block_size, num_sub_blocks, nullptr, d_matches,
                     d_match_counts, max_matches_per_pos, search_depth);
}

int gc_lzma2_match_find_packed_dev(gc_ctx* ctx, const uint8_t* d_input, uint32_t* d_packed, uint32_t sub_block_size,
                                   uint64_t total_input_size, uint32_t search_depth) {
    if (!ctx || !d_input || !d_packed) return fail(GC_EINVAL, "null argument");
    if (int rc = check_sub(sub_block_size)) return rc;
    if (!aligned4(d_input)) return fail(GC_EINVAL, "d_input must be 4-byte aligned");
    uint64_t nsub = div_up(total_input_size, sub_block_size);
    if (nsub == 0) return GC_OK;
    if (nsub > 0xFFFFFFFFull) return fail(GC_EINVAL, "input too large");
    return run_lzma2(ctx, d_input, total_input_size, sub_block_size, (uint32_t)nsub, d_packed, nullptr, nullptr, 1,
                     search_depth);
}

int gc_lzma2_find_matches_packed(gc_ctx* ctx, const uint8_t* chunk, size_t n, uint32_t* packed) {
    if (!ctx || (!chunk && n) || (!packed && n)) return fail(GC_EINVAL, "null argument");
    if (n == 0) return GC_OK;
    std::lock_guard<std::mutex> g(ctx->mu);
    (void)hipSetDevice(ctx->device);
    int rc = stage_input(ctx, chunk, n);
    if (!rc) rc = ctx->out_a.ensure(n * 4);
    if (!rc)
        rc = gc_lzma2_match_find_packed_dev(ctx, (const uint8_t*)ctx->input.p, (uint32_t*)ctx->out_a.p, GC_SUB_BLOCK,
                                            n, GC_LZMA2_DEPTH);
    if (!rc) {
        hipError_t e = hipMemcpyAsync(packed, ctx->out_a.p, n * 4, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = hip_fail(e, "gc_lzma2_find_matches_packed: copy out");
    }
    return rc;
}

int gc_lzma2_triage_dev(gc_ctx* ctx, const uint8_t* d_input, uint64_t n, uint8_t* d_any_match) {
    if (!ctx || !d_input || !d_any_match) return fail(GC_EINVAL, "null argument");
    if (!aligned4(d_input)) return fail(GC_EINVAL, "d_input must be 4-byte aligned");
    const uint64_t nsub = div_up(n, GC_SUB_BLOCK);
    if (nsub == 0) return GC_OK;
    if (nsub > 0xFFFFFFFFull) return fail(GC_EINVAL, "input too large");
    return run_lzma2(ctx, d_input, n, GC_SUB_BLOCK, (uint32_t)nsub, nullptr, nullptr, nullptr, 1, GC_LZMA2_DEPTH,
                     d_any_match);
}

int gc_lzma2_triage(gc_ctx* ctx, const uint8_t* chunk, size_t n, uint8_t* any_match) {
    if (!ctx || (!chunk && n) || (!any_match && n)) return fail(GC_EINVAL, "null argument");
    if (n == 0) return GC_OK;
    const size_t nsub = (size_t)div_up(n, GC_SUB_BLOCK);
    std::lock_guard<std::mutex> g(ctx->mu);
    (void)hipSetDevice(ctx->device);
    int rc = stage_input(ctx, chunk, n);
    if (!rc) rc = ctx->out_a.ensure(nsub);
    if (!rc) rc = gc_lzma2_triage_dev(ctx, (const uint8_t*)ctx->input.p, n, (uint8_t*)ctx->out_a.p);
    if (!rc) {
        hipError_t e = hipMemcpyAsync(any_match, ctx->out_a.p, nsub, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = hip_fail(e, "gc_lzma2_triage: copy out");
    }
    return rc;
}

int gc_gpu_find_matches(gc_ctx* ctx, const uint8_t* chunk, size_t n, uint32_t* match_distances,
                        uint32_t* match_lengths, uint32_t* match_counts) {
    if (!ctx || !match_distances || !match_lengths || !match_counts) return fail(GC_EINVAL, "null argument");
    const size_t nsub = (size_t)div_up(n, GC_SUB_BLOCK);
    const size_t slots = nsub * GC_SUB_BLOCK;
    std::vector<uint32_t> packed(n);
    int rc = gc_lzma2_find_matches_packed(ctx, chunk, n, packed.data());
    if (rc) return rc;
    // GpuMatchData layout: slot k of position q at q*8 + k, unused slots 0.
    memset(match_distances, 0, slots * GC_LZMA2_MAX_MATCHES * sizeof(uint32_t));
    memset(match_lengths, 0, slots * GC_LZMA2_MAX_MATCHES * sizeof(uint32_t));
    memset(match_counts, 0, slots * sizeof(uint32_t));
    for (size_t q = 0; q < n; q++) {
        uint32_t v = packed[q];
        if (!v) continue;
        match_distances[q * GC_LZMA2_MAX_MATCHES] = v >> 16;
        match_lengths[q * GC_LZMA2_MAX_MATCHES] = v & 0xFFFFu;
        match_counts[q] = 1;
    }
    return GC_OK;
}

int gc_zstd_find_sequences(gc_ctx* ctx, const uint8_t* chunk, size_t n, uint32_t level, gc_zstd_sequence* seqs,
                           uint32_t* seq_counts) {
    if (!ctx || (!chunk && n) || !seqs || !seq_counts) return fail(GC_EINVAL, "null argument");
    if (n == 0) return GC_OK;
    const uint32_t max_seq = GC_SUB_BLOCK / 3;  // compress_zstd_custom.rs:55
    const uint64_t nsub = div_up(n, GC_SUB_BLOCK);
    std::lock_guard<std::mutex> g(ctx->mu);
    (void)hipSetDevice(ctx->device);
    int rc = stage_input(ctx, chunk, n);
    const size_t seq_bytes = (size_t)nsub * max_seq * sizeof(gc_zstd_sequence);
    if (!rc) rc = ctx->out_a.ensure(seq_bytes);
    if (!rc) rc = ctx->out_b.ensure(nsub * 4);
    if (!rc)
        rc = run_zstd(ctx, (const uint8_t*)ctx->input.p, n, GC_SUB_BLOCK, (uint32_t)nsub, ctx->out_a.p,
                      GC_SEQ_FORMAT_ZSTD, nullptr, (uint32_t*)ctx->out_b.p, nullptr, zstd_depth_for_level(level),
                      max_seq);
    if (!rc) {
        hipError_t e = hipMemcpyAsync(seq_counts, ctx->out_b.p, nsub * 4, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(seqs, ctx->out_a.p, seq_bytes, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = hip_fail(e, "gc_zstd_find_sequences: copy out");
    }
    return rc;
}

int gc_zstd_find_sequences_packed(gc_ctx* ctx, const uint8_t* chunk, size_t n, uint32_t level, gc_zstd_sequence* seqs,
                                  size_t seq_capacity, uint32_t* seq_counts, size_t* total) {
    if (!ctx || (!chunk && n) || !seq_counts || !total) return fail(GC_EINVAL, "null argument");
    *total = 0;
    if (n == 0) return GC_OK;
    const uint32_t max_seq = GC_SUB_BLOCK / 3;  // compress_zstd_custom.rs:55
    const uint64_t nsub = div_up(n, GC_SUB_BLOCK);
    std::lock_guard<std::mutex> g(ctx->mu);
    (void)hipSetDevice(ctx->device);
    int rc = stage_input(ctx, chunk, n);
    const size_t seq_bytes = (size_t)nsub * max_seq * sizeof(gc_zstd_sequence);
    if (!rc) rc = ctx->out_a.ensure(seq_bytes);
    if (!rc) rc = ctx->out_b.ensure(nsub * 4);
    if (!rc)
        rc = run_zstd(ctx, (const uint8_t*)ctx->input.p, n, GC_SUB_BLOCK, (uint32_t)nsub, ctx->out_a.p,
                      GC_SEQ_FORMAT_ZSTD, nullptr, (uint32_t*)ctx->out_b.p, nullptr, zstd_depth_for_level(level),
                      max_seq);
    if (rc) return rc;
    hipError_t e = hipMemcpyAsync(seq_counts, ctx->out_b.p, nsub * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(e, "gc_zstd_find_sequences_packed: counts");
    std::vector<uint64_t> offs(nsub);
    uint64_t t = 0;
    for (uint64_t s = 0; s < nsub; s++) {
        offs[s] = t;
        t += seq_counts[s];
    }
    *total = (size_t)t;
    if (!seqs) return GC_OK;
    if (t > seq_capacity) return fail(GC_ESIZE, "sequence capacity below the number of sequences found");
    if (t == 0) return GC_OK;
    rc = upload(ctx, ctx->out_c, offs.data(), nsub * 8);
    if (!rc) rc = ctx->out_d.ensure(t * sizeof(gc_zstd_sequence));
    if (!rc)
        rc = launch_gather(ctx, ctx->stream, ctx->out_a.p, (const uint32_t*)ctx->out_b.p, (const uint64_t*)ctx->out_c.p,
                           ctx->out_d.p, (uint32_t)nsub, max_seq);
    if (rc) return rc;
    e = hipMemcpyAsync(seqs, ctx->out_d.p, t * sizeof(gc_zstd_sequence), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? GC_OK : hip_fail(e, "gc_zstd_find_sequences_packed: sequences");
}

int gc_zstd_encode_block_dev(gc_ctx* ctx, const uint8_t* d_input, const void* d_sequences, uint32_t seq_format,
                             const uint8_t* d_literals, const uint32_t* d_seq_counts, const uint32_t* d_lit_counts,
                             uint8_t* d_output, uint32_t* d_output_sizes, uint32_t chunk_size, uint32_t num_chunks,
                             uint64_t total_input_size, uint32_t max_sequences_per_chunk, uint32_t max_frame_size) {
    if (!ctx || !d_input || !d_sequences || !d_literals || !d_seq_counts || !d_lit_counts || !d_output ||
        !d_output_sizes)
        return fail(GC_EINVAL, "null argument");
    if (int rc = check_sub(chunk_size)) return rc;
    if (seq_format > GC_SEQ_FORMAT_REF) return fail(GC_EINVAL, "unknown seq_format");
    if ((uint64_t)max_frame_size < (uint64_t)chunk_size + 16)
        return fail(GC_EINVAL, "max_frame_size must be at least chunk_size + 16 (the raw frame)");
    if ((uint64_t)num_chunks > div_up(total_input_size, chunk_size))
        return fail(GC_EINVAL, "num_chunks exceeds the sub-chunks of total_input_size");
    if (num_chunks == 0) return GC_OK;
    std::vector<gc_item> items;
    make_items(items, total_input_size, chunk_size, num_chunks, max_sequences_per_chunk, 0, true);
    int rc = upload(ctx, ctx->items, items.data(), items.size() * sizeof(gc_item));
    if (!rc) rc = ctx->out_b.ensure((size_t)num_chunks * 4);  // the frames' order
    if (rc) return rc;
    return launch_zstd_encode(ctx, ctx->stream, d_input, (const gc_item*)ctx->items.p, num_chunks, d_sequences,
                              seq_format, d_literals, d_seq_counts, d_lit_counts, d_output, d_output_sizes,
                              max_frame_size, max_sequences_per_chunk, (uint32_t*)ctx->out_b.p);
}

int gc_zstd_frames_gpu(gc_ctx* ctx, const uint8_t* chunk, size_t n, uint32_t level, uint8_t* out, size_t cap,
                       uint64_t* sizes, size_t* out_len) {
    if (!ctx || (!chunk && n) || !out || !sizes || !out_len) return fail(GC_EINVAL, "null argument");
    *out_len = 0;
    if (n == 0) return GC_OK;
    return frames_pipelined(ctx, chunk, n, level, out, cap, sizes, out_len);
}

int gc_nvzs_compress_gpu(gc_ctx* ctx, const uint8_t* src, size_t n, uint32_t level, uint8_t* out, size_t cap,
                         size_t* out_len) {
    if (!ctx || (!src && n) || !out || !out_len) return fail(GC_EINVAL, "null argument");
    const uint64_t nsub = div_up(n, GC_SUB_BLOCK);
    const size_t head = 28 + 8 * nsub;
    if (cap < head + n + 16 * nsub) return fail(GC_ESIZE, "output capacity below gc_nvzs_gpu_bound");
    memcpy(out, "NVZS", 4);  // pipeline.rs:187-260
    auto put64 = [](uint8_t* p, uint64_t v) {
        for (int i = 0; i < 8; i++) p[i] = (uint8_t)(v >> (8 * i));
    };
    put64(out + 4, n);
    put64(out + 12, GC_SUB_BLOCK);
    put64(out + 20, nsub);
    std::vector<uint64_t> sizes(nsub, 0);
    size_t w = 0;
    if (n) {
        int rc = frames_pipelined(ctx, src, n, level, out + head, cap - head, sizes.data(), &w);
        if (rc) return rc;
    }
    for (uint64_t s = 0; s < nsub; s++) put64(out + 28 + 8 * s, sizes[s]);
    *out_len = head + w;
    return GC_OK;
}

int gc_lzma2_encode_dev(gc_ctx* ctx, const uint8_t* d_input, const uint32_t* d_packed, uint8_t* d_output,
                        uint32_t* d_output_sizes, uint64_t total_input_size) {
    if (!ctx || !d_input || !d_packed || !d_output || !d_output_sizes) return fail(GC_EINVAL, "null argument");
    const uint64_t nsub = div_up(total_input_size, GC_SUB_BLOCK);
    if (nsub == 0) return GC_OK;
    if (nsub > 0xFFFFFFFFull) return fail(GC_EINVAL, "input too large");
    std::vector<gc_item> items;
    make_items(items, total_input_size, GC_SUB_BLOCK, (uint32_t)nsub, 0, 0, false);
    int rc = upload(ctx, ctx->items, items.data(), items.size() * sizeof(gc_item));
    if (!rc) rc = ctx->out_b.ensure(nsub * 8);  // the items' costs and their order
    if (rc) return rc;
    return launch_lzma2_encode_ordered(ctx, ctx->stream, d_input, (const gc_item*)ctx->items.p, (uint32_t)nsub,
                                       d_packed, d_output, d_output_sizes, (uint32_t*)ctx->out_b.p);
}

int gc_lzma2_streams_gpu(gc_ctx* ctx, const uint8_t* chunk, size_t n, uint8_t* out, size_t cap, uint64_t* sizes,
                         size_t* out_len) {
    if (!ctx || (!chunk && n) || !out || !sizes || !out_len) return fail(GC_EINVAL, "null argument");
    *out_len = 0;
    if (n == 0) return GC_OK;
    const uint64_t nsub = div_up(n, GC_SUB_BLOCK);
    if (nsub > 0xFFFFFFFFull) return fail(GC_EINVAL, "input too large");
    std::lock_guard<std::mutex> g(ctx->mu);
    (void)hipSetDevice(ctx->device);
    return lz_two_halves(ctx, chunk, n, out, cap, sizes, out_len);
}

int gc_nvlz_compress_gpu(gc_ctx* ctx, const uint8_t* src, size_t n, uint32_t dict_mb, uint8_t* out, size_t cap,
                         size_t* out_len) {
    if (!ctx || (!src && n) || !out || !out_len) return fail(GC_EINVAL, "null argument");
    if (cap < gc_container_bound(n, GC_ALGO_LZMA2)) return fail(GC_ESIZE, "output capacity below gc_container_bound");
    const uint64_t nsub = div_up(n, GC_SUB_BLOCK);
    const size_t head = 28 + 8 * nsub;
    memcpy(out, "NVLZ", 4);  // pipeline_lzma2.rs:344-351
    auto put64 = [](uint8_t* p, uint64_t v) {
        for (int i = 0; i < 8; i++) p[i] = (uint8_t)(v >> (8 * i));
    };
    put64(out + 4, n);
    put64(out + 12, GC_SUB_BLOCK);
    put64(out + 20, nsub);
    // Matches never cross a 64 KiB sub-block, so the dictionary-size chunks of
    // pipeline_lzma2.rs:40 change no byte of the output; the whole input goes
    // through in pieces of up to 2 GiB instead, each as gc_lzma2_streams_gpu
    // codes a chunk (two halves on two streams).
    (void)dict_mb;
    const size_t piece = (size_t)2 << 30;
    std::vector<uint64_t> sizes(nsub, 0);
    size_t o = head;
    std::lock_guard<std::mutex> g(ctx->mu);
    (void)hipSetDevice(ctx->device);
    for (size_t c0 = 0; c0 < n; c0 += piece) {
        const size_t len = n - c0 < piece ? n - c0 : piece;
        size_t w = 0;
        int rc = lz_two_halves(ctx, src + c0, len, out + o, cap - o, sizes.data() + c0 / GC_SUB_BLOCK, &w);
        if (rc) return rc;
        o += w;
    }
    for (uint64_t s = 0; s < nsub; s++) put64(out + 28 + 8 * s, sizes[s]);
    *out_len = o;
    return GC_OK;
}

size_t gc_nvzs_gpu_bound(size_t n) { return 28 + 8 * (size_t)div_up(n, GC_SUB_BLOCK) + n + 16 * (size_t)div_up(n, GC_SUB_BLOCK); }

}  // extern "C"
