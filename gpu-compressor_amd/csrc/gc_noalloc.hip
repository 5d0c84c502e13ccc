// gc_noalloc.hip -- the allocation-free form of the kernel-level Zstd entry
// point, for callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc).
//
// gc_zstd_match_find_dev keeps its scratch (work items, per-position
// candidates and info words) in grow-only buffers that it allocates on
// demand.  Here the scratch is one static device array, `scratch`, of fixed
// size; each device holds its own copy from the time the code object is
// loaded until the process ends.  Host code never passes the array's name to
// a kernel (on the host it is only a stand-in for the symbol): its device
// address is looked up once per device with hipGetSymbolAddress when a
// context opens, and that address is what the kernels get.
//
// The work items are built on the device (gc_make_items), so a call issues
// kernel launches on the context's stream and nothing else: no allocation,
// no copy, no synchronisation.  Every call on a device uses the same array,
// so calls that overlap on one device (two streams, two threads) race; calls
// ordered on one stream do not.  Every scratch word a call reads is written
// earlier in the same call, so nothing needs clearing between calls.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/gcompress.h"
#include "gc_runtime.h"

// Layout (each part 256-byte aligned, with 256 bytes to spare):
// items (64 B per sub-block), candidates (2 B per position), info words
// (4 B per position); positions are counted at a pitch of the sub-block size
// rounded up to 8.
#define NOALLOC_ITEMS_BYTES ((uint64_t)GC_NOALLOC_MAX_SUB_BLOCKS * 64u + 256u)
#define NOALLOC_CAND_BYTES ((uint64_t)GC_NOALLOC_MAX_POSITIONS * 2u + 256u)
#define NOALLOC_INFO_BYTES ((uint64_t)GC_NOALLOC_MAX_POSITIONS * 4u + 256u)
#define NOALLOC_SCRATCH_BYTES (NOALLOC_ITEMS_BYTES + NOALLOC_CAND_BYTES + NOALLOC_INFO_BYTES)

static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];

namespace {

const int kMaxDevices = 64;
std::mutex g_mu;
uint8_t* g_scratch[kMaxDevices];  // device address of `scratch` on each device, once looked up
hipError_t g_lookup[kMaxDevices];  // the lookup's result, for the error message

uint64_t up256(uint64_t v) { return (v + 255) & ~(uint64_t)255; }

}  // namespace

namespace gc {

// With `device` current (gc_open has just made it so).
int noalloc_attach(int device) {
    if (device < 0 || device >= kMaxDevices) return GC_EINVAL;
    std::lock_guard<std::mutex> g(g_mu);
    if (g_scratch[device]) return GC_OK;
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_lookup[device] = e;
    if (e != hipSuccess || !p) return GC_EHIP;
    g_scratch[device] = (uint8_t*)p;
    return GC_OK;
}

}  // namespace gc

using namespace gc;

extern "C" int gc_zstd_match_find_dev_noalloc(gc_ctx* ctx, const uint8_t* d_input, void* d_sequences,
                                              uint32_t seq_format, uint8_t* d_literals, uint32_t* d_seq_counts,
                                              uint32_t* d_lit_counts, uint32_t chunk_size, uint32_t num_chunks,
                                              uint64_t total_input_size, uint32_t search_depth,
                                              uint32_t max_sequences_per_chunk) {
    if (!ctx || !d_input || !d_sequences || !d_seq_counts) return fail(GC_EINVAL, "null argument");
    if (chunk_size == 0 || chunk_size > GC_SUB_BLOCK)
        return fail(GC_EINVAL, "sub-block size must be 1..65536 (positions are kept in 16 bits)");
    if ((uintptr_t)d_input & 3u) return fail(GC_EINVAL, "d_input must be 4-byte aligned");
    if (seq_format > GC_SEQ_FORMAT_REF) return fail(GC_EINVAL, "unknown seq_format");
    if ((uint64_t)num_chunks > (total_input_size + chunk_size - 1) / chunk_size)
        return fail(GC_EINVAL, "num_chunks exceeds the sub-chunks of total_input_size");
    if (num_chunks == 0) return GC_OK;
    const uint64_t pitch = ((uint64_t)chunk_size + 7) & ~(uint64_t)7;
    const uint64_t positions = (uint64_t)num_chunks * pitch;
    if (num_chunks > GC_NOALLOC_MAX_SUB_BLOCKS || positions > GC_NOALLOC_MAX_POSITIONS)
        return fail(GC_EINVAL,
                    "gc_zstd_match_find_dev_noalloc: input beyond its fixed scratch (at most 4096 sub-chunks "
                    "and 8 Mi positions, e.g. 8 MiB in 64 KiB sub-chunks); use gc_zstd_match_find_dev");
    uint8_t* base = (ctx->device >= 0 && ctx->device < kMaxDevices) ? g_scratch[ctx->device] : nullptr;
    if (!base) {
        const hipError_t e = (ctx->device >= 0 && ctx->device < kMaxDevices) ? g_lookup[ctx->device] : hipSuccess;
        return hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol,
                        "gc_zstd_match_find_dev_noalloc: the scratch array's address on this device is unknown");
    }
    gc_item* items = (gc_item*)base;
    uint16_t* cand = (uint16_t*)(base + up256(NOALLOC_ITEMS_BYTES));
    uint32_t* info = (uint32_t*)((uint8_t*)cand + up256(NOALLOC_CAND_BYTES));
    hipStream_t st = ctx->stream;
    int rc = launch_make_items(ctx, st, items, total_input_size, chunk_size, num_chunks, max_sequences_per_chunk,
                               search_depth);
    if (!rc) rc = launch_cand(ctx, st, d_input, items, num_chunks, num_chunks, cand);
    if (!rc) rc = launch_zstd_info(ctx, st, d_input, items, num_chunks, num_chunks, cand, info);
    if (!rc)
        rc = launch_zstd_parse(ctx, st, d_input, items, num_chunks, (num_chunks + 3) / 4, info, d_sequences, seq_format,
                               d_literals, d_seq_counts, d_lit_counts, max_sequences_per_chunk);
    return rc;
}
