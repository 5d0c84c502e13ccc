// kernels.hip -- gfx950 match-finding kernels.
//
// Two finders, same results as the reference's device kernels:
//   lzma2_match_find  (lzma2_match_find.cu:54-194)
//   zstd_match_find   (zstd_compress.cu:179-355)
//
// The reference runs one thread per 64 KiB sub-block over every position.
// Here the work is split by what depends on what:
//
//  * The hash candidate of a position p is a pure function of the sub-block:
//      cand(p) = max{ q < p : q + 3 < n, H(q) = H(p) }   (or "none")
//    because both reference kernels insert every position q with q + 3 < n
//    into the 14-bit table exactly once, in increasing order, before any
//    later lookup.  gc_cand_kernel computes it tile by tile with the
//    reference's own table (16384 x u16 in LDS); inside a wavefront the lanes
//    that share a hash are grouped with ballots, so only the first lane of a
//    group reads the table and only the last one writes it.  The kernel does
//    nothing but LDS work and prefetched, coalesced input loads.
//  * Everything else per position is elementwise given cand(p): the LZMA2
//    match (gc_lzma2_len_kernel) and the Zstd verify + short length
//    (gc_zstd_info_kernel).  These kernels use no LDS and run at full
//    occupancy, which hides the latency of their gathers.
//  * The Zstd parse (heads, repeat offsets, the lazy step) is serial per
//    sub-block: gc_zstd_parse_kernel walks each sub-block with one wavefront,
//    testing 128 positions per step against the current repeat offsets and
//    jumping straight to the first one that matches.
//
// Work arrives as a list of sub-blocks (gc_item), so one launch serves one
// chunk (plain form) or every sub-block of many inputs (batched form).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gc_device.h"

// Two code objects are built from this file: the parse kernels
// (GC_BUILD_MAIN=0), whose walk is compiled without structurizing uniform
// branches (fewer register copies on the loop's paths), and the others
// (GC_BUILD_PARSE=0).  A plain compile has both.
#ifndef GC_BUILD_MAIN
#define GC_BUILD_MAIN 1
#endif
#ifndef GC_BUILD_PARSE
#define GC_BUILD_PARSE 1
#endif

#define GC_HASH_BITS 14
#define GC_TABLE (1u << GC_HASH_BITS)
#define GC_NONE 0xFFFFu
#ifndef GC_DUP_BITS
#define GC_DUP_BITS 9
#endif
#define GC_DUP_SIZE (1u << GC_DUP_BITS)
#define GC_LZMA_MAX 273u
#define GC_ZSTD_SHORT 64u  // Zstd lengths >= this are re-measured by the parse

#define WAVE 64
#define CAND_AWAVES (GC_CAND_THREADS / WAVE - 1)  // candidate kernel: analysis wavefronts (+ 1 table)
#define CAND_THREADS GC_CAND_THREADS
#define CAND_WIDE_AWAVES (GC_CAND_WIDE_THREADS / WAVE - 1)  // the wide form (launches of few sub-blocks)
#define INFO_THREADS 1024                          // per-position kernels: one workgroup per sub-block
#define STAGE_WORDS ((GC_SUB_BLOCK + 64) / 4)      // a sub-block staged in LDS, with 64 bytes to spare
#define PARSE_WAVES 4
#ifndef PARSE_UNIFORM_ITEM
#define PARSE_UNIFORM_ITEM 0
#endif

// ---------------------------------------------------------------------------
// Helpers
// ---------------------------------------------------------------------------

// A sub-block's bytes as seen by one wavefront: a uniform (scalar) base
// rounded down to 4 bytes and the misalignment, so that every read is a
// 32-bit offset from a scalar base (no 64-bit address arithmetic per lane).
// `avail` is the number of readable bytes from the sub-block start.
// Input bytes are read through global-address-space pointers: a load through
// a generic pointer is a FLAT load, which also counts against the LDS
// counter, so the first LDS wait after it (a table or ring read) would wait
// for it too -- a prefetch from HBM would be waited for at once.
#define GC_GLOBAL __attribute__((address_space(1)))
typedef const uint8_t GC_GLOBAL* gbytes;
typedef const uint32_t GC_GLOBAL* gwords;
typedef uint32_t __attribute__((aligned(1))) u32_unaligned;

struct SrcView {
    gbytes base;  // sub-block start - mis, 4-byte aligned
    uint32_t mis;         // sub-block start & 3
    uint64_t avail;
};

// The byte at arena + off, as an address computed in integers: an item's
// offset is relative to its launch's arena or, with a null arena, the
// absolute address of an input that lives in another allocation.
__device__ __forceinline__ const uint8_t* at(const uint8_t* arena, uint64_t off) {
    return (const uint8_t*)((uintptr_t)arena + off);
}

__device__ __forceinline__ SrcView src_view(const uint8_t* arena, uint64_t src_off, uint64_t avail) {
    SrcView v;
    const uintptr_t a = (uintptr_t)arena + src_off;
    v.mis = (uint32_t)(a & 3);
    v.base = (gbytes)(a - v.mis);
    v.avail = avail;
    return v;
}

// The aligned dword at byte offset a (a multiple of 4) of the view's base.
__device__ __forceinline__ uint32_t dword_at(const SrcView& v, uint32_t a) {
    return *(gwords)(v.base + a);
}

// 4 bytes at sub-block offset `off`, little endian, built from the two
// aligned dwords around them.  SAFE callers guarantee avail >= n + 32 for a
// sub-block of n bytes, which covers every read below (at most 24 bytes from
// an offset below n); otherwise a read that would pass the end is built byte
// by byte from the bytes that exist.
__device__ __attribute__((noinline)) uint32_t ld32_tail(gbytes p, uint32_t off, uint64_t avail) {
    uint32_t v = 0;
    for (uint32_t i = 0; i < 4; i++)
        if ((uint64_t)off + i < avail) v |= (uint32_t)p[i] << (8 * i);
    return v;
}

template <bool SAFE>
__device__ __forceinline__ uint32_t ld32(const SrcView& v, uint32_t off) {
    if (SAFE || (uint64_t)off + 8 <= v.avail) {
        const uint32_t t = off + v.mis;
        gwords q = (gwords)(v.base + (t & ~3u));  // q[0], q[1]: one 8-byte load
        return __builtin_amdgcn_alignbyte(q[1], q[0], t);
    }
    return ld32_tail(v.base + v.mis + off, off, v.avail);  // the last bytes of the arena, byte by byte
}

// ld32 as one unaligned dword load (gfx950 under ROCm runs with unaligned
// memory access enabled: a dword load at any byte address returns the 4
// bytes there), for a prefetch whose word is used as loaded.  The bytes read
// lie inside those of ld32's aligned pair, so the same SAFE bound covers
// them.
template <bool SAFE>
__device__ __forceinline__ uint32_t ld32u(const SrcView& v, uint32_t off) {
    if (SAFE || (uint64_t)off + 8 <= v.avail) {
        uint32_t w;
        w = *(const GC_GLOBAL u32_unaligned*)(v.base + (off + v.mis));  // scalar base + 32-bit lane offset
        return w;
    }
    return ld32_tail(v.base + v.mis + off, off, v.avail);
}

// ld32 in two halves, for a prefetch: the load now (the two aligned dwords
// and the shift), the byte alignment when the word is used.  Computing it
// right away would wait for the load there.
struct Raw32 {
    uint32_t lo, hi, sh;
    __device__ __forceinline__ uint32_t word() const { return __builtin_amdgcn_alignbyte(hi, lo, sh); }
};
template <bool SAFE>
__device__ __forceinline__ Raw32 ld32_raw(const SrcView& v, uint32_t off) {
    if constexpr (SAFE) {
        const uint32_t t = off + v.mis;
        gwords q = (gwords)(v.base + (t & ~3u));
        return Raw32{q[0], q[1], t};
    } else {
        return Raw32{ld32<false>(v, off), 0u, 0u};
    }
}

// 16 bytes at sub-block offset `off` as four little-endian words, from the
// five aligned dwords around them.
template <bool SAFE>
__device__ __forceinline__ void ld128(const SrcView& v, uint32_t off, uint32_t o[4]) {
    if (SAFE || (uint64_t)off + 24 <= v.avail) {
        const uint32_t t = off + v.mis;
        gwords q = (gwords)(v.base + (t & ~3u));
        const uint32_t sh = t;
        const uint32_t d0 = q[0], d1 = q[1], d2 = q[2], d3 = q[3], d4 = q[4];
        o[0] = __builtin_amdgcn_alignbyte(d1, d0, sh);
        o[1] = __builtin_amdgcn_alignbyte(d2, d1, sh);
        o[2] = __builtin_amdgcn_alignbyte(d3, d2, sh);
        o[3] = __builtin_amdgcn_alignbyte(d4, d3, sh);
        return;
    }
    for (int k = 0; k < 4; k++) o[k] = ld32<false>(v, off + 4 * k);
}

// (v_alignbyte_b32 shifts by 8 x the low two bits of its third operand, so
// the byte offsets above and below are passed to it unmasked.)

// Per-item reads may skip bounds checks when the arena extends 32 bytes past
// the sub-block (every sub-block but the last of an arena).
__device__ __forceinline__ bool item_safe(uint64_t avail, uint32_t n) { return avail >= (uint64_t)n + 32; }

// Item i, every field made wavefront-uniform (held in scalar registers).
__device__ __forceinline__ gc_item load_item(const gc_item* __restrict__ items, uint32_t i) {
    gc_item it = items[i];
    uint32_t* w = (uint32_t*)&it;
#pragma unroll
    for (int k = 0; k < (int)(sizeof(gc_item) / 4); k++) w[k] = (uint32_t)__builtin_amdgcn_readfirstlane((int)w[k]);
    return it;
}

__device__ __forceinline__ uint32_t hash4(uint32_t v) { return (v * 2654435761u) >> (32 - GC_HASH_BITS); }

__device__ __forceinline__ uint32_t lane_id() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

__device__ __forceinline__ uint64_t ballot(bool b) { return __builtin_amdgcn_ballot_w64(b); }

// A wavefront-uniform value kept in a vector register, and back.
__device__ __forceinline__ uint32_t in_vgpr(uint32_t s) {
    uint32_t v;
    __asm__("v_mov_b32 %0, %1" : "=v"(v) : "s"(s));
    return v;
}
__device__ __forceinline__ uint32_t uniform(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }

// max of wavefront-uniform values on the scalar unit (the compiler otherwise
// takes a max of three to the vector unit, with a move per scalar operand)
__device__ __forceinline__ uint32_t smax(uint32_t a, uint32_t b) {
    uint32_t r;
    __asm__("s_max_u32 %0, %1, %2" : "=s"(r) : "s"(a), "s"(b));
    return r;
}

__device__ __forceinline__ uint32_t bcast(uint32_t v, uint32_t lane) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)lane);
}

// Lane i gets lane i - 1's v, lane 0 gets first (DPP wave_shr:1).
__device__ __forceinline__ uint32_t wave_shr1(uint32_t v, uint32_t first) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)first, (int)v, 0x138, 0xF, 0xF, false);
}

// Sum over the wavefront, uniform result: pairs, quads (quad_perm), 8 and
// 16 lanes (row half mirror, row mirror), then rows 1 and 3 take rows 0
// and 2 (row_bcast:15) and row 3 the first half (row_bcast:31); lane 63 holds
// the total.  DPP needs no lane-address registers.
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false);   // quad_perm 1,0,3,2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false);   // quad_perm 2,3,0,1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xF, 0xF, false);  // row_half_mirror
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x140, 0xF, 0xF, false);  // row_mirror
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false);  // row_bcast:15
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false);  // row_bcast:31
    return bcast(v, WAVE - 1);
}

// Inclusive prefix sum over the wavefront's lanes (lane i: v of lanes 0..i):
// shifts within rows of 16 (row_shr 1, 2, 4, 8), then rows 1 and 3 add the
// row before (row_bcast:15) and rows 2 and 3 the first half (row_bcast:31).
__device__ __forceinline__ uint32_t wave_scan(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, false);  // row_shr:1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, false);  // row_shr:2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, false);  // row_shr:4
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, false);  // row_shr:8
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false);  // row_bcast:15
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false);  // row_bcast:31
    return v;
}

// lane `lane` of a set to sa and of b to sb (all uniform): one instruction
// per register, where a compare and a select would be two or three.  (Two
// scalar operands exceed the constant bus: the lane index is passed in M0.)
__device__ __forceinline__ void write_lane2(uint32_t& a, uint32_t sa, uint32_t& b, uint32_t sb, uint32_t lane) {
    __asm__("v_writelane_b32 %0, %2, %4\n\tv_writelane_b32 %1, %3, %4"
            : "+v"(a), "+v"(b)
            : "s"(sa), "s"(sb), "{m0}"(lane));
}

// Barrier for LDS traffic only: outstanding global loads (prefetches) stay
// in flight across it.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// ---------------------------------------------------------------------------
// Item sources.  Static: workgroup b takes items b, b + G, b + 2G, ...
// Queue (persistent kernels): each workgroup (or, in the parse, each
// wavefront) takes the next item from the device work queue until it runs
// dry.  A taker can take at most n items, so its loop is bounded by n + 1
// fetches; passing that bound sets error bit 0 and ends the loop.  No taker
// ever waits for another: every loop ends whatever the other workgroups do.
// ---------------------------------------------------------------------------
// The taker that leaves last (of `takers`) finds every other one gone: it
// checks the count and sets the counters back to 0 for the next launch.
__device__ __forceinline__ void queue_leave(gc_queue* q, uint32_t ph, uint32_t n, uint32_t mine, uint32_t takers) {
    atomicAdd(&q->taken[ph], mine);
    __threadfence();
    if (atomicAdd(&q->done[ph], 1u) == takers - 1) {
        __threadfence();
        if (atomicAdd(&q->taken[ph], 0u) != n) atomicOr(&q->err, 2u);
        atomicExch(&q->next[ph], 0u);
        atomicExch(&q->taken[ph], 0u);
        atomicExch(&q->done[ph], 0u);
    }
}

// Workgroup-level: body(i) for every item i the workgroup takes (i counts
// from the phase's first item, q->base[ph]); all threads call this, body may
// use barriers.
__device__ __forceinline__ uint32_t queue_base(const gc_queue* q, uint32_t ph) {
    return uniform(__atomic_load_n(&q->base[ph], __ATOMIC_RELAXED));
}

template <class F>
__device__ __forceinline__ void wg_queue_items(gc_queue* q, uint32_t ph, F&& body) {
    __shared__ uint32_t s_take;
    const uint32_t n = uniform(__atomic_load_n(&q->n[ph], __ATOMIC_RELAXED));
    uint32_t mine = 0;
    for (uint32_t iter = 0;; iter++) {
        if (threadIdx.x == 0) {
            uint32_t i = atomicAdd(&q->next[ph], 1u);
            if (iter > n) {
                atomicOr(&q->err, 1u);
                i = n;
            }
            s_take = i;
        }
        __syncthreads();
        const uint32_t i = uniform(s_take);
        __syncthreads();
        if (i >= n) break;
        body(i);
        mine++;
    }
    if (threadIdx.x == 0) queue_leave(q, ph, n, mine, gridDim.x);
}

// A sub-block staged in LDS (byte 0 = sub-block start, 4-byte aligned, at
// least 32 readable bytes past the sub-block).
struct LdsBytes {
    const uint32_t* w;
};

__device__ __forceinline__ uint32_t ld32(const LdsBytes& v, uint32_t off) {
    const uint32_t* q = v.w + (off >> 2);
    return __builtin_amdgcn_alignbyte(q[1], q[0], off);
}

__device__ __forceinline__ void ld128(const LdsBytes& v, uint32_t off, uint32_t o[4]) {
    const uint32_t* q = v.w + (off >> 2);
    const uint32_t sh = off;
    const uint32_t d0 = q[0], d1 = q[1], d2 = q[2], d3 = q[3], d4 = q[4];
    o[0] = __builtin_amdgcn_alignbyte(d1, d0, sh);
    o[1] = __builtin_amdgcn_alignbyte(d2, d1, sh);
    o[2] = __builtin_amdgcn_alignbyte(d3, d2, sh);
    o[3] = __builtin_amdgcn_alignbyte(d4, d3, sh);
}

// Copies bytes [0, round_up(n + 32, 16)) of a sub-block into LDS (bytes past
// the arena read as 0); the whole workgroup takes part.
// With room past the end (the usual case) every thread issues all its loads
// (clamped, unconditional) before the first LDS write, so their latencies
// overlap; otherwise the reads near the arena's end go byte by byte.
template <uint32_t THREADS>
__device__ __forceinline__ void stage_sub_block(uint32_t* lds, const SrcView& src, uint32_t n) {
    const uint32_t end = (n + 32 + 15) & ~15u;
    constexpr uint32_t PASSES = (GC_SUB_BLOCK + 32 + 15) / (16 * THREADS) + 1;
    if (src.avail >= (uint64_t)end + 8) {
        uint32_t w[PASSES][4];
#pragma unroll
        for (uint32_t k = 0; k < PASSES; k++) {
            const uint32_t o = (threadIdx.x + k * THREADS) * 16;
            ld128<true>(src, o < end ? o : 0u, w[k]);
        }
#pragma unroll
        for (uint32_t k = 0; k < PASSES; k++) {
            const uint32_t o = (threadIdx.x + k * THREADS) * 16;
            if (o < end) *(uint4*)(lds + o / 4) = make_uint4(w[k][0], w[k][1], w[k][2], w[k][3]);
        }
        return;
    }
    for (uint32_t o = threadIdx.x * 16; o < end; o += THREADS * 16) {
        uint32_t w[4];
        ld128<false>(src, o, w);
        *(uint4*)(lds + o / 4) = make_uint4(w[0], w[1], w[2], w[3]);
    }
}

// Common-prefix length of src[a..] and src[b..] from `start` (bytes before
// it known equal), at most `limit`; one lane.
template <bool SAFE>
__device__ __forceinline__ void ld128v(const SrcView& v, uint32_t off, uint32_t o[4]) {
    ld128<SAFE>(v, off, o);
}
template <bool SAFE>
__device__ __forceinline__ void ld128v(const LdsBytes& v, uint32_t off, uint32_t o[4]) {
    ld128(v, off, o);
}

template <bool SAFE, class V>
__device__ __forceinline__ uint32_t lcp_lane(const V& src, uint32_t a, uint32_t b, uint32_t start, uint32_t limit) {
    uint32_t len = start;
    while (len < limit) {  // 16 bytes per step
        uint32_t x[4], y[4];
        ld128v<SAFE>(src, a + len, x);
        ld128v<SAFE>(src, b + len, y);
        const uint32_t d0 = x[0] ^ y[0], d1 = x[1] ^ y[1], d2 = x[2] ^ y[2], d3 = x[3] ^ y[3];
        const uint32_t k = d0   ? (uint32_t)__builtin_ctz(d0) >> 3
                           : d1 ? 4 + ((uint32_t)__builtin_ctz(d1) >> 3)
                           : d2 ? 8 + ((uint32_t)__builtin_ctz(d2) >> 3)
                           : d3 ? 12 + ((uint32_t)__builtin_ctz(d3) >> 3)
                                : 16u;
        len += k;
        if (k < 16) break;
    }
    return len < limit ? len : limit;
}

// Common-prefix length, whole wavefront cooperating (256 bytes per step).
// a, b, start and limit must be uniform.
template <bool SAFE>
__device__ __forceinline__ uint32_t lcp_wave(const SrcView& src, uint32_t a, uint32_t b, uint32_t limit,
                                             uint32_t start) {
    const uint32_t lane = lane_id();
    uint32_t len = start;
    uint64_t mm = 0;
    uint32_t x = 0;
    while (len < limit) {  // one exit: the structurizer adds no flags
        const uint32_t o = len + 4 * lane;
        x = 0;
        if (o < limit) x = ld32<SAFE>(src, a + o) ^ ld32<SAFE>(src, b + o);
        mm = __builtin_amdgcn_uicmp(x, 0u, 33);  // the lanes whose words differ
        len = mm ? len : len + 4 * WAVE;
        if (mm) break;
    }
    if (mm) {
        const uint32_t k = (uint32_t)__builtin_ctzll(mm);
        len += 4 * k + ((uint32_t)__builtin_ctz(bcast(x, k)) >> 3);
    }
    return len < limit ? len : limit;
}

#if GC_BUILD_MAIN
// ---------------------------------------------------------------------------
// Candidate kernel: cand16[it.cpos + p] = cand(p) (GC_NONE when there is none
// or p + 3 >= n) for every position of every item.
//
// One workgroup per sub-block, tiles of 64 x (analysis wavefronts) positions,
// two roles:
//  * CAND_AWAVES analysis wavefronts, 64 positions each: hash, and group the
//    lanes of the wavefront that share a hash, so that each position knows
//    its predecessor inside its 64 (if any) and whether a later lane of the
//    64 shares its hash.  Packed words go to LDS.
//  * One table wavefront applies a tile's CAND_AWAVES steps of 64 to the
//    reference's table (16384 x u16 in LDS) in order: the first lane of a
//    group reads the slot, the last one writes it.  The order of table
//    reads and writes is this wavefront's program order, so the steps need
//    no barrier between them.
// The analysis of tile t + 1 runs while the table wavefront works on tile t:
// one barrier per tile.
// ---------------------------------------------------------------------------
// AW analysis wavefronts: tiles of AW x 64 positions.
template <uint32_t AW>
struct CandLds {
    uint16_t table[GC_TABLE];          // the reference's 14-bit table
    uint32_t ana[2][AW * WAVE];        // analysis words, double buffered
    uint8_t dup[AW][GC_DUP_SIZE + 4];  // write/read-back duplicate test (+ a junk slot)
};

// Analysis word of a position, decoded as far as the table wavefront's
// steps need: the byte offset of its hash's table entry (2 x hash, bits
// 0-14), the candidate unless it comes from the table (bits 15-30: the
// earlier lane's position when one shares the hash, GC_NONE when the
// position is not valid, else ANA_FROM_TABLE, which no position reaches: a
// candidate is at most n - 5), and whether the table entry is written (the
// sign bit, one compare: the position is valid and no later lane of its tile
// step shares its hash).
#define ANA_WRITE (1u << 31)
#define ANA_FROM_TABLE 0xFFFEu

// ALL: every lane's position is valid (p + 3 < n), as in all but a
// sub-block's last tiles: the validity masks and selects drop out.
template <bool ALL, uint32_t AW>
__device__ __forceinline__ uint32_t cand_analyse(CandLds<AW>& s, uint32_t wave, uint32_t word, uint32_t p,
                                                 uint32_t n) {
    const uint32_t lane = lane_id();
    const uint32_t h = hash4(word);
    const bool valid = ALL || p + 3 < n;

    // Duplicate test: every valid lane writes its id under a slice of its
    // hash (one slot array per wavefront) and reads it back; one lane of each
    // slot wins.  Lanes that share a whole hash share the slot, so every
    // group of two or more equal hashes has a lane that lost.  A loser reads
    // the winner's hash across the wavefront: if it is its own, the loser
    // seeds its group.  Lanes that lost to another hash repeat the test among
    // themselves under another slice of their hash: a group none of whose
    // lanes won the first slot lies wholly among them and shares the second
    // slot, so one of its lanes loses there and seeds it.  Every group of two
    // or more gets a seed whichever lane the hardware lets win a slot, and a
    // lane alone with its hash seeds nothing unless it loses both tests (the
    // seeds are what the exact groups below iterate over).
    // Other lanes write the same bytes between this lane's store and load,
    // which the compiler cannot see: the empty asm with a memory clobber
    // keeps it from forwarding the store.  LDS operations of one wavefront
    // execute in program order, so the load sees every lane's store.
    // Lanes that take no part in a test write and read the junk slot instead
    // of branching around the LDS access; the lane masks below leave them out.
    // The lane masks come straight from compares (uicmp) and are combined on
    // the scalar unit; a group's mask selects its lanes as it is (inverse
    // ballot): a ballot of a conjunction, or a per-lane test of a mask bit,
    // would cost further vector instructions.
    uint8_t* dup = s.dup[wave];
    const uint64_t vmask = ALL ? ~0ull : __builtin_amdgcn_uicmp(p + 3, n, 36);  // unsigned <
    const uint32_t slot = valid ? (h & (GC_DUP_SIZE - 1)) : GC_DUP_SIZE;
    dup[slot] = (uint8_t)lane;
    __asm__ volatile("" ::: "memory");
    const uint32_t back = dup[slot];
    const uint32_t hw = (uint32_t)__builtin_amdgcn_ds_bpermute((int)(back << 2), (int)h);
    const uint64_t lost = __builtin_amdgcn_uicmp(back, lane, 33) & vmask;  // !=
    const uint64_t other = lost & __builtin_amdgcn_uicmp(hw, h, 33);       // lost to another hash
    const uint32_t slot2 =
        __builtin_amdgcn_inverse_ballot_w64(other) ? (h >> (GC_HASH_BITS - GC_DUP_BITS)) & (GC_DUP_SIZE - 1) : GC_DUP_SIZE;
    __asm__ volatile("" ::: "memory");
    dup[slot2] = (uint8_t)lane;
    __asm__ volatile("" ::: "memory");
    const uint32_t back2 = dup[slot2];

    // Exact groups (one step per distinct hash among the seeds).
    const uint64_t lane_bit = 1ull << lane;
    uint64_t group = lane_bit;
    uint64_t rest = (lost & ~other) | (__builtin_amdgcn_uicmp(back2, lane, 33) & other);
    while (rest) {
        const uint32_t lead = (uint32_t)__builtin_ctzll(rest);
        const uint32_t hl = bcast(h, lead);
        const uint64_t g = __builtin_amdgcn_uicmp(h, hl, 32) & vmask;  // ==
        if (__builtin_amdgcn_inverse_ballot_w64(g)) group = g;
        rest &= ~g;
    }
    const uint64_t below = group & (lane_bit - 1);
    const uint64_t above = group & ~(lane_bit | (lane_bit - 1));
    // (the earlier lane's position, used when there is one; | 1 keeps clz defined)
    const uint32_t pred = p - lane + 63u - (uint32_t)__builtin_clzll(below | 1u);
    const uint32_t c = !valid ? GC_NONE : below ? pred : ANA_FROM_TABLE;
    return (valid && !above ? ANA_WRITE : 0u) | (c << 15) | (h << 1);
}

// The table wavefront's work on one tile at t0 (analysis words in `ana`).
template <uint32_t AW>
__device__ __forceinline__ void cand_table(CandLds<AW>& s, const uint32_t* ana, uint32_t t0, uint32_t n,
                                           uint16_t* out) {
    const uint32_t lane = lane_id();
    uint32_t a[AW], v[AW];
#pragma unroll
    for (int j = 0; j < (int)AW; j++) a[j] = ana[j * WAVE + lane];
#pragma unroll
    for (int j = 0; j < (int)AW; j++) {
        // every lane reads its entry (lanes whose candidate is decided already
        // ignore it); the last lane of each hash writes its position
        const uint16_t* e = (const uint16_t*)((const uint8_t*)s.table + (a[j] & (2 * GC_TABLE - 2)));
        v[j] = *e;
        if ((int32_t)a[j] < 0) *(uint16_t*)e = (uint16_t)(t0 + j * WAVE + lane);  // ANA_WRITE
    }
    // (one base address and immediate offsets; the bound on the scalar unit,
    // tested per lane only in a sub-block's last tile)
    uint16_t* o = out + t0 + lane;
    const int32_t left = (int32_t)(n - t0);
    if (left >= (int32_t)(AW * WAVE)) {
#pragma unroll
        for (int j = 0; j < (int)AW; j++) {
            const uint32_t c = (a[j] >> 15) & 0xFFFFu;
            o[j * WAVE] = (uint16_t)(c == ANA_FROM_TABLE ? v[j] : c);
        }
    } else {
#pragma unroll
        for (int j = 0; j < (int)AW; j++) {
            const uint32_t c = (a[j] >> 15) & 0xFFFFu;
            if ((int32_t)lane < left - j * (int32_t)WAVE) o[j * WAVE] = (uint16_t)(c == ANA_FROM_TABLE ? v[j] : c);
        }
    }
}

template <bool SAFE, uint32_t AW>
__device__ __forceinline__ void cand_item(CandLds<AW>& s, const gc_item& it, const uint8_t* arena, uint16_t* cand16) {
    constexpr uint32_t TILE = AW * WAVE, THREADS = TILE + WAVE;
    const uint32_t tid = threadIdx.x;
    const uint32_t wave = (uint32_t)__builtin_amdgcn_readfirstlane((int)(tid / WAVE));
    const uint32_t lane = lane_id();
    const SrcView src = src_view(arena, it.src_off, it.avail);
    const uint32_t n = it.n;
    uint16_t* out = cand16 + it.cpos;
    uint32_t* t = (uint32_t*)s.table;
    for (uint32_t k = tid; k < GC_TABLE / 2; k += THREADS) t[k] = 0xFFFFFFFFu;
    // (lane 0's position decides for the wavefront whether all 64 are valid)
    auto analyse = [&](uint32_t word, uint32_t p) {
        return uniform(p) + (WAVE - 1) + 3 < n ? cand_analyse<true, AW>(s, wave, word, p, n)
                                               : cand_analyse<false, AW>(s, wave, word, p, n);
    };
    // analysis of tile 0; words two tiles ahead.  The loads are unconditional
    // (positions past the end read position 0, masked when used): a load
    // under a per-lane condition is waited for where the branches join.
    Raw32 next{0u, 0u, 0u};
    if (wave < AW) {
        const uint32_t p = wave * WAVE + lane;
        const uint32_t word = ld32<SAFE>(src, p + 3 < n ? p : 0u);
        const uint32_t q = TILE + p;
        next = ld32_raw<SAFE>(src, q + 3 < n ? q : 0u);
        s.ana[0][p] = analyse(word, p);
    }
    lds_barrier();
    uint32_t buf = 0;
    for (uint32_t t0 = 0; t0 < n; t0 += TILE) {
        if (wave < AW) {
            const uint32_t t1 = t0 + TILE;
            if (t1 < n) {
                const uint32_t p = t1 + wave * WAVE + lane;
                const uint32_t word = next.word();
                const uint32_t q = p + TILE;
                next = ld32_raw<SAFE>(src, q + 3 < n ? q : 0u);
                s.ana[buf ^ 1][wave * WAVE + lane] = analyse(word, p);
            }
        } else {
            cand_table(s, s.ana[buf], t0, n, out);
        }
        lds_barrier();
        buf ^= 1;
    }
}

template <uint32_t AW>
__device__ __forceinline__ void cand_items(const uint8_t* arena, const gc_item* __restrict__ items, uint32_t n_items,
                                           uint16_t* cand16) {
    __shared__ CandLds<AW> s;
    for (uint32_t i = blockIdx.x; i < n_items; i += gridDim.x) {
        const gc_item it = load_item(items, i);
        if (item_safe(it.avail, it.n))
            cand_item<true, AW>(s, it, arena, cand16);
        else
            cand_item<false, AW>(s, it, arena, cand16);
        lds_barrier();  // the table is reset for the next item
    }
}

extern "C" __global__ void __launch_bounds__(CAND_THREADS)
gc_cand_kernel(const uint8_t* arena, const gc_item* __restrict__ items, uint32_t n_items, uint16_t* cand16) {
    cand_items<CAND_AWAVES>(arena, items, n_items, cand16);
}

// The same with twice the analysis wavefronts per workgroup: half the tiles,
// so half the serial steps per sub-block.  For launches of few sub-blocks,
// where each workgroup's latency sets the time; with many sub-blocks the
// narrower form keeps more workgroups resident per CU and is faster.
extern "C" __global__ void __launch_bounds__(GC_CAND_WIDE_THREADS)
gc_cand_wide_kernel(const uint8_t* arena, const gc_item* __restrict__ items, uint32_t n_items, uint16_t* cand16) {
    cand_items<CAND_WIDE_AWAVES>(arena, items, n_items, cand16);
}

// Persistent form: items from the device work queue (phase GC_Q_CAND).  The
// served form launches the same code under a name of its own (the `_sq`
// kernels), so that a kernel trace tells its ticks from batched passes.
__device__ __forceinline__ void cand_queue(const uint8_t* arena, const gc_item* __restrict__ items, gc_queue* q,
                                           uint16_t* cand16) {
    __shared__ CandLds<CAND_AWAVES> s;
    items += queue_base(q, GC_Q_CAND);
    wg_queue_items(q, GC_Q_CAND, [&](uint32_t i) {
        const gc_item it = load_item(items, i);
        if (item_safe(it.avail, it.n))
            cand_item<true, CAND_AWAVES>(s, it, arena, cand16);
        else
            cand_item<false, CAND_AWAVES>(s, it, arena, cand16);
        lds_barrier();
    });
}

extern "C" __global__ void __launch_bounds__(CAND_THREADS)
gc_cand_q_kernel(const uint8_t* arena, const gc_item* __restrict__ items, gc_queue* q, uint16_t* cand16) {
    cand_queue(arena, items, q, cand16);
}

extern "C" __global__ void __launch_bounds__(CAND_THREADS)
gc_cand_sq_kernel(const uint8_t* arena, const gc_item* __restrict__ items, gc_queue* q, uint16_t* cand16) {
    cand_queue(arena, items, q, cand16);
}

// ---------------------------------------------------------------------------
// Per-position passes.  Each thread takes a strip of STRIP consecutive
// positions and walks it backwards: where position p+1 matches along the same
// diagonal (its target is p's target + 1), the common length at p is one more
// than at p+1, capped (the caps of p and p+1 differ by at most one byte), so
// a long match costs one direct measurement per strip instead of one per
// position.
// ---------------------------------------------------------------------------
#define STRIP 8

// The 4 bytes at p0 + j from the strip's dwords.
__device__ __forceinline__ uint32_t strip_word(const uint32_t d[4], uint32_t j) {
    const uint32_t k = j >> 2;
    const uint32_t hi = k + 1 < 4 ? d[k + 1] : 0;
    return __builtin_amdgcn_alignbyte(hi, d[k], j & 3);
}

// ---------------------------------------------------------------------------
// LZMA2 (lzma2_match_find.cu:96-185) per position, given cand(p):
//   p + 1 < n, and either the hash candidate shares >= 2 bytes (distance
//   p - cand, length = common run capped at min(n - p, 273)), or no such
//   match and src[p-1] == src[p] == src[p+1] (distance 1, run capped alike).
// Output: packed[it.opos + p] = distance << 16 | length (0 = none) and/or the
// reference layout (slot 0 of LzmaMatch[max_matches] and counts).
// ---------------------------------------------------------------------------
// One workgroup of INFO_THREADS per sub-block, staged in LDS first (as for
// the Zstd info words below).  LENGTHS = false: only whether the sub-block
// has any match (the consumer's triage, compress_lzma2_custom.rs:577), no
// lengths measured and nothing per position written.  Returns this thread's
// "some position of mine has a match".
template <bool LENGTHS, bool REF>
__device__ __forceinline__ bool lzma2_item(const gc_item& it, const LdsBytes& src, const uint16_t* cand16,
                                           uint32_t* packed, gc_lzma_match* ref_m, uint32_t* ref_c,
                                           uint32_t max_matches) {
    const uint32_t n = it.n;
    const uint32_t depth = it.depth;
    bool any = false;
    // a whole strip's candidates in one 16-byte load and its packed words in
    // two 16-byte stores where the item's slots are 16-byte aligned (the
    // candidate slots are; an output at a caller's offset may not be)
    const bool wide = ((uintptr_t)(cand16 + it.cpos) & 15) == 0 &&
                      (!LENGTHS || REF || ((uintptr_t)(packed + it.opos) & 15) == 0);
    // INNER: every position of the wavefront's strips lies in a full, wide
    // strip with p > 0 and p + GC_LZMA_MAX <= n (the end tests and the length
    // cap drop out); chosen per step for the wavefront (base: lane 0's strip)
    //
    // Lengths: along a diagonal (the target of p + 1 is the target of p, plus
    // 1) the length at p is one more than at p + 1, capped, so only the last
    // position of a run of such positions (a run end) is measured.  A run can
    // cross strips: each lane finds, from a ballot of the lanes whose whole
    // strip continues, the first lane after it whose run ends inside its strip
    // and takes that lane's length at its first position (ds_bpermute), plus 8
    // per whole strip between.  Each wavefront walks its own region from the
    // last step, so a run that crosses steps continues from the step before
    // (carry_t, carry_len: the target and length at the first position of the
    // step above): a long match costs one measurement per region, not one per
    // strip.
    auto strip = [&](uint32_t p0, auto inner, uint32_t& carry_t, uint32_t& carry_len) {
        constexpr bool INNER = decltype(inner)::value;
        uint32_t d[4];
        d[0] = src.w[p0 / 4];
        d[1] = src.w[p0 / 4 + 1];
        d[2] = src.w[p0 / 4 + 2];
        d[3] = src.w[p0 / 4 + 3];
        const uint32_t prev0 = (INNER || p0 > 0) ? ld32(src, p0 - 1) : 0;
        const bool full = INNER || (wide && p0 + STRIP <= n);
        uint32_t tgt[STRIP], vc[STRIP];
        if (full) {
            const uint4 cw = *(const uint4*)(cand16 + it.cpos + p0);
            const uint32_t cv[4] = {cw.x, cw.y, cw.z, cw.w};
#pragma unroll
            for (uint32_t j = 0; j < STRIP; j++)
                tgt[j] = ((INNER || p0 + j + 3 < n) && depth > 0) ? (cv[j >> 1] >> (16 * (j & 1))) & 0xFFFFu : GC_NONE;
        } else {
#pragma unroll
            for (uint32_t j = 0; j < STRIP; j++) {
                const uint32_t p = p0 + j;
                tgt[j] = (p + 3 < n && depth > 0) ? cand16[it.cpos + p] : GC_NONE;
            }
        }
#pragma unroll
        for (uint32_t j = 0; j < STRIP; j++) vc[j] = tgt[j] != GC_NONE ? ld32(src, tgt[j]) : 0;
        // the match target of each position (GC_NONE: no match)
        uint32_t t[STRIP];
#pragma unroll
        for (int j = 0; j < (int)STRIP; j++) {
            const uint32_t p = p0 + j;
            t[j] = GC_NONE;
            if (INNER || p + 1 < n) {
                const uint32_t cur = strip_word(d, j);
                if (tgt[j] != GC_NONE && ((vc[j] ^ cur) & 0xFFFFu) == 0) {
                    t[j] = tgt[j];  // the probe reports a match only from 2 common bytes on
                } else if (INNER || p > 0) {
                    const uint32_t prev = j > 0 ? strip_word(d, j - 1) : prev0;
                    if (((prev ^ cur) & 0xFFFFu) == 0) t[j] = p - 1;  // src[p-1] == src[p] == src[p+1]
                }
            }
            any |= t[j] != GC_NONE;
        }
        if constexpr (!LENGTHS) return;
        auto maxlen = [&](uint32_t p) { return INNER ? GC_LZMA_MAX : (n - p < GC_LZMA_MAX ? n - p : GC_LZMA_MAX); };
        // cont bit j: position j's diagonal continues at j + 1 (for j = 7, the
        // next lane's first position; for lane 63, the first of the step above)
        const uint32_t tn = (uint32_t)__builtin_amdgcn_update_dpp((int)carry_t, (int)t[0], 0x130, 0xF, 0xF, false);
        uint32_t cont = 0;
#pragma unroll
        for (int j = 0; j < (int)STRIP; j++) {
            const uint32_t nt = j + 1 < (int)STRIP ? t[j + 1] : tn;
            if (t[j] != GC_NONE && nt != GC_NONE && nt == t[j] + 1) cont |= 1u << j;
        }
        // run ends, measured
        uint32_t L[STRIP];
#pragma unroll
        for (int j = 0; j < (int)STRIP; j++)
            L[j] = (t[j] != GC_NONE && !((cont >> j) & 1)) ? lcp_lane<true>(src, t[j], p0 + j, 2, maxlen(p0 + j)) : 0;
        // this lane's length at its first position where its run ends inside
        // the strip (else unused)
        uint32_t v = 0;
#pragma unroll
        for (int j = STRIP - 1; j >= 0; j--)
            v = t[j] == GC_NONE ? 0 : ((cont >> j) & 1) ? min(v + 1, GC_LZMA_MAX) : L[j];
        // the length at p0 + STRIP: `run` whole strips continue after this
        // lane's, then a lane (or the step above) whose run ends
        const uint64_t whole = ballot(cont == (1u << STRIP) - 1);
        const uint32_t lane = lane_id();
        const uint64_t after = lane + 1 < WAVE ? whole >> (lane + 1) : 0;
        const uint32_t run = (uint32_t)__builtin_ctzll(~after);
        const uint32_t vm = (uint32_t)__builtin_amdgcn_ds_bpermute((int)((lane + run + 1) << 2), (int)v);
        const uint32_t in = min((lane + run + 1 < WAVE ? vm : carry_len) + STRIP * run, GC_LZMA_MAX);
        uint32_t out[STRIP];
        uint32_t nl = in;
#pragma unroll
        for (int j = STRIP - 1; j >= 0; j--) {
            const uint32_t p = p0 + j;
            const uint32_t len = t[j] == GC_NONE ? 0 : ((cont >> j) & 1) ? min(nl + 1, maxlen(p)) : L[j];
            out[j] = len ? ((p - t[j]) << 16) | len : 0;
            nl = len;
        }
        carry_t = bcast(t[0], 0);
        carry_len = bcast(nl, 0);
        if (!REF && full) {
            uint4* q = (uint4*)(packed + it.opos + p0);
            q[0] = make_uint4(out[0], out[1], out[2], out[3]);
            q[1] = make_uint4(out[4], out[5], out[6], out[7]);
            return;
        }
#pragma unroll
        for (uint32_t j = 0; j < STRIP; j++) {
            const uint32_t p = p0 + j;
            if (p >= n) break;
            const uint64_t q = it.opos + p;
            if (!REF || packed) packed[q] = out[j];
            if (REF) {
                ref_c[q] = out[j] ? 1u : 0u;
                if (out[j]) {
                    gc_lzma_match m;
                    m.distance = out[j] >> 16;
                    m.length = out[j] & 0xFFFFu;
                    ref_m[q * max_matches] = m;
                }
            }
        }
    };
    // wavefront w's region: [w * REGION, (w + 1) * REGION), steps of WAVE
    // strips from the last
    constexpr uint32_t STEP_POS = WAVE * STRIP;
    constexpr uint32_t REGION = GC_SUB_BLOCK / (INFO_THREADS / WAVE);
    const uint32_t rbase = uniform((threadIdx.x / WAVE) * REGION);
    uint32_t carry_t = GC_NONE, carry_len = 0;
    for (int k = (int)(REGION / STEP_POS) - 1; k >= 0; k--) {
        const uint32_t sbase = rbase + (uint32_t)k * STEP_POS;
        if (sbase >= n) continue;
        const uint32_t p0 = sbase + lane_id() * STRIP;
        if (wide && sbase > 0 && sbase + STEP_POS + GC_LZMA_MAX <= n)
            strip(p0, std::true_type{}, carry_t, carry_len);
        else
            strip(p0, std::false_type{}, carry_t, carry_len);
    }
    // positions past the end of a short sub-block (lzma2_match_find.cu:188-190)
    if (LENGTHS && REF)
        for (uint32_t p = n + threadIdx.x; p < it.out_span; p += INFO_THREADS) ref_c[it.opos + p] = 0;
    return any;
}

// One kernel per output (scalar registers bound the occupancy): packed
// words (gc_lzma2_len_kernel), the reference layout with optional packed
// words (gc_lzma2_len_ref_kernel), or the triage byte alone
// (gc_lzma2_triage_kernel).  triage (optional): one byte per item at
// it.count_pos, 1 if any position matches.
template <bool LENGTHS, bool REF>
__device__ __forceinline__ void lzma2_len_items(const uint8_t* arena, const gc_item* __restrict__ items,
                                                uint32_t n_items, const uint16_t* cand16, uint32_t* packed,
                                                gc_lzma_match* ref_m, uint32_t* ref_c, uint32_t max_matches,
                                                uint8_t* triage, gc_queue* q = nullptr) {
    __shared__ uint32_t sb[STAGE_WORDS];
    auto body = [&](uint32_t i) {
        const gc_item it = load_item(items, i);
        __syncthreads();  // the last item's reads of sb are done
        stage_sub_block<INFO_THREADS>(sb, src_view(arena, it.src_off, it.avail), it.n);
        __syncthreads();
        const bool any = lzma2_item<LENGTHS, REF>(it, LdsBytes{sb}, cand16, packed, ref_m, ref_c, max_matches);
        if (triage) {
            const int wg_any = __syncthreads_or(any ? 1 : 0);
            if (threadIdx.x == 0) triage[it.count_pos] = wg_any ? 1 : 0;
        }
    };
    if (q) {
        items += queue_base(q, GC_Q_LZMA2);
        wg_queue_items(q, GC_Q_LZMA2, body);
    }
    else
        for (uint32_t i = blockIdx.x; i < n_items; i += gridDim.x) body(i);
}

#define GC_LZMA2_KERNEL(NAME, LENGTHS, REF)                                                                        \
    extern "C" __global__ void __launch_bounds__(INFO_THREADS, 8)                                                \
        NAME(const uint8_t* arena, const gc_item* __restrict__ items, uint32_t n_items, const uint16_t* cand16,  \
             uint32_t* packed, gc_lzma_match* ref_m, uint32_t* ref_c, uint32_t max_matches, uint8_t* triage) {   \
        lzma2_len_items<LENGTHS, REF>(arena, items, n_items, cand16, packed, ref_m, ref_c, max_matches, triage); \
    }
GC_LZMA2_KERNEL(gc_lzma2_len_kernel, true, false)
GC_LZMA2_KERNEL(gc_lzma2_len_ref_kernel, true, true)
GC_LZMA2_KERNEL(gc_lzma2_triage_kernel, false, false)

// Persistent form of the packed-word kernel: items from the device work queue
// (phase GC_Q_LZMA2).
extern "C" __global__ void __launch_bounds__(INFO_THREADS, 8)
gc_lzma2_len_q_kernel(const uint8_t* arena, const gc_item* __restrict__ items, gc_queue* q, const uint16_t* cand16,
                      uint32_t* packed) {
    lzma2_len_items<true, false>(arena, items, 0, cand16, packed, nullptr, nullptr, 1, nullptr, q);
}

extern "C" __global__ void __launch_bounds__(INFO_THREADS, 8)
gc_lzma2_len_sq_kernel(const uint8_t* arena, const gc_item* __restrict__ items, gc_queue* q, const uint16_t* cand16,
                       uint32_t* packed) {
    lzma2_len_items<true, false>(arena, items, 0, cand16, packed, nullptr, nullptr, 1, nullptr, q);
}

// ---------------------------------------------------------------------------
// Zstd per position, given cand(p): the candidate as the reference's table
// returns it (an empty slot reads as position 0, zstd_compress.cu:198-200),
// and whether it verifies 3 bytes (:241-245) with the common length capped
// at GC_ZSTD_SHORT: info = len << 16 | (p - cand) & 0xFFFF, len 0 = no (the
// low half then means nothing).  The low half of a verified word is the
// match offset (1..65535), the form the parse works in.  A length equal to
// the cap is re-measured by the parse.  Written for p + 3 < n.
//
// One workgroup of INFO_THREADS per sub-block, the sub-block staged in LDS
// first: the candidate reads and the length measurements land anywhere in
// the 64 KiB before p, and from LDS they cost no L2 or HBM traffic (from
// global memory, the sub-blocks of every workgroup in flight on an XCD do
// not fit its L2).
// ---------------------------------------------------------------------------

// Returns how many of this thread's positions have a verified candidate.
template <uint32_t W>
__device__ __forceinline__ uint32_t zstd_info_lds(const LdsBytes& L, uint32_t n, const uint16_t* cand,
                                                  uint32_t* out) {
    uint32_t verified = 0;
    // cand and out start at multiples of 8 positions and each item owns its
    // positions rounded up to 8: whole strips are read and written with
    // vector accesses.  The next strip's candidates are loaded one strip
    // early into the other of two register sets (a strip past the end
    // reloads the current one): a loop-carried copy of registers whose load
    // is in flight would wait for it.
    auto load_cand = [&](uint32_t p0, uint32_t cv[W / 2]) {
        if constexpr (W == 8) {
            const uint4 cw = *(const uint4*)(cand + p0);
            cv[0] = cw.x, cv[1] = cw.y, cv[2] = cw.z, cv[3] = cw.w;
        } else {
            const uint2 cw = *(const uint2*)(cand + p0);
            cv[0] = cw.x, cv[1] = cw.y;
        }
    };
    auto strip = [&](uint32_t p0, const uint32_t cv[W / 2]) {
        uint32_t d[4];
        d[0] = L.w[p0 / 4];
        d[1] = L.w[p0 / 4 + 1];
        d[2] = L.w[p0 / 4 + 2];
        d[3] = L.w[p0 / 4 + 3];
        uint32_t c[W], vc[W];
#pragma unroll
        for (uint32_t j = 0; j < W; j++) {
            const uint32_t cj = (cv[j >> 1] >> (16 * (j & 1))) & 0xFFFFu;
            c[j] = (cj == GC_NONE || p0 + j + 3 >= n) ? 0 : cj;
        }
#pragma unroll
        for (uint32_t j = 0; j < W; j++) vc[j] = ld32(L, c[j]);
        uint32_t o[W];
        uint32_t next_len = 0;
#pragma unroll
        for (int j = W - 1; j >= 0; j--) {
            const uint32_t p = p0 + j;
            uint32_t len = 0;
            if (p + 3 < n && p > 0 && ((vc[j] ^ strip_word(d, j)) & 0xFFFFFFu) == 0) {
                const uint32_t lim = n - p < GC_ZSTD_SHORT ? n - p : GC_ZSTD_SHORT;
                if (j + 1 < (int)W && next_len && c[j + 1] == c[j] + 1)
                    len = next_len + 1 < lim ? next_len + 1 : lim;
                else
                    len = lcp_lane<true>(L, c[j], p, 3, lim);
            }
            next_len = len;
            verified += len ? 1u : 0u;
            o[j] = (len << 16) | ((p - c[j]) & 0xFFFFu);
        }
        uint4* q = (uint4*)(out + p0);
        q[0] = make_uint4(o[0], o[1], o[2], o[3]);
        if constexpr (W == 8) q[1] = make_uint4(o[4], o[5], o[6], o[7]);
    };
    constexpr uint32_t STEP = INFO_THREADS * W;
    uint32_t ca[W / 2], cb[W / 2];
    uint32_t p0 = threadIdx.x * W;
    load_cand(p0 + 3 < n ? p0 : 0u, ca);
    // (the empty asm keeps each prefetch where it is: scheduled later, the
    // next use would wait for it)
    for (; p0 + 3 < n; p0 += 2 * STEP) {
        const uint32_t p1 = p0 + STEP;
        load_cand(p1 + 3 < n ? p1 : p0, cb);
        __asm__ volatile("" ::: "memory");
        strip(p0, ca);
        if (p1 + 3 >= n) break;
        const uint32_t p2 = p1 + STEP;
        load_cand(p2 + 3 < n ? p2 : p1, ca);
        __asm__ volatile("" ::: "memory");
        strip(p1, cb);
    }
    return verified;
}

// cost (optional): per item (at it.count_pos) the number of positions whose
// candidate verified, which is what the parse's work grows with.
// (cost is indexed by the item's index in `items`, the index the parse's
// order refers to)
template <uint32_t W>
__device__ __forceinline__ void zstd_info_items(const uint8_t* arena, const gc_item* __restrict__ items,
                                                uint32_t n_items, const uint16_t* cand16, uint32_t* info,
                                                uint32_t* cost, gc_queue* q = nullptr) {
    __shared__ uint32_t sb[STAGE_WORDS];
    __shared__ uint32_t verified;
    auto body = [&](uint32_t i) {
        const gc_item it = load_item(items, i);
        __syncthreads();  // the last item's reads of sb and `verified` are done
        if (threadIdx.x == 0) verified = 0;
        stage_sub_block<INFO_THREADS>(sb, src_view(arena, it.src_off, it.avail), it.n);
        __syncthreads();
        const uint32_t v = zstd_info_lds<W>(LdsBytes{sb}, it.n, cand16 + it.cpos, info + it.opos);
        if (cost) {
            if (v) atomicAdd(&verified, v);
            __syncthreads();
            if (threadIdx.x == 0) cost[i] = verified;
        }
    };
    if (q) {
        items += queue_base(q, GC_Q_INFO);
        wg_queue_items(q, GC_Q_INFO, body);
    }
    else
        for (uint32_t i = blockIdx.x; i < n_items; i += gridDim.x) body(i);
}

// Strips of 8 or 4 positions: the same words, different register use.
extern "C" __global__ void __launch_bounds__(INFO_THREADS, 8)
gc_zstd_info_kernel(const uint8_t* arena, const gc_item* __restrict__ items, uint32_t n_items,
                    const uint16_t* cand16, uint32_t* info, uint32_t* cost) {
    zstd_info_items<8>(arena, items, n_items, cand16, info, cost);
}

extern "C" __global__ void __launch_bounds__(INFO_THREADS, 8)
gc_zstd_info4_kernel(const uint8_t* arena, const gc_item* __restrict__ items, uint32_t n_items,
                     const uint16_t* cand16, uint32_t* info, uint32_t* cost) {
    zstd_info_items<4>(arena, items, n_items, cand16, info, cost);
}

// Persistent forms: items from the device work queue (phase GC_Q_INFO).
extern "C" __global__ void __launch_bounds__(INFO_THREADS, 8)
gc_zstd_info_q_kernel(const uint8_t* arena, const gc_item* __restrict__ items, gc_queue* q, const uint16_t* cand16,
                      uint32_t* info, uint32_t* cost) {
    zstd_info_items<8>(arena, items, 0, cand16, info, cost, q);
}

extern "C" __global__ void __launch_bounds__(INFO_THREADS, 8)
gc_zstd_info4_q_kernel(const uint8_t* arena, const gc_item* __restrict__ items, gc_queue* q, const uint16_t* cand16,
                       uint32_t* info, uint32_t* cost) {
    zstd_info_items<4>(arena, items, 0, cand16, info, cost, q);
}

extern "C" __global__ void __launch_bounds__(INFO_THREADS, 8)
gc_zstd_info_sq_kernel(const uint8_t* arena, const gc_item* __restrict__ items, gc_queue* q, const uint16_t* cand16,
                       uint32_t* info, uint32_t* cost) {
    zstd_info_items<8>(arena, items, 0, cand16, info, cost, q);
}

// The parse's item order, costliest first (counting sort of cost >> 9 into
// 129 buckets, descending): the sub-blocks that take longest start first and
// the launch does not end on a few long ones.  One workgroup of 256 threads
// (a few microseconds of work): in the served form it is launched while the
// other tick slots' kernels fill the CUs, and a 1024-thread workgroup waited
// up to 12 ms for a CU with 16 free wavefront slots.
#define ORDER_BUCKETS 129
#define ORDER_THREADS 256
// (q: the item count is the queue's parse count)
extern "C" __global__ void __launch_bounds__(ORDER_THREADS) gc_order_items(const uint32_t* cost, uint32_t n, uint32_t* order,
                                                                  const gc_queue* q) {
    __shared__ uint32_t start[ORDER_BUCKETS];
    if (q) n = uniform(__atomic_load_n(&q->n[GC_Q_PARSE], __ATOMIC_RELAXED));
    for (uint32_t b = threadIdx.x; b < ORDER_BUCKETS; b += blockDim.x) start[b] = 0;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) atomicAdd(&start[min(cost[i] >> 9, ORDER_BUCKETS - 1u)], 1u);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (int b = ORDER_BUCKETS - 1; b >= 0; b--) {
            const uint32_t c = start[b];
            start[b] = acc;
            acc += c;
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x)
        order[atomicAdd(&start[min(cost[i] >> 9, ORDER_BUCKETS - 1u)], 1u)] = i;
}

#endif  // GC_BUILD_MAIN

#if GC_BUILD_PARSE
// ---------------------------------------------------------------------------
// Zstd parse (zstd_compress.cu:230-338), one wavefront per sub-block.
//
// The wavefront holds an aligned window of 64 positions x = wbase + lane
// (wbase a multiple of 64): the info word of x, the 3 bytes at x and, for
// each current repeat offset r, whether the 3 bytes at x - r are the same.
// The next window's info words and bytes are loaded one window early.  The
// first lane at or after pos whose position matches (hash candidate
// verified, or a repeat offset) is the next head: every position before it
// is a literal.  After a sequence the repeat offsets become (new, old0,
// old1): each lane renames its two kept flags and gathers the new one, so no
// data moves between lanes until the walk leaves the window.  Sequences are
// staged one per lane and stored 64 at a time.
//
// The walk's decisions are wavefront-uniform and run on the scalar unit,
// which one CU shares between its four SIMDs: that unit's instruction count
// per sequence sets this kernel's speed, so per-lane work stays in vector
// registers and loop-carried flags are plain integers.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void copy_wave(uint8_t* dst, gbytes src, uint32_t len) {
    for (uint32_t i = lane_id(); i < len; i += WAVE) dst[i] = src[i];
}

// 3 bytes at x - r equal w (the 3 bytes at x), for a head candidate x
// (x + 3 < n); the read is clamped so that it is always in the sub-block.
template <bool SAFE>
__device__ __forceinline__ uint32_t rep_eq(const SrcView& src, uint32_t x, uint32_t r, uint32_t w) {
    const bool ok = r <= x;
    const uint32_t v = ld32<SAFE>(src, ok ? x - r : 0) & 0xFFFFFFu;
    return (ok && v == w) ? 1u : 0u;
}

// The parse's recent history in LDS: a ring of PARSE_RING bytes per
// wavefront, byte p of the sub-block at index p % PARSE_RING, plus a copy of
// its first 8 bytes past the end so that a dword pair read at any index stays
// inside.  While the walk is in the window [wbase, wbase + 64) the ring holds
// every byte of [wbase + 72 - PARSE_RING, wbase + 66) (the last lane's
// 3-byte word gives the 2 bytes past the window): a repeat offset of at most
// RING_REACH is compared for every lane of the window without a global load,
// which on this path is a dependent round trip to HBM.
#define PARSE_RING 4096u
#define RING_REACH (PARSE_RING - WAVE - 8)

#define RING_BYTES (PARSE_RING + 16)  // + the copy of bytes 0..7 + a junk slot
#define RING_JUNK (PARSE_RING + 8)

// The 3 bytes at p (little endian) from the ring.
__device__ __forceinline__ uint32_t ring_w24(const uint8_t* ring, uint32_t p) {
    const uint32_t* r = (const uint32_t*)(ring + (p & (PARSE_RING - 4)));
    return __builtin_amdgcn_alignbyte(r[1], r[0], p) & 0xFFFFFFu;
}

// A segment of a sub-block's parse (the segmented form, gc_zstd_parse_seg_kernel):
// the walk starts at `start` with the given repeat offsets, makes every decision
// at a position below `stop`, writes its sequences as (head, offset | length
// << 16) records and reports how it ended: the end of its last sequence when
// that passes `stop`, else `stop` (a literal run reached it), and the offsets.
//
// A fix-up walk (SEG_FIX) also has the records of an earlier walk of the same
// segment from another start state (xs, x0..x2): its states are that start and
// the (end, offsets) after each of its records.  The parse's future from a
// state depends on nothing else, so once the fix-up walk reaches one of those
// states the earlier walk's records from there on are the fix-up walk's own:
// it stops and reports which (conv: the index of the first such record).
enum { SEG_NONE = 0, SEG_SPEC = 1, SEG_FIX = 2 };
struct SegWalk {
    uint32_t start, r0, r1, r2, stop;  // in
    uint2* rec;                        // in: room for (stop - start) / 3 + 2 records
    const uint2* srec;                 // in (SEG_FIX): the earlier walk's records
    uint32_t scnt, xs, x0, x1, x2;     // in (SEG_FIX): their count and the earlier walk's start state
    uint32_t count, ypos, y0, y1, y2;  // out (wavefront-uniform)
    uint32_t conv;                     // out (SEG_FIX): first record of the earlier walk that follows, or ~0u
};

template <bool SAFE, uint32_t FMT, bool LITS, bool LAZY, uint32_t SEG = SEG_NONE>
__device__ __forceinline__ void zstd_parse_item(const gc_item& it, const uint8_t* arena, const uint32_t* info,
                                                void* seq_out, uint8_t* literals, uint32_t* seq_counts,
                                                uint32_t* lit_counts, uint32_t max_seq, uint8_t* ring,
                                                SegWalk* sw = nullptr) {
    const SrcView src = src_view(arena, it.src_off, it.avail);
    const uint32_t n = it.n;
    const uint32_t lane = lane_id();
    const uint8_t* inf = (const uint8_t*)(info + it.opos);
    uint32_t rep0 = 1, rep1 = 4, rep2 = 8;  // zstd_compress.cu:225
    uint32_t pos = 0, nseq = 0, wbase = 0;  // the window is [wbase, wbase + WAVE)
    uint32_t last_end = 0;                  // SEG: end of the last sequence (the start while there is none)
    if constexpr (SEG != SEG_NONE) {
        rep0 = sw->r0, rep1 = sw->r1, rep2 = sw->r2;
        pos = sw->start;
        wbase = pos & ~(WAVE - 1);
        last_end = pos;
    }
    // SEG_FIX: the earlier walk's records 64 at a time from record cb: lane l
    // holds record cb + l's end (~0u past the last) and offset; po1 and po2
    // are the offsets of records cb - 1 and cb - 2 (before the first record,
    // the earlier start's offsets).  States are compared at the ends of this
    // walk's sequences, whose positions only grow.
    uint32_t cb = 0, ce = 0, co = 0, po1 = 0, po2 = 0, conv = ~0u;
    auto conv_load = [&]() {
        const uint32_t i = cb + lane;
        const uint2 r = i < sw->scnt ? sw->srec[i] : make_uint2(0u, 0u);
        ce = i < sw->scnt ? r.x + (r.y >> 16) : ~0u;
        co = r.y & 0xFFFFu;
    };
    // is (p, rep0, rep1, rep2) a state of the earlier walk? (sets conv)
    auto conv_at = [&](uint32_t p) -> bool {
        while (bcast(ce, WAVE - 1) < p) {  // (a lane past the last record reads ~0u: the loop ends)
            po1 = bcast(co, WAVE - 1);
            po2 = bcast(co, WAVE - 2);
            cb += WAVE;
            conv_load();
        }
        const uint64_t m = ballot(ce == p);
        if (!m) return false;
        const uint32_t k = (uint32_t)__builtin_ctzll(m);
        const uint32_t o0 = bcast(co, k);
        const uint32_t o1 = k >= 1 ? bcast(co, k - 1) : po1;
        const uint32_t o2 = k >= 2 ? bcast(co, k - 2) : (k == 1 ? po1 : po2);
        if (o0 != rep0 || o1 != rep1 || o2 != rep2) return false;
        conv = cb + k + 1;
        return true;
    };
    if constexpr (SEG == SEG_FIX) {
        po1 = sw->x0, po2 = sw->x1;
        conv_load();
        if (pos == sw->xs && rep0 == sw->x0 && rep1 == sw->x1 && rep2 == sw->x2)
            conv = 0;  // the earlier walk's own start
        else
            conv_at(pos);
    }
    // literal bookkeeping for the literal copy, in vector registers (the same
    // in every lane): the scalar register count limits this kernel's occupancy
    uint32_t lit_start = in_vgpr(0u), nlit = in_vgpr(0u);
    uint8_t* lit_dst = LITS ? literals + it.lit_pos : nullptr;

    // staged sequences: lane j holds sequence (nseq & ~63) + j as
    // offset | length << 16 (both fit 16 bits: a head is at position >= 1, so
    // offset and length are below 65536) and its head position, each written
    // into its lane by one instruction.  Literal lengths follow when the
    // records are stored: head minus the end of the sequence before.
    uint32_t q_ol = 0, q_h = 0;
    // (uniform, in vector registers like the literal bookkeeping)
    uint32_t batch_end = in_vgpr(0u);  // end of the last sequence before the staged ones
    uint32_t matched = in_vgpr(0u);    // bytes covered by the stored sequences
    auto store_staged = [&](uint32_t first, uint32_t cnt) {  // sequences first .. first + cnt - 1
        if constexpr (SEG != SEG_NONE) {
            if (lane < cnt) sw->rec[first + lane] = make_uint2(q_h, q_ol);
            return;
        }
        const uint32_t len = q_ol >> 16;
        const uint32_t end = q_h + len;
        const uint32_t ll = q_h - wave_shr1(end, batch_end);
        if (lane < cnt) {
            const uint64_t si = it.seq_pos + first + lane;
            if constexpr (FMT == GC_SEQ_REF) {
                gc_ref_sequence* o = (gc_ref_sequence*)seq_out + si;
                o->literal_length = ll;
                o->match_length_minus3 = len - 3;
                o->offset = q_ol & 0xFFFFu;
            } else {
                *(uint4*)((gc_zstd_sequence*)seq_out + si) = make_uint4(q_ol & 0xFFFFu, ll, len, 0);
            }
        }
        matched = in_vgpr(uniform(matched) + wave_sum(lane < cnt ? len : 0u));
        batch_end = in_vgpr(bcast(end, cnt - 1));
    };

    // per lane, for x = wbase + lane: e = info(x), e1 = info(x + 1) (the
    // lazy step's word), w = the 3 bytes at x (0 past the end; the low byte
    // goes to the ring) with bit 24 set where x + 3 >= n, so that no 3-byte
    // comparison with it succeeds there; f = repeat flags; ae, ae1, aw the
    // same for the next window.  Info words of positions with x + 3 >= n read
    // as 0.
    uint32_t e = 0, e1 = 0, w = 0, f = 0, ae = 0, ae1 = 0, aw = 0;
    uint32_t xv = lane;  // wbase + lane
    // peek issues the loads only: the words go straight into the registers
    // that carry them (for the next window, ae, ae1, aw) and are masked by
    // finish where they are used.  Masking them at the load would wait for
    // the loads there, a full memory round trip on every window change.
    // The 4 bytes at x come from one unaligned dword load (gfx950 under ROCm
    // runs with unaligned memory access enabled), so no byte shift is left
    // to do either.
    auto peek = [&](uint32_t x, uint32_t& ev, uint32_t& ev1, uint32_t& wv) {
        const bool act = x + 3 < n;
        const uint32_t xc = act ? x : 0;  // reads stay in the sub-block (x + 1 < n here)
        const uint32_t* ip = (const uint32_t*)(inf + (xc << 2));
        ev = ip[0];
        wv = ld32u<SAFE>(src, x < n ? x : 0u);
        if constexpr (LAZY) ev1 = ip[1];
    };
    // finish masks what peek left for the window at nb (x = nb + lane).  The
    // info words' low halves are already match offsets (x - cand), so a
    // head's offset is one scalar AND.  A window whose every position has
    // x + 4 < n (all but a sub-block's last window or two: a uniform branch)
    // keeps its words as loaded, but for the bytes' 4th byte.
    auto finish = [&](uint32_t nb) {
        if (nb + WAVE + 4 <= n) {
            w &= 0xFFFFFFu;
            return;
        }
        const uint32_t x = nb + lane;
        const bool act = x + 3 < n;
        e = act ? e : 0;
        w = (x < n ? (w & 0xFFFFFFu) : 0) | (act ? 0 : 0x1000000u);
        if constexpr (LAZY) e1 = x + 4 < n ? e1 : 0;
    };
    // Repeat offsets beyond the ring's reach are compared with a global read
    // where they are needed.  (Prefetching their next window's words at each
    // window entry cost more than it saved: every later wait for a direct
    // read, which counts loads in order, waited for those loads too.)
    auto far_w24 = [&](uint32_t y, uint32_t r) -> uint32_t {  // the 3 bytes at y - r (clamped read)
        return ld32<SAFE>(src, (y >= r && y < n) ? y - r : 0u) & 0xFFFFFFu;
    };
    // (far: some current offset is beyond the ring's reach; one test per
    // window change instead of one per offset)
    auto any_far = [&]() { return smax(rep0, smax(rep1, rep2)) > RING_REACH; };
    // the repeat flags of the window at x - lane: from the ring, or for a
    // far offset a global read
    // (wv comes from peek: bit 24 is set where x + 3 >= n, so no flag is set
    // there)
    auto flags = [&](uint32_t x, uint32_t wv, bool far) -> uint32_t {
        // every lane reads (indexes wrap inside the ring), then masks
        uint32_t v0 = ring_w24(ring, x - rep0), v1 = ring_w24(ring, x - rep1), v2 = ring_w24(ring, x - rep2);
        if (far) {
            if (rep0 > RING_REACH) v0 = far_w24(x, rep0);
            if (rep1 > RING_REACH) v1 = far_w24(x, rep1);
            if (rep2 > RING_REACH) v2 = far_w24(x, rep2);
        }
        // (a window at or past every offset, as all but a sub-block's first
        // windows are for the usual offsets, needs no x >= r masks)
        if (wbase >= smax(rep0, smax(rep1, rep2)))
            return (v0 == wv ? 1u : 0u) | (v1 == wv ? 2u : 0u) | (v2 == wv ? 4u : 0u);
        return (((x >= rep0) & (v0 == wv)) ? 1u : 0u) | (((x >= rep1) & (v1 == wv)) ? 2u : 0u) |
               (((x >= rep2) & (v2 == wv)) ? 4u : 0u);
    };
    // entering the window at nb (w already holds its bytes): the bytes from
    // `from` (a multiple of 64) up to nb are fetched into the ring first
    auto enter = [&](uint32_t nb, uint32_t from) {
        for (uint32_t p0 = from; p0 < nb; p0 += 4 * WAVE) {
            const uint32_t p = p0 + 4 * lane;
            if (p < nb) {  // p + 3 < nb <= n
                const uint32_t v = ld32<SAFE>(src, p);
                const uint32_t j = p & (PARSE_RING - 1);
                *(uint32_t*)(ring + j) = v;
                if (j < 8) *(uint32_t*)(ring + PARSE_RING + j) = v;
            }
        }
        // the window's bytes, and from the last lane the 2 bytes past it (the
        // other lanes write the junk slot rather than branch around the
        // write); the copy of bytes 0..7 past the ring's end only in the
        // windows that write them (one in 64, a uniform branch)
        const uint32_t jw = nb & (PARSE_RING - 1), je = (nb + WAVE) & (PARSE_RING - 1);
        ring[jw + lane] = (uint8_t)w;
        *(uint16_t*)(ring + (lane == WAVE - 1 ? je : RING_JUNK)) = (uint16_t)(w >> 8);
        if (jw == 0) {
            if (lane < 8) ring[PARSE_RING + lane] = (uint8_t)w;
        }
        if (je == 0) {
            if (lane == WAVE - 1) *(uint16_t*)(ring + PARSE_RING) = (uint16_t)(w >> 8);
        }
        __asm__ volatile("" ::: "memory");  // ring writes before the reads
    };
    // heads satisfy pos + 3 < n; the walk ends by lowering lim (SEG: and below stop)
    uint32_t lim = (n > 3 && max_seq > 0) ? n - 3 : 0;
    if constexpr (SEG != SEG_NONE) {
        lim = min(lim, sw->stop);
        if (conv != ~0u) lim = 0;  // (SEG_FIX) nothing to walk
        if (pos < lim) {  // the window holding the start, the ring filled as far back as it reaches
            xv = wbase + lane;
            peek(xv, e, e1, w);
            finish(wbase);
            enter(wbase, wbase + WAVE < PARSE_RING ? 0u : wbase + WAVE - PARSE_RING);
            f = flags(xv, w, any_far());
            __asm__ volatile("" ::: "memory");
            peek(wbase + WAVE + lane, ae, ae1, aw);
        }
    } else if (lim) {
        peek(lane, e, e1, w);
        finish(0);
        enter(0, 0);
        f = flags(lane, w, false);  // (offsets 1, 4, 8)
        __asm__ volatile("" ::: "memory");  // the window first, then the prefetch
        peek(WAVE + lane, ae, ae1, aw);
    }
    // move to the window at nb (a multiple of 64 past the current one); with
    // `prefetched`, the next window's prefetched words serve it when nb is
    // that window
    auto move_to = [&](uint32_t nb, bool prefetched) {
        const uint32_t wend = wbase + WAVE;
        const bool next = prefetched && nb == wend;
        uint32_t from = nb;
        if (next) {
            e = ae;
            e1 = ae1;
            w = aw;
        } else {
            peek(nb + lane, e, e1, w);
            // the windows jumped over, as far as the ring reaches
            from = nb + WAVE - PARSE_RING < wend || nb + WAVE < PARSE_RING ? wend : nb + WAVE - PARSE_RING;
        }
        wbase = nb;
        xv = nb + lane;
        finish(nb);
        enter(nb, from);
        f = flags(xv, w, any_far());
        __asm__ volatile("" ::: "memory");
        peek(nb + WAVE + lane, ae, ae1, aw);
    };
    // Look-ahead over a stretch of literals: from sb (a window start), 256
    // positions per step, 4 per lane, with every load of a step issued
    // together (a window-by-window slide waits for one load per window).
    // Returns the first window that may hold a head at or after sb (a
    // verified hash candidate, or a position whose 3 bytes equal those at
    // one of the current repeat offsets), or lim if there is none before
    // lim.  A lane whose positions straddle a repeat offset counts as a hit:
    // the window's own flags then decide.
    auto scan = [&](uint32_t sb) -> uint32_t {
        for (; sb < lim; sb += 4 * WAVE) {
            const uint32_t x0 = sb + 4 * lane;
            const uint32_t xc = x0 < lim ? x0 : 0u;  // x0 + 3 < n: reads stay in the sub-block
            const uint4 iw = *(const uint4*)(inf + (xc << 2));
            const uint32_t o0 = ld32<SAFE>(src, xc), o1 = ld32<SAFE>(src, xc + 4);
            const uint32_t ya = xc >= rep0 ? xc - rep0 : 0u, yb = xc >= rep1 ? xc - rep1 : 0u,
                           yc = xc >= rep2 ? xc - rep2 : 0u;
            const uint32_t a0 = ld32<SAFE>(src, ya), a1 = ld32<SAFE>(src, ya + 4);
            const uint32_t b0 = ld32<SAFE>(src, yb), b1 = ld32<SAFE>(src, yb + 4);
            const uint32_t c0 = ld32<SAFE>(src, yc), c1 = ld32<SAFE>(src, yc + 4);
            bool hit = false;
#pragma unroll
            for (uint32_t j = 0; j < 4; j++) {
                const uint32_t wj = __builtin_amdgcn_alignbyte(o1, o0, j) & 0xFFFFFFu;
                const uint32_t ij = j == 0 ? iw.x : j == 1 ? iw.y : j == 2 ? iw.z : iw.w;
                const bool m0 = ((__builtin_amdgcn_alignbyte(a1, a0, j) & 0xFFFFFFu) == wj) & (xc >= rep0);
                const bool m1 = ((__builtin_amdgcn_alignbyte(b1, b0, j) & 0xFFFFFFu) == wj) & (xc >= rep1);
                const bool m2 = ((__builtin_amdgcn_alignbyte(c1, c0, j) & 0xFFFFFFu) == wj) & (xc >= rep2);
                hit |= (x0 + j < lim) & ((ij > 0xFFFFu) | m0 | m1 | m2);
            }
            hit |= (x0 < lim) & (((x0 < rep0) & (x0 + 3 >= rep0)) | ((x0 < rep1) & (x0 + 3 >= rep1)) |
                                 ((x0 < rep2) & (x0 + 3 >= rep2)));
            const uint64_t hm = ballot(hit);
            if (hm) return (sb + 4 * (uint32_t)__builtin_ctzll(hm)) & ~(WAVE - 1);
        }
        return lim;
    };
    // next sequence count at which the staged records are stored or the
    // walk stops (max_seq)
    uint32_t ev_at = max_seq < WAVE ? max_seq : WAVE;
    while (pos < lim) {
        // two ballots of plain compares: a ballot of their conjunction is
        // built as a select and a third compare
        const uint64_t m = ballot(xv >= pos) & ballot(((e >> 16) | f) != 0);
        if (!m) {
            // no head left in the window: every position up to its end is a
            // literal.  Slide to the next window (its words are prefetched);
            // if that one has no head either, look further ahead at once.
            // (Testing the new window here keeps a "just slid" flag, and its
            // reset on every head, out of the walk.)
            wbase += WAVE;
            pos = wbase;
            xv += WAVE;
            e = ae;
            e1 = ae1;
            w = aw;
            finish(wbase);
            enter(wbase, wbase);
            f = flags(xv, w, any_far());
            __asm__ volatile("" ::: "memory");
            peek(wbase + WAVE + lane, ae, ae1, aw);
            if (pos < lim && ballot(((e >> 16) | f) != 0) == 0) {
                const uint32_t nb = scan(wbase + WAVE);
                if (nb >= lim) {
                    pos = lim;  // literals to the end
                } else {
                    move_to(nb, false);  // (the prefetched words are not kept across the scan)
                    pos = nb;
                }
            }
        } else {
            const uint32_t k = (uint32_t)__builtin_ctzll(m);
            uint32_t head = wbase + k;
            if constexpr (SEG != SEG_NONE) {
                if (head >= lim) {  // the next head is the next segment's: a literal run reaches stop
                    pos = lim;
                    continue;
                }
            }
            const uint32_t ek = bcast(e, k);
            uint32_t len, off, ol;  // ol: the staged record word, off | len << 16
            if (ek > 0xFFFFu) {
                // hash candidate verified (ek's low half is its offset, so ek
                // is the record word unless the length is measured further)
                off = ek & 0xFFFFu;
                len = ek >> 16;
                ol = ek;
                if (len >= GC_ZSTD_SHORT) {
                    len = lcp_wave<SAFE>(src, head - off, head, n - head, GC_ZSTD_SHORT);
                    ol = off | (len << 16);
                }
            } else {
                // the first repeat offset that verifies (rep0, rep1, rep2 in order)
                const uint32_t fk = bcast(f, k);  // (read here: only this path needs the head's flags)
                off = (fk & 1) ? rep0 : ((fk & 2) ? rep1 : rep2);
                if (off <= RING_REACH) {
                    // the bytes up to the window end are in the ring
                    const uint32_t span = (wbase + WAVE < n ? wbase + WAVE : n) - head;
                    const uint32_t ba = ring[(head - off + lane) & (PARSE_RING - 1)];
                    const uint32_t bb = ring[(head + lane) & (PARSE_RING - 1)];
                    // lanes 3 .. span - 1 whose bytes differ (masks from compares, combined
                    // on the scalar unit)
                    const uint64_t mm = __builtin_amdgcn_uicmp(ba, bb, 33) & __builtin_amdgcn_uicmp(lane, span, 36) &
                                        ~7ull;
                    if (mm)
                        len = (uint32_t)__builtin_ctzll(mm);
                    else
                        len = lcp_wave<SAFE>(src, head - off, head, n - head, span > 3 ? span : 3u);
                } else {
                    len = lcp_wave<SAFE>(src, head - off, head, n - head, 3);
                }
                ol = off | (len << 16);
            }
            if constexpr (LAZY) {
                // one lazy step at head + 1 (zstd_compress.cu:278-307)
                const uint32_t e2 = bcast(e1, k);
                // it can win only if the word at head + 1 is at least
                // min(len + 2, GC_ZSTD_SHORT) long (a length of GC_ZSTD_SHORT
                // or more reads as GC_ZSTD_SHORT): one compare of packed words
                // rules out most heads (ol >> 16 is len; an overflow only lets
                // a head through to the exact test below)
                const uint32_t need = min((ol & 0xFFFF0000u) + 0x20000u, GC_ZSTD_SHORT << 16);
                if (e2 >= need) {
                    const uint32_t off2 = e2 & 0xFFFFu;  // (the offset from head + 1)
                    uint32_t len2 = e2 >> 16;
                    if (len2 >= GC_ZSTD_SHORT)
                        len2 = lcp_wave<SAFE>(src, head + 1 - off2, head + 1, n - head - 1, GC_ZSTD_SHORT);
                    if (len2 > len + 1) {
                        head += 1;
                        len = len2;
                        off = off2;
                        ol = off | (len << 16);
                    }
                }
            }
            // emit (nseq < max_seq holds here)
            if constexpr (LITS) {
                const uint32_t ll = head - lit_start;
                copy_wave(lit_dst + uniform(nlit), src.base + src.mis + uniform(lit_start), uniform(ll));
                nlit += ll;
            }
            write_lane2(q_ol, ol, q_h, head, nseq & (WAVE - 1));
            nseq++;
            if (nseq == ev_at) {
                if ((nseq & (WAVE - 1)) == 0) store_staged(nseq - WAVE, WAVE);
                if (nseq == max_seq) lim = 0;  // the reference stops at the next head
                ev_at = nseq + WAVE < max_seq ? nseq + WAVE : max_seq;
            }
            // repeat offsets (off, o0, o1)
            rep2 = rep1;
            rep1 = rep0;
            rep0 = off;
            pos = head + len;
            if constexpr (SEG != SEG_NONE) last_end = pos;
            if constexpr (SEG == SEG_FIX) {
                if (conv_at(pos)) break;  // the earlier walk's records follow from here
            }
            if constexpr (LITS) lit_start = in_vgpr(pos);
            if (pos >= wbase + WAVE) {
                // the walk left the window: move to the window holding pos
                // (reads past the sub-block end are clamped by peek)
                move_to(pos & ~(WAVE - 1), true);
            } else {
                // the new offset's flags (a flag depends only on the offset
                // and the position, so a repeat match's offset gets the same
                // bits it had: recomputed, not renamed, which keeps the rare
                // repeat path's choice out of the common one)
                uint32_t b0;
                if (off <= RING_REACH) {
                    const uint32_t v = ring_w24(ring, xv - off);
                    // (w has bit 24 set where xv + 3 >= n; the x >= off mask
                    // only where the window starts before the offset)
                    if (off <= wbase)
                        b0 = v == w ? 1u : 0u;
                    else
                        b0 = ((xv >= off) & (v == w)) ? 1u : 0u;
                } else if (off == rep1) {  // a far offset equal to the old rep0 or rep1: its flags
                    b0 = f & 1u;
                } else if (off == rep2) {
                    b0 = (f >> 1) & 1u;
                } else {
                    const bool act = xv + 3 < n;
                    b0 = rep_eq<SAFE>(src, act ? xv : 0u, off, w & 0xFFFFFFu) & (act ? 1u : 0u);
                }
                f = b0 | ((f & 3u) << 1);
            }
        }
    }
    if (nseq & (WAVE - 1)) store_staged(nseq & ~(uint32_t)(WAVE - 1), nseq & (WAVE - 1));
    if constexpr (SEG != SEG_NONE) {
        sw->count = nseq;
        sw->conv = conv;
        sw->ypos = last_end >= sw->stop ? last_end : sw->stop;
        sw->y0 = rep0, sw->y1 = rep1, sw->y2 = rep2;
        return;
    }
    // trailing literals (zstd_compress.cu:341-344): every byte not covered
    // by a sequence is a literal
    const uint32_t lit_end = nseq ? uniform(batch_end) : 0u;
    const uint32_t nlit_total = n - uniform(matched);
    if constexpr (LITS)
        if (n > lit_end) copy_wave(lit_dst + uniform(nlit), src.base + src.mis + lit_end, n - lit_end);
    if (lane == 0) {
        seq_counts[it.count_pos] = nseq;
        if (lit_counts) lit_counts[it.count_pos] = nlit_total;
    }
}

template <uint32_t FMT, bool LITS>
__device__ __forceinline__ void zstd_parse_one(const uint8_t* arena, const gc_item* __restrict__ items, uint32_t k,
                                               const uint32_t* info, void* seq_out, uint8_t* literals,
                                               uint32_t* seq_counts, uint32_t* lit_counts, uint32_t max_seq,
                                               uint8_t* ring);

template <uint32_t FMT, bool LITS>
__device__ __forceinline__ void zstd_parse_items(const uint8_t* arena, const gc_item* __restrict__ items,
                                                 uint32_t n_items, const uint32_t* info, void* seq_out,
                                                 uint8_t* literals, uint32_t* seq_counts, uint32_t* lit_counts,
                                                 uint32_t max_seq, const uint32_t* __restrict__ order,
                                                 uint32_t (*ring_mem)[RING_BYTES / 4]) {
    // the wavefront index is uniform: items are read into scalar registers
    const uint32_t wave = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
    uint8_t* ring = (uint8_t*)ring_mem[wave];
    for (uint32_t i = blockIdx.x * PARSE_WAVES + wave; i < n_items; i += gridDim.x * PARSE_WAVES)
        zstd_parse_one<FMT, LITS>(arena, items, order ? order[i] : i, info, seq_out, literals, seq_counts, lit_counts,
                                  max_seq, ring);
}

// Persistent form: each wavefront takes items from the device work queue
// (phase GC_Q_PARSE) in `order`, at most n + 1 fetches.
template <uint32_t FMT, bool LITS>
__device__ __forceinline__ void zstd_parse_queue(const uint8_t* arena, const gc_item* __restrict__ items, gc_queue* q,
                                                 const uint32_t* info, void* seq_out, uint32_t* seq_counts,
                                                 uint32_t max_seq, const uint32_t* __restrict__ order,
                                                 uint32_t (*ring_mem)[RING_BYTES / 4]) {
    const uint32_t wave = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
    uint8_t* ring = (uint8_t*)ring_mem[wave];
    const uint32_t n = uniform(__atomic_load_n(&q->n[GC_Q_PARSE], __ATOMIC_RELAXED));
    items += queue_base(q, GC_Q_PARSE);
    uint32_t mine = 0;
    for (uint32_t iter = 0;; iter++) {
        uint32_t i = 0;
        if (lane_id() == 0) i = atomicAdd(&q->next[GC_Q_PARSE], 1u);
        i = uniform(i);
        if (iter > n) {
            if (lane_id() == 0) atomicOr(&q->err, 1u);
            break;
        }
        if (i >= n) break;
        zstd_parse_one<FMT, LITS>(arena, items, order ? order[i] : i, info, seq_out, nullptr, seq_counts, nullptr,
                                  max_seq, ring);
        mine++;
    }
    if (lane_id() == 0) queue_leave(q, GC_Q_PARSE, n, mine, gridDim.x * PARSE_WAVES);
}

template <uint32_t FMT, bool LITS>
__device__ __forceinline__ void zstd_parse_one(const uint8_t* arena, const gc_item* __restrict__ items, uint32_t k,
                                               const uint32_t* info, void* seq_out, uint8_t* literals,
                                               uint32_t* seq_counts, uint32_t* lit_counts, uint32_t max_seq,
                                               uint8_t* ring) {
    {
#if PARSE_UNIFORM_ITEM
        const gc_item it = load_item(items, k);
#else
        const gc_item it = items[k];  // fields in vector registers: fewer scalar registers in use
#endif
        const bool safe = item_safe(it.avail, it.n);
        if (it.depth > 1) {
            if (safe)
                zstd_parse_item<true, FMT, LITS, true>(it, arena, info, seq_out, literals, seq_counts, lit_counts,
                                                       max_seq, ring);
            else
                zstd_parse_item<false, FMT, LITS, true>(it, arena, info, seq_out, literals, seq_counts, lit_counts,
                                                        max_seq, ring);
        } else {
            if (safe)
                zstd_parse_item<true, FMT, LITS, false>(it, arena, info, seq_out, literals, seq_counts, lit_counts,
                                                        max_seq, ring);
            else
                zstd_parse_item<false, FMT, LITS, false>(it, arena, info, seq_out, literals, seq_counts, lit_counts,
                                                         max_seq, ring);
        }
    }
}

// One kernel per record layout and literal output, each with its own
// register allocation (scalar registers bound this kernel's occupancy).
// order (optional): the items in the order to take them.
//
// 8 waves per SIMD are asked for (the second launch bound): the compiler then
// keeps to 64 vector registers (the look-ahead scan would take 75) and to 78
// scalar ones, spilling about 28 scalar values to vector lanes.  Scalar
// registers are what this kernel's occupancy has turned on: the same parse
// at 98 took 12.4 ms per 2 GiB launch, at 94 11.0 ms and at 78 10.8 ms
// (the compiler counts each of them as 8 waves per SIMD).
#define GC_PARSE_KERNEL(NAME, FMT, LITS)                                                                          \
    extern "C" __global__ void __launch_bounds__(PARSE_WAVES* WAVE, 8)                                          \
        NAME(const uint8_t* arena, const gc_item* __restrict__ items, uint32_t n_items, const uint32_t* info,    \
             void* seq_out, uint32_t seq_fmt, uint8_t* literals, uint32_t* seq_counts, uint32_t* lit_counts,     \
             uint32_t max_seq, const uint32_t* __restrict__ order) {                                             \
        __shared__ uint32_t ring_mem[PARSE_WAVES][RING_BYTES / 4];                                          \
        zstd_parse_items<FMT, LITS>(arena, items, n_items, info, seq_out, literals, seq_counts, lit_counts,      \
                                    max_seq, order, ring_mem);                                                   \
    }
GC_PARSE_KERNEL(gc_zstd_parse_kernel, GC_SEQ_ZSTD, false)
GC_PARSE_KERNEL(gc_zstd_parse_lits_kernel, GC_SEQ_ZSTD, true)
GC_PARSE_KERNEL(gc_zstd_parse_ref_kernel, GC_SEQ_REF, false)
GC_PARSE_KERNEL(gc_zstd_parse_ref_lits_kernel, GC_SEQ_REF, true)

// Persistent form (the batched and served forms' record layout, no literals).
extern "C" __global__ void __launch_bounds__(PARSE_WAVES* WAVE, 8)
gc_zstd_parse_q_kernel(const uint8_t* arena, const gc_item* __restrict__ items, gc_queue* q, const uint32_t* info,
                       void* seq_out, uint32_t* seq_counts, uint32_t max_seq, const uint32_t* __restrict__ order) {
    __shared__ uint32_t ring_mem[PARSE_WAVES][RING_BYTES / 4];
    zstd_parse_queue<GC_SEQ_ZSTD, false>(arena, items, q, info, seq_out, seq_counts, max_seq, order, ring_mem);
}

extern "C" __global__ void __launch_bounds__(PARSE_WAVES* WAVE, 8)
gc_zstd_parse_sq_kernel(const uint8_t* arena, const gc_item* __restrict__ items, gc_queue* q, const uint32_t* info,
                        void* seq_out, uint32_t* seq_counts, uint32_t max_seq, const uint32_t* __restrict__ order) {
    __shared__ uint32_t ring_mem[PARSE_WAVES][RING_BYTES / 4];
    zstd_parse_queue<GC_SEQ_ZSTD, false>(arena, items, q, info, seq_out, seq_counts, max_seq, order, ring_mem);
}

// ---------------------------------------------------------------------------
// Segmented parse: one workgroup per sub-block, its walk split over
// GC_SEG_WAVES wavefronts.  Used where a launch has too few sub-blocks to
// fill the GPU with one walk per wavefront (one chunk of the plain drop-in):
// the serial walk's length, not the GPU's throughput, then sets the time.
//
// The parse's decisions from a state (a position where it looks for the next
// sequence, and the three repeat offsets) depend on nothing else: a
// position's hash candidate is a pure function of the sub-block.  So:
//   1. every wavefront j walks segment j (a 64-aligned range of positions)
//      from a guessed state: its first position with the initial offsets
//      (segment 0's guess is the parse's true start);
//   2. rounds of fix-ups.  A walk that finds no sequence passes its start's
//      offsets through to its end state.  Each round, segment j expects to
//      start from the end state of the nearest earlier segment that has
//      sequences (or segment 0's), moved on to j's first position; when its
//      records start from another state, it walks again from the expected
//      one until it reaches a state of its earlier walk (then that walk's
//      records from there on stand: usually within a few sequences) or its
//      end (then the new walk replaces the earlier one).  All fix-ups of a
//      round run at once.  When no segment has to walk, the walks chain from
//      segment 0's true start, so they are the serial walk; once the segments
//      before j are final, j's expected start is its true one and j is final
//      after its next fix-up, so there are at most GC_SEG_WAVES - 1 rounds.
//      (Incompressible data with one rare sequence: every later segment must
//      carry its offsets, and does so in one round, not one round each.)
//   3. each segment's records (the fix-up's, then the earlier walk's from the
//      state where they met) are written out in order after a prefix sum of
//      the segments' counts, with literal lengths, up to max_seq sequences
//      (the serial parse stops emitting there), and the literals if asked.
// The result is the serial parse's, record for record.
//
// Scratch: per workgroup, GC_SEG_WAVES x 2 record buffers of
// GC_SEG_REC_CAP (head, offset | length << 16) records (GC_SEG_SCRATCH
// bytes), the walk and its fix-up alternating between the two.
// ---------------------------------------------------------------------------
template <bool SAFE, bool LAZY, uint32_t MODE>
__device__ __forceinline__ void seg_walk(const gc_item& it, const uint8_t* arena, const uint32_t* info, uint8_t* ring,
                                         SegWalk* sw) {
    zstd_parse_item<SAFE, GC_SEQ_ZSTD, false, LAZY, MODE>(it, arena, info, nullptr, nullptr, nullptr, nullptr, ~0u,
                                                          ring, sw);
}

template <bool LAZY, uint32_t MODE>
__device__ __forceinline__ void seg_walk_any(bool safe, const gc_item& it, const uint8_t* arena, const uint32_t* info,
                                             uint8_t* ring, SegWalk* sw) {
    if (safe)
        seg_walk<true, LAZY, MODE>(it, arena, info, ring, sw);
    else
        seg_walk<false, LAZY, MODE>(it, arena, info, ring, sw);
}

template <bool LAZY>
__device__ __forceinline__ void zstd_parse_seg(const uint8_t* arena, const gc_item* __restrict__ items,
                                               uint32_t n_items, const uint32_t* info, void* seq_out, uint32_t fmt,
                                               uint8_t* literals, uint32_t* seq_counts, uint32_t* lit_counts,
                                               uint32_t max_seq, uint2* scratch) {
    __shared__ uint32_t ring_mem[GC_SEG_WAVES][RING_BYTES / 4];
    __shared__ uint32_t s_y[GC_SEG_WAVES][4];  // each segment's end state (position, offsets)
    __shared__ uint32_t s_tr[GC_SEG_WAVES];  // 1: the segment's records are none (its end state passes its start's offsets)
    __shared__ uint32_t s_cnt[GC_SEG_WAVES], s_end[GC_SEG_WAVES], s_len[GC_SEG_WAVES];
    __shared__ uint32_t s_out[2];  // sequences written, bytes they cover
    __shared__ uint32_t s_last;    // end of the last sequence written
    const uint32_t wave = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
    const uint32_t lane = lane_id();
    uint8_t* ring = (uint8_t*)ring_mem[wave];
    uint2* bufs = scratch + ((uint64_t)blockIdx.x * GC_SEG_WAVES + wave) * 2 * GC_SEG_REC_CAP;

    for (uint32_t k = blockIdx.x; k < n_items; k += gridDim.x) {
        const gc_item it = load_item(items, k);
        const uint32_t n = it.n;
        const bool safe = item_safe(it.avail, n);
        // segment bounds: 64-aligned, at most 65536 / GC_SEG_WAVES bytes (the last ones may be empty)
        const uint32_t seg = (((n + GC_SEG_WAVES - 1) / GC_SEG_WAVES) + WAVE - 1) & ~(WAVE - 1);
        const uint32_t start = min(wave * seg, n), stop = min(start + seg, n);

        // 1. the walk from the guessed state
        SegWalk sw;
        sw.start = start, sw.r0 = 1, sw.r1 = 4, sw.r2 = 8, sw.stop = stop;
        sw.rec = bufs;
        seg_walk_any<LAZY, SEG_SPEC>(safe, it, arena, info, ring, &sw);
        uint32_t xs = start, x0 = 1, x1 = 4, x2 = 8;  // the state the records in spec_buf start from
        uint32_t spec_buf = 0, scnt = sw.count, fcnt = 0, tail = 0;
        uint32_t yp = sw.ypos, y0 = sw.y0, y1 = sw.y1, y2 = sw.y2;
        if (lane == 0) s_y[wave][0] = yp, s_y[wave][1] = y0, s_y[wave][2] = y1, s_y[wave][3] = y2;

        // 2. fix-up rounds.  Each segment's expected start: the end state of the
        // nearest earlier segment that has sequences (or segment 0), moved on to
        // this segment's start -- a segment without sequences passes its start's
        // offsets through, so its successors need not wait for it to be redone.
        // A segment whose current records start elsewhere walks again from there.
        uint32_t wp = start, w0 = 1, w1 = 4, w2 = 8;  // the state the segment's records start from
        for (uint32_t round = 0; round < GC_SEG_WAVES; round++) {
            if (lane == 0) s_tr[wave] = (fcnt + scnt - tail) == 0 ? 1u : 0u;
            __syncthreads();  // end states and record counts of the last round
            uint32_t ep = 0, e0 = 0, e1 = 0, e2 = 0;
            bool need = false;
            if (wave > 0) {
                uint32_t m = wave - 1;
                while (m > 0 && s_tr[m]) m--;
                ep = max(s_y[m][0], start), e0 = s_y[m][1], e1 = s_y[m][2], e2 = s_y[m][3];
                need = ep != wp || e0 != w0 || e1 != w1 || e2 != w2;
            }
            if (!__syncthreads_or(need)) break;  // (also: every expected start read before any end state changes)
            if (need) {
                SegWalk fw;
                fw.start = ep, fw.r0 = e0, fw.r1 = e1, fw.r2 = e2, fw.stop = stop;
                fw.rec = bufs + (spec_buf ^ 1) * GC_SEG_REC_CAP;
                fw.srec = bufs + spec_buf * GC_SEG_REC_CAP;
                fw.scnt = scnt, fw.xs = xs, fw.x0 = x0, fw.x1 = x1, fw.x2 = x2;
                seg_walk_any<LAZY, SEG_FIX>(safe, it, arena, info, ring, &fw);
                wp = ep, w0 = e0, w1 = e1, w2 = e2;
                if (fw.conv != ~0u) {  // met the earlier walk: its end state stands
                    fcnt = fw.count;
                    tail = fw.conv;
                } else {  // a whole new walk of the segment
                    spec_buf ^= 1;
                    scnt = fw.count, fcnt = 0, tail = 0;
                    xs = ep, x0 = e0, x1 = e1, x2 = e2;
                    yp = fw.ypos, y0 = fw.y0, y1 = fw.y1, y2 = fw.y2;
                    if (lane == 0) s_y[wave][0] = yp, s_y[wave][1] = y0, s_y[wave][2] = y1, s_y[wave][3] = y2;
                }
            }
        }

        // 3. the segment's records: fix[0, fcnt) then spec[tail, scnt)
        const uint2* fix = bufs + (spec_buf ^ 1) * GC_SEG_REC_CAP;
        const uint2* spec = bufs + spec_buf * GC_SEG_REC_CAP;
        const uint32_t cnt = fcnt + scnt - tail;
        auto rec_at = [&](uint32_t i) -> uint2 { return i < fcnt ? fix[i] : spec[tail + i - fcnt]; };
        uint32_t lens = 0;
        for (uint32_t i0 = 0; i0 < cnt; i0 += WAVE) {
            const uint32_t i = i0 + lane;
            lens += i < cnt ? rec_at(i).y >> 16 : 0u;
        }
        lens = wave_sum(lens);
        uint32_t last = 0;
        if (cnt) {
            const uint2 r = rec_at(cnt - 1);
            last = r.x + (r.y >> 16);
        }
        if (lane == 0) {
            s_cnt[wave] = cnt, s_end[wave] = last, s_len[wave] = lens;
            if (wave == 0) s_out[0] = 0, s_out[1] = 0, s_last = 0;
        }
        __syncthreads();
        uint32_t base = 0, mbase = 0, pend = 0;  // sequences, matched bytes and the last end before this segment
        for (uint32_t j = 0; j < wave; j++) {
            base += s_cnt[j];
            mbase += s_len[j];
            pend = s_cnt[j] ? s_end[j] : pend;
        }
        uint32_t emitted = 0, covered = 0;
        uint8_t* lit_dst = literals ? literals + it.lit_pos : nullptr;
        const SrcView sv = src_view(arena, it.src_off, it.avail);
        gbytes src_bytes = sv.base + sv.mis;
        for (uint32_t i0 = 0; i0 < cnt && base + i0 < max_seq; i0 += WAVE) {
            const uint32_t i = i0 + lane;
            const bool valid = i < cnt && base + i < max_seq;
            const uint2 r = valid ? rec_at(i) : make_uint2(0u, 0u);
            const uint32_t len = r.y >> 16, off = r.y & 0xFFFFu;
            const uint32_t end = r.x + len;
            const uint32_t prev = wave_shr1(end, pend);
            const uint32_t mcum = wave_scan(valid ? len : 0u);  // this segment's matched bytes up to record i
            if (valid) {
                const uint64_t si = it.seq_pos + base + i;
                if (fmt == GC_SEQ_REF) {
                    gc_ref_sequence* o = (gc_ref_sequence*)seq_out + si;
                    o->literal_length = r.x - prev;
                    o->match_length_minus3 = len - 3;
                    o->offset = off;
                } else {
                    *(uint4*)((gc_zstd_sequence*)seq_out + si) = make_uint4(off, r.x - prev, len, 0);
                }
            }
            const uint32_t nv = (uint32_t)__builtin_popcountll(ballot(valid));
            if (lit_dst) {
                // record i's literals [prev, head) go where the bytes before prev that no
                // sequence covers end: prev - (matched bytes before record i)
                const uint32_t dst = prev - (mbase + covered + mcum - len);
                for (uint32_t l = 0; l < nv; l++) {
                    const uint32_t a = bcast(prev, l), h = bcast(r.x, l), d = bcast(dst, l);
                    if (h > a && d + (h - a) <= n) copy_wave(lit_dst + d, src_bytes + a, h - a);
                }
            }
            covered += bcast(mcum, nv - 1);
            emitted += nv;
            pend = bcast(end, nv - 1);  // (valid lanes are a prefix, lane 0 among them)
        }
        if (lane == 0 && emitted) {
            atomicAdd(&s_out[0], emitted);
            atomicAdd(&s_out[1], covered);
            atomicMax(&s_last, pend);
        }
        __syncthreads();
        const uint32_t nseq = s_out[0], matched = s_out[1], lend = s_last;
        if (threadIdx.x == 0) {
            seq_counts[it.count_pos] = nseq;
            if (lit_counts) lit_counts[it.count_pos] = n - matched;
        }
        // trailing literals: every byte after the last sequence written
        if (lit_dst && n > lend && lend >= matched) {
            uint8_t* d = lit_dst + (lend - matched);
            for (uint32_t i = threadIdx.x; i < n - lend; i += blockDim.x) d[i] = src_bytes[lend + i];
        }
        __syncthreads();  // LDS reused by the next sub-block
    }
}

#define GC_PARSE_SEG_KERNEL(NAME, LAZY)                                                                          \
    extern "C" __global__ void __launch_bounds__(GC_SEG_WAVES* WAVE, GC_SEG_WAVES / 4)                                          \
        NAME(const uint8_t* arena, const gc_item* __restrict__ items, uint32_t n_items, const uint32_t* info,   \
             void* seq_out, uint32_t seq_fmt, uint8_t* literals, uint32_t* seq_counts, uint32_t* lit_counts,    \
             uint32_t max_seq, uint2* scratch) {                                                                \
        zstd_parse_seg<LAZY>(arena, items, n_items, info, seq_out, seq_fmt, literals, seq_counts, lit_counts,   \
                             max_seq, scratch);                                                                 \
    }
GC_PARSE_SEG_KERNEL(gc_zstd_parse_seg_kernel, true)
GC_PARSE_SEG_KERNEL(gc_zstd_parse_seg1_kernel, false)

#endif  // GC_BUILD_PARSE

#if GC_BUILD_MAIN
// ---------------------------------------------------------------------------
// Zstd frames on the GPU (the capability of zstd_compress.cu:756-1038, where
// the coder builds per dimension an RLE or FSE_Compressed table from the
// sub-block's own code histogram, :529-740, 835-903), widened: per dimension
// the cheapest of the predefined, RLE and FSE_Compressed tables (estimated
// bits; an accuracy log from the sequence count), offsets as repeat codes
// where the format's offset history allows (RFC 8878 3.1.2.5), and literals
// Huffman-coded in four streams (weights written directly, bytes below 129)
// or RLE when smaller than raw; a raw block when nothing is smaller.
//
// One wavefront per sub-block.  Work that is small and serial -- ordering
// the literal bytes by count, the Huffman merge, code lengths cut to 11 bits,
// normalising a histogram, writing a table description, spreading a table's
// cells -- runs in one lane over the wavefront's LDS workspace, step for step
// as the oracle's coder (oracle/gc_oracle.c orc_zstd_frame) does it.  The
// parallel work is the wavefront's: histograms, ranks, literal streams and
// the sequences' bits, 64 at a time: every lane computes its field, a prefix
// sum over the lanes places it, the lanes OR their bits into an LDS buffer
// whose whole dwords then go to the frame.  Sequences are taken 64 at a time
// from the last: the FSE state chain (each state follows from the next
// sequence's) runs on the scalar unit with one LDS read per table and
// sequence; the offset history runs forward, once to count the offset codes
// (its state kept at every 64th sequence) and once per 64 from that state.
// ---------------------------------------------------------------------------
#define ZE_THREADS 256
#define ZE_WAVES (ZE_THREADS / WAVE)
// per-wavefront bitstream buffer: carried bits (< 32) + 64 fields of at most
// 74 bits (9 + 9 + 8 state bits, 16 + 16 + 16 extra bits), + the dwords past
// the last that a lane's 4-dword OR reaches
#define ZE_BITBUF 160
#define ZE_MAXCHUNK ((GC_SUB_BLOCK / 3 + WAVE - 1) / WAVE)  // 64-sequence chunks of a sub-block

__constant__ uint32_t ze_ll_base[36] = {0,  1,  2,   3,   4,   5,   6,    7,    8,    9,     10,    11,
                                        12, 13, 14,  15,  16,  18,  20,   22,   24,   28,    32,    40,
                                        48, 64, 128, 256, 512, 1024, 2048, 4096, 8192, 16384, 32768, 65536};
__constant__ uint8_t ze_ll_bits[36] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,  0,  1,  1,
                                       1, 1, 2, 2, 3, 3, 4, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16};
__constant__ uint32_t ze_ml_base[53] = {3,  4,  5,  6,  7,  8,   9,   10,  11,  12,   13,   14,   15,   16,
                                       17, 18, 19, 20, 21, 22,  23,  24,  25,  26,   27,   28,   29,   30,
                                       31, 32, 33, 34, 35, 37,  39,  41,  43,  47,   51,   59,   67,   83,
                                       99, 131, 259, 515, 1027, 2051, 4099, 8195, 16387, 32771, 65539};
__constant__ uint8_t ze_ml_bits[53] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,
                                       0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1,
                                       2, 2, 3, 3, 4, 4, 5, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16};
// predefined distributions, LL then ML then OF (-1: one cell, "less than 1")
__constant__ int8_t ze_norm[36 + 53 + 29] = {
    4, 3, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 1, 1, 1, 2, 2, 2, 2, 2, 2, 2, 2, 2, 3, 2, 1, 1, 1, 1, 1, -1, -1, -1, -1,
    1, 4, 3, 2, 2, 2, 2, 2, 2, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1,
    1, 1, 1, 1, 1, 1, 1, 1, 1, 1, -1, -1, -1, -1, -1, -1, -1,
    1, 1, 1, 1, 1, 1, 2, 2, 2, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, -1, -1, -1, -1, -1};

// Tables and symbol ranges: 0 LL (36 codes, table log <= 9), 1 ML (53, 9), 2 OF (32 slots, 29 predefined, 8).
#define ZE_SYM0(t) ((t) == 0 ? 0u : ((t) == 1 ? 36u : 89u))  // first per-symbol slot
#define ZE_ST0(t) ((t) == 0 ? 0u : ((t) == 1 ? 512u : 1024u))  // first next-state slot (dynamic tables)
#define ZE_NSYM 121u                                           // 36 + 53 + 32

// FSE encoding table seen from the encoder: st (next states of each
// symbol's cells in cell order, state = table size + cell), per symbol dnb
// (bits out = (state + dnb) >> 16) and dfs (first next-state index - the
// symbol's normalized count).
struct ZeShared {  // the predefined tables, built once per workgroup
    uint16_t st[64 + 64 + 32];
    uint32_t dnb[ZE_NSYM];
    int32_t dfs[ZE_NSYM];
};

struct ZeWave {  // one wavefront's workspace
    uint32_t bits[ZE_BITBUF];
    uint32_t reps[ZE_MAXCHUNK][3];  // the offset history at each 64th sequence
    uint2 sym[3][WAVE];             // per table, a chunk's sequences' (dnb, dfs)
    uint32_t ebits[3][WAVE];        // per table, their state bits (value | count << 12)
    uint32_t rq[4];                 // the offset history after a chunk
    union {
        struct {  // literals
            uint32_t hist[256];
            uint32_t sc[256];  // counts in (count, byte) order
            uint32_t nw[256];  // Huffman internal nodes' weights
            uint16_t srt[256];
            uint16_t code[256];
            uint8_t lp[256], np[256], dn[256], len[256];
        } h;
        struct {  // sequences
            uint32_t hist[ZE_NSYM];
            uint32_t dnb[ZE_NSYM];
            int32_t dfs[ZE_NSYM];
            int16_t nc[ZE_NSYM];
            uint16_t st[512 + 512 + 256];
            uint8_t cell[512];
            uint8_t desc[3][128];  // table descriptions as written
        } f;
    } u;
};

__device__ __forceinline__ uint32_t ze_highbit(uint32_t v) { return 31u - (uint32_t)__builtin_clz(v); }  // v >= 1

// 256 x log2(x), linear between powers of two (the cost estimates)
__device__ __forceinline__ uint32_t ze_log256(uint32_t x) {
    const uint32_t h = ze_highbit(x);
    return (h << 8) + (((x << 8) >> h) - 256u);
}

__device__ __forceinline__ uint32_t ze_code(const uint32_t* base, uint32_t ncode, uint32_t v) {
    uint32_t c = 0;
    for (uint32_t k = 1; k < ncode; k++) c = base[k] <= v ? k : c;
    return c;
}

// One table's encoder view from normalized counts nc[0, nsym) with accuracy
// log `log` (run by one thread): cells spread as the decoder spreads them
// ("less than 1" symbols at the top), next states in cell order per symbol.
// (Not inlined: the coding kernel then needs 139 vector registers, which
// leaves three waves per SIMD, as many as its LDS allows.)
__device__ __attribute__((noinline)) void ze_ctab_serial(const int16_t* nc, uint32_t nsym, uint32_t log, uint16_t* st, uint32_t* dnb,
                               int32_t* dfs, uint8_t* cell) {
    const uint32_t size = 1u << log, mask = size - 1;
    uint32_t high = size - 1;
    for (uint32_t s = 0; s < nsym; s++)
        if (nc[s] == -1) cell[high--] = (uint8_t)s;
    const uint32_t step = (size >> 1) + (size >> 3) + 3;
    uint32_t pos = 0;
    for (uint32_t s = 0; s < nsym; s++)
        for (int k = 0; k < nc[s]; k++) {
            cell[pos] = (uint8_t)s;
            do pos = (pos + step) & mask;
            while (pos > high);
        }
    // dfs first holds each symbol's first next-state index (a running cursor)
    int32_t total = 0;
    for (uint32_t s = 0; s < nsym; s++) {
        dfs[s] = total;
        total += nc[s] == -1 ? 1 : (nc[s] > 0 ? nc[s] : 0);
    }
    for (uint32_t u = 0; u < size; u++) st[dfs[cell[u]]++] = (uint16_t)(size + u);
    total = 0;
    for (uint32_t s = 0; s < nsym; s++) {
        if (nc[s] == 0) {
            dnb[s] = ((log + 1) << 16) - size, dfs[s] = 0;
        } else if (nc[s] == -1 || nc[s] == 1) {
            dnb[s] = (log << 16) - size, dfs[s] = total - 1, total += 1;
        } else {
            const uint32_t mb = log - ze_highbit((uint32_t)nc[s] - 1);
            dnb[s] = (mb << 16) - ((uint32_t)nc[s] << mb), dfs[s] = total - nc[s], total += nc[s];
        }
    }
}

__device__ void ze_build(ZeShared& s, int16_t* tmp_nc, uint8_t* tmp_cell) {
    const uint32_t tid = threadIdx.x;
    if (tid < 3) {  // one thread per table
        const uint32_t nsym = tid == 0 ? 36u : (tid == 1 ? 53u : 29u);
        const uint32_t n0 = tid == 0 ? 0u : (tid == 1 ? 36u : 89u);
        int16_t* nc = tmp_nc + ZE_SYM0(tid);
        for (uint32_t k = 0; k < nsym; k++) nc[k] = ze_norm[n0 + k];
        ze_ctab_serial(nc, nsym, tid == 2 ? 5u : 6u, s.st + (tid == 0 ? 0u : (tid == 1 ? 64u : 128u)),
                       s.dnb + ZE_SYM0(tid), s.dfs + ZE_SYM0(tid), tmp_cell + 64 * tid);
    }
    __syncthreads();
}

// Bit writer of one wavefront: each lane's field (lo: bits 0-63, hi: 64-95,
// nbit in all) is appended in lane order (rev: from the highest lane down);
// whole dwords go to dst from byte wpos.  Fields past `limit` set `over`.
struct ZeBits {
    uint32_t* buf;
    uint8_t* dst;
    uint32_t carry, wpos, limit;
    bool over;

    __device__ void put(uint64_t lo, uint32_t hi, uint32_t nbit, bool rev) {
        const uint32_t incl = wave_scan(nbit);
        const uint32_t batch = bcast(incl, WAVE - 1);
        const uint32_t at = carry + (rev ? batch - incl : incl - nbit);
        const uint32_t d = at >> 5, sh = at & 31;
        const uint32_t l0 = (uint32_t)lo, l1 = (uint32_t)(lo >> 32);
        const uint32_t w0 = l0 << sh;
        const uint32_t w1 = sh ? __builtin_amdgcn_alignbit(l1, l0, 32 - sh) : l1;
        const uint32_t w2 = sh ? __builtin_amdgcn_alignbit(hi, l1, 32 - sh) : hi;
        const uint32_t w3 = sh ? hi >> (32 - sh) : 0u;
        if (nbit) {
            atomicOr(&buf[d], w0);
            atomicOr(&buf[d + 1], w1);
            atomicOr(&buf[d + 2], w2);
            atomicOr(&buf[d + 3], w3);
        }
        __asm__ volatile("" ::: "memory");
        const uint32_t total = carry + batch, full = total >> 5;
        if (wpos + 4 * full >= limit) {
            over = true;
        } else {
            for (uint32_t k = lane_id(); k < full; k += WAVE) {
                const uint32_t v = buf[k];
                uint8_t* q = dst + wpos + 4 * k;
                q[0] = (uint8_t)v, q[1] = (uint8_t)(v >> 8), q[2] = (uint8_t)(v >> 16), q[3] = (uint8_t)(v >> 24);
            }
        }
        const uint32_t part = buf[full];
        __asm__ volatile("" ::: "memory");
        for (uint32_t k = lane_id(); k <= full + 3; k += WAVE) buf[k] = k == 0 ? part : 0u;
        __asm__ volatile("" ::: "memory");
        carry = total & 31;
        wpos += 4 * full;
    }
    // a field of lane 0 only (the other lanes add nothing)
    __device__ void put1(uint32_t v, uint32_t nb) {
        const bool me = lane_id() == 0;
        put(me ? (uint64_t)(nb < 32 ? v & ((1u << nb) - 1) : v) : 0ull, 0u, me ? nb : 0u, false);
    }
    // the end: the partial bits to the next byte boundary; returns the end position
    __device__ uint32_t finish() {
        const uint32_t tail = (carry + 7) / 8;
        if (wpos + tail >= limit) over = true;
        if (!over && lane_id() < tail) dst[wpos + lane_id()] = (uint8_t)(buf[lane_id() >> 2] >> (8 * (lane_id() & 3)));
        return wpos + tail;
    }
    __device__ void reset(uint32_t at) {
        for (uint32_t k = lane_id(); k < ZE_BITBUF; k += WAVE) buf[k] = 0;
        __asm__ volatile("" ::: "memory");
        carry = 0;
        wpos = at;
    }
};

// Huffman code lengths (at most 11 bits) of the present bytes, in one lane
// (w.h.sc and w.h.srt already hold the n present bytes by (count, byte));
// returns the longest.  As the oracle: two-queue merge, a leaf before a node
// of equal weight; depths over 11 cut to 11, the code made complete again
// one symbol at a time; lengths to the symbols in sorted order, the longest
// to the least frequent.
__device__ uint32_t ze_huff_lengths_serial(ZeWave& w, uint32_t n) {
    uint32_t li = 0, ni = 0, nn = 0;
    for (uint32_t k = 0; k + 1 < n; k++) {
        uint32_t wt = 0;
        for (int pick = 0; pick < 2; pick++) {
            if (li < n && (ni >= nn || w.u.h.sc[li] <= w.u.h.nw[ni])) {
                wt += w.u.h.sc[li];
                w.u.h.lp[li++] = (uint8_t)nn;
            } else {
                wt += w.u.h.nw[ni];
                w.u.h.np[ni++] = (uint8_t)nn;
            }
        }
        w.u.h.nw[nn++] = wt;
    }
    w.u.h.dn[nn - 1] = 0;
    for (int32_t j = (int32_t)nn - 2; j >= 0; j--) w.u.h.dn[j] = (uint8_t)min(w.u.h.dn[w.u.h.np[j]] + 1u, 200u);
    uint32_t blc[12];
    for (uint32_t l = 0; l < 12; l++) blc[l] = 0;
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t d = w.u.h.dn[w.u.h.lp[i]] + 1u;
        blc[d < 11 ? d : 11]++;
    }
    uint32_t k2 = 0;
    for (uint32_t l = 1; l <= 11; l++) k2 += blc[l] << (11 - l);
    while (k2 > 2048) {
        uint32_t l = 10;
        while (blc[l] == 0) l--;
        blc[l]--, blc[l + 1]++, k2 -= 1u << (10 - l);
    }
    while (k2 < 2048) {
        uint32_t l = 11;
        while (blc[l] == 0) l--;
        blc[l]--, blc[l - 1]++, k2 += 1u << (11 - l);
    }
    uint32_t maxlen = 11;
    while (blc[maxlen] == 0) maxlen--;
    uint32_t i = 0;
    for (uint32_t l = maxlen; l >= 1; l--)
        for (uint32_t c = 0; c < blc[l]; c++) w.u.h.len[w.u.h.srt[i++]] = (uint8_t)l;
    return maxlen;
}

// FSE table description of nc[0, nsym) (RFC 8878 4.1.1), in one lane, into
// out (byte-aligned); returns its bytes.
__device__ uint32_t ze_ncount_serial(const int16_t* nc, uint32_t nsym, uint32_t log, uint8_t* out) {
    uint64_t acc = log - 5;
    uint32_t bits = 4, o = 0;
    auto put = [&](uint32_t v, uint32_t nb) {
        acc |= (uint64_t)(v & ((1u << nb) - 1)) << bits;
        bits += nb;
        while (bits >= 8) out[o++] = (uint8_t)acc, acc >>= 8, bits -= 8;
    };
    int32_t remaining = (1 << log) + 1, threshold = 1 << log;
    uint32_t nb = log + 1, s = 0;
    bool prev0 = false;
    while (s < nsym && remaining > 1) {
        if (prev0) {
            uint32_t z = 0;
            while (s < nsym && nc[s] == 0) s++, z++;
            while (z >= 3) put(3, 2), z -= 3;
            put(z, 2);
        }
        const int32_t count = nc[s++];
        const int32_t max = (2 * threshold - 1) - remaining;
        remaining -= count < 0 ? -count : count;
        int32_t v = count + 1;
        if (v >= threshold) v += max;
        put((uint32_t)v, nb - (v < max ? 1u : 0u));
        prev0 = count == 0;
        while (remaining < threshold) nb--, threshold >>= 1;
    }
    if (bits) out[o++] = (uint8_t)acc;
    return o;
}

// Per-dimension choice for table t (in one lane): 0 predefined, 1 RLE
// (*rle = the symbol), 2 FSE_Compressed (normalized counts in w.u.f.nc,
// description in w.u.f.desc[t], *dlen bytes, accuracy log *log).
__device__ uint32_t ze_choose_serial(ZeWave& w, uint32_t t, uint32_t nseq, uint32_t* log, uint32_t* rle,
                                     uint32_t* dlen) {
    const uint32_t s0 = ZE_SYM0(t), nsym = t == 0 ? 36u : (t == 1 ? 53u : 29u);
    const uint32_t n0 = t == 0 ? 0u : (t == 1 ? 36u : 89u), plog = t == 2 ? 5u : 6u, maxlog = t == 2 ? 8u : 9u;
    const uint32_t* hist = w.u.f.hist + s0;
    int16_t* nc = w.u.f.nc + s0;
    uint32_t npresent = 0, maxs = 0;
    for (uint32_t s = 0; s < nsym; s++)
        if (hist[s]) npresent++, maxs = s;
    if (npresent == 1) {
        *rle = maxs;
        return 1;
    }
    uint64_t cp = 0;
    for (uint32_t s = 0; s <= maxs; s++)
        if (hist[s]) {
            const int32_t pn = ze_norm[n0 + s];
            cp += (uint64_t)hist[s] * ((plog << 8) - ze_log256(pn < 0 ? 1u : (uint32_t)pn));
        }
    uint32_t lg = ze_highbit(nseq);
    lg = lg > 7 ? lg - 2 : 5;
    const uint32_t need = ze_highbit(npresent - 1) + 2;
    if (lg < need) lg = need;
    if (lg > maxlog) lg = maxlog;
    // normalize (as the oracle: rounded shares, at least 1, the difference from the first largest count)
    const uint32_t size = 1u << lg;
    int32_t dist = 0;
    for (uint32_t s = 0; s <= maxs; s++) {
        if (!hist[s]) {
            nc[s] = 0;
            continue;
        }
        uint32_t v = (uint32_t)(((uint64_t)hist[s] * size + nseq / 2) / nseq);
        if (!v) v = 1;
        nc[s] = (int16_t)v;
        dist += (int32_t)v;
    }
    while (dist != (int32_t)size) {
        const bool up = dist < (int32_t)size;
        uint32_t best = 0, bc = 0;
        for (uint32_t s = 0; s <= maxs; s++)
            if (hist[s] > bc && (up || nc[s] > 1)) bc = hist[s], best = s;
        nc[best] += up ? 1 : -1;
        dist += up ? 1 : -1;
    }
    for (uint32_t s = maxs + 1; s < nsym; s++) nc[s] = 0;
    *dlen = ze_ncount_serial(nc, maxs + 1, lg, w.u.f.desc[t]);
    uint64_t cc = (uint64_t)*dlen * 8 * 256;
    for (uint32_t s = 0; s <= maxs; s++)
        if (hist[s]) cc += (uint64_t)hist[s] * ((lg << 8) - ze_log256((uint32_t)nc[s]));
    *log = lg;
    return cc < cp ? 2u : 0u;
}

// The offset history (RFC 8878 3.1.2.5) over lane j's sequence (off, ll),
// for j < cnt in order: lane j gets its Offset_Value (1-3: repeat offset 1-3,
// with ll == 0 meaning r1, r2, r0 - 1; else off + 3); r0..r2 (uniform) end as
// the history after the chunk.  The new first offset is always the
// sequence's own; the other two follow from two flags (a repeat of r0 with
// literals keeps the history, a repeat of r1 swaps the first two), written
// without branches: the scalar unit selects.
// Without that one exception (a sequence with no literals whose offset is
// the previous sequence's: the history then holds that offset twice), the
// history is the three most recent distinct offsets, most recent first, and
// the first is always the previous sequence's own.  Such lists of the
// sequences before each lane come from a prefix scan over the lanes (merge:
// the later list, then the earlier list's offsets not in it), six steps of
// ds_bpermute; a chunk with the exception, or a history holding an offset
// twice, is walked one sequence at a time in lane 0.
__device__ __forceinline__ void ze_mru_merge(uint32_t& o0, uint32_t& o1, uint32_t& o2, uint32_t a0, uint32_t a1,
                                             uint32_t a2) {
    // o (later, o0 != 0) then a (earlier); 0 = no offset
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const uint32_t a = k == 0 ? a0 : (k == 1 ? a1 : a2);
        const bool add = a != 0 && a != o0 && a != o1 && a != o2;
        const bool at1 = add && o1 == 0, at2 = add && o1 != 0 && o2 == 0;
        o1 = at1 ? a : o1;
        o2 = at2 ? a : o2;
    }
}

__device__ __forceinline__ uint32_t ze_rep_value(uint32_t o, bool lit, uint32_t a, uint32_t b, uint32_t c) {
    uint32_t v = o + 3;
    v = o == (lit ? c : a - 1) ? 3u : v;
    v = o == (lit ? b : c) ? 2u : v;
    v = o == (lit ? a : b) ? 1u : v;
    return v;
}

__device__ __forceinline__ uint32_t ze_rep_chunk(ZeWave& w, uint32_t off, uint32_t ll, uint32_t cnt, uint32_t& r0,
                                                 uint32_t& r1, uint32_t& r2) {
    const uint32_t lane = lane_id();
    const bool valid = lane < cnt;
    // the exception: no literals and the previous offset (r0 for lane 0)
    const uint32_t prev = wave_shr1(off, r0);
    // (an offset of 0, which no sequence should have, would read as "no
    // offset" in the lists: walked one at a time too)
    const bool odd = (valid && ((ll == 0 && off == prev) || off == 0)) || r0 == r1 || r1 == r2 || r0 == r2;
    if (ballot(odd) == 0) {
        // inclusive scan: lane i's list of the distinct offsets of lanes 0..i
        uint32_t h0 = valid ? off : 0u, h1 = 0, h2 = 0;
#pragma unroll
        for (uint32_t d = 1; d < WAVE; d <<= 1) {
            const int src = (int)((lane >= d ? lane - d : lane) << 2);
            const uint32_t a0 = (uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)h0);
            const uint32_t a1 = (uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)h1);
            const uint32_t a2 = (uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)h2);
            if (lane >= d) ze_mru_merge(h0, h1, h2, a0, a1, a2);
        }
        // then the history before the chunk, and the history before each lane
        if (valid) ze_mru_merge(h0, h1, h2, r0, r1, r2);
        const uint32_t a = wave_shr1(h0, r0), b = wave_shr1(h1, r1), c = wave_shr1(h2, r2);
        const uint32_t v = ze_rep_value(off, ll > 0, a, b, c);
        r0 = bcast(h0, cnt - 1), r1 = bcast(h1, cnt - 1), r2 = bcast(h2, cnt - 1);
        return v;
    }
    // one sequence at a time: the chunk's (off, ll) staged in LDS, walked by
    // lane 0 (the chains' buffers are free until the chains run)
    uint2* q = (uint2*)&w.ebits[1][0];
    uint32_t* res = w.ebits[0];
    q[lane] = make_uint2(off, ll);
    __asm__ volatile("" ::: "memory");  // (one wavefront: its LDS operations run in order)
    if (lane == 0) {
        uint32_t a = r0, b = r1, c = r2;
        for (uint32_t j = 0; j < cnt; j++) {
            const uint2 x = q[j];
            const uint32_t o = x.x;
            const bool lit = x.y > 0;
            res[j] = ze_rep_value(o, lit, a, b, c);
            const bool keep = lit && o == a, swap = !keep && o == b;
            const uint32_t n1 = keep ? b : a, n2 = (keep || swap) ? c : b;
            a = o, b = n1, c = n2;
        }
        w.rq[0] = a, w.rq[1] = b, w.rq[2] = c;
    }
    __asm__ volatile("" ::: "memory");
    r0 = uniform(w.rq[0]), r1 = uniform(w.rq[1]), r2 = uniform(w.rq[2]);
    return res[lane];
}

template <uint32_t FMT>
__device__ __forceinline__ void ze_load_seq(const void* seqs, uint64_t r, uint32_t& off, uint32_t& ll, uint32_t& ml) {
    if constexpr (FMT == GC_SEQ_REF) {
        const gc_ref_sequence q = ((const gc_ref_sequence*)seqs)[r];
        ll = q.literal_length, ml = q.match_length_minus3 + 3, off = q.offset;
    } else {
        const uint4 q = ((const uint4*)seqs)[r];
        off = q.x, ll = q.y, ml = q.z;
    }
}

__device__ __forceinline__ void ze_copy_bytes(uint8_t* dst, const uint8_t* src, uint32_t len) {
    const uint32_t lane = lane_id();
    uint32_t k = lane;
    for (; k + 3 * WAVE < len; k += 4 * WAVE) {  // four loads in flight per lane
        const uint8_t a = src[k], b = src[k + WAVE], c = src[k + 2 * WAVE], d = src[k + 3 * WAVE];
        dst[k] = a, dst[k + WAVE] = b, dst[k + 2 * WAVE] = c, dst[k + 3 * WAVE] = d;
    }
    for (; k < len; k += WAVE) dst[k] = src[k];
}

// The literals section at dst[o, limit) for lit[0, nlit): returns its end, or
// `limit` when it does not fit (the block then goes raw).  Not inlined (as
// ze_ctab_serial: fewer vector registers in the kernel).
__device__ __attribute__((noinline)) uint32_t ze_literals(ZeWave& w, ZeBits& bw, uint8_t* dst, uint32_t o, uint32_t limit, const uint8_t* lit,
                                uint32_t nlit) {
    const uint32_t lane = lane_id();
    const uint32_t rh = nlit < 32 ? 1u : (nlit < 4096 ? 2u : 3u);
    for (uint32_t k = lane; k < 256; k += WAVE) w.u.h.hist[k] = 0;
    __asm__ volatile("" ::: "memory");
    for (uint32_t k = lane; k < nlit; k += WAVE) atomicAdd(&w.u.h.hist[lit[k]], 1u);
    __asm__ volatile("" ::: "memory");
    uint32_t nz = 0, maxs = 0;
    for (uint32_t k0 = 0; k0 < 256; k0 += WAVE) {
        const uint64_t m = ballot(w.u.h.hist[k0 + lane] != 0);
        nz += (uint32_t)__builtin_popcountll(m);
        if (m) maxs = k0 + 63u - (uint32_t)__builtin_clzll(m);
    }
    uint32_t type = (nlit > 0 && nz == 1) ? 1u : 0u;
    if (nlit >= 64 && nz >= 2 && maxs <= 128) {
        for (uint32_t k = lane; k < 256; k += WAVE) w.u.h.len[k] = 0;
        // rank of every present byte by (count, byte): lane l ranks bytes l,
        // l + 64, l + 128, l + 192, reading each count once (a count read per
        // byte pair in the inner loop, hoisted out of an outer loop over the
        // lane's bytes, held all 256 counts in registers)
        uint32_t hs[4], rank[4];
#pragma unroll
        for (uint32_t q = 0; q < 4; q++) hs[q] = w.u.h.hist[q * WAVE + lane], rank[q] = 0;
#pragma unroll 4
        for (uint32_t t = 0; t < 256; t++) {
            const uint32_t ht = w.u.h.hist[t];
#pragma unroll
            for (uint32_t q = 0; q < 4; q++)
                rank[q] += (ht != 0 && (ht < hs[q] || (ht == hs[q] && t < q * WAVE + lane))) ? 1u : 0u;
        }
#pragma unroll
        for (uint32_t q = 0; q < 4; q++)
            if (hs[q]) {
                w.u.h.srt[rank[q]] = (uint16_t)(q * WAVE + lane);
                w.u.h.sc[rank[q]] = hs[q];
            }
        __asm__ volatile("" ::: "memory");
        uint32_t maxlen = 0;
        if (lane == 0) maxlen = ze_huff_lengths_serial(w, nz);
        maxlen = uniform(maxlen);
        __asm__ volatile("" ::: "memory");
        // canonical codes (the decoder's: weights from 1 up take consecutive
        // table ranges, bytes in order within a weight)
        if (lane == 0) {
            uint32_t start[13], cntw[13];
            for (uint32_t k = 0; k < 13; k++) cntw[k] = 0;
            for (uint32_t s = 0; s < 256; s++)
                if (w.u.h.len[s]) cntw[maxlen + 1 - w.u.h.len[s]]++;
            uint32_t pos = 0;
            for (uint32_t k = 1; k <= maxlen; k++) start[k] = pos, pos += cntw[k] << (k - 1);
            for (uint32_t s = 0; s < 256; s++) {
                const uint32_t l = w.u.h.len[s];
                if (l) {
                    const uint32_t wt = maxlen + 1 - l;
                    w.u.h.code[s] = (uint16_t)(start[wt] >> (wt - 1));
                    start[wt] += 1u << (wt - 1);
                }
            }
        }
        __asm__ volatile("" ::: "memory");
        // stream sizes first: whether the section is smaller than raw, and its header size
        const uint32_t seg = (nlit + 3) / 4;
        uint32_t sbits[4];
        for (uint32_t k = 0; k < 4; k++) {
            const uint32_t a = k * seg, e = k == 3 ? nlit : (k + 1) * seg;
            uint32_t b = 0;
            for (uint32_t i = a + lane; i < e; i += WAVE) b += w.u.h.len[lit[i]];
            sbits[k] = wave_sum(b) + 1;  // + the end mark
        }
        const uint32_t tree = 1 + (maxs + 1) / 2;
        uint32_t cs = tree + 6;
        for (uint32_t k = 0; k < 4; k++) cs += (sbits[k] + 7) / 8;
        const uint32_t hh = (nlit <= 1023 && cs <= 1023) ? 3u : ((nlit <= 16383 && cs <= 16383) ? 4u : 5u);
        if (hh + cs < rh + nlit) {
            if (o + hh + cs >= limit) return limit;
            const uint64_t h = hh == 3 ? (2ull | (1ull << 2) | ((uint64_t)nlit << 4) | ((uint64_t)cs << 14))
                                       : (hh == 4 ? (2ull | (2ull << 2) | ((uint64_t)nlit << 4) | ((uint64_t)cs << 18))
                                                  : (2ull | (3ull << 2) | ((uint64_t)nlit << 4) | ((uint64_t)cs << 22)));
            if (lane < hh) dst[o + lane] = (uint8_t)(h >> (8 * lane));
            o += hh;
            // weights of bytes 0 .. maxs - 1, 4 bits each, the first in the high nibble
            if (lane == 0) dst[o] = (uint8_t)(127 + maxs);
            for (uint32_t s = 2 * lane; s < maxs; s += 2 * WAVE) {
                const uint32_t l0 = w.u.h.len[s], l1 = s + 1 < maxs ? w.u.h.len[s + 1] : 0u;
                const uint32_t w0 = l0 ? maxlen + 1 - l0 : 0u, w1 = l1 ? maxlen + 1 - l1 : 0u;
                dst[o + 1 + s / 2] = (uint8_t)((w0 << 4) | w1);
            }
            o += tree;
            const uint32_t jt = o;
            o += 6;
            for (uint32_t k = 0; k < 4; k++) {
                const uint32_t a = k * seg, e = k == 3 ? nlit : (k + 1) * seg;
                bw.reset(o);
                // the stream codes the bytes from the last one back (the decoder reads it backward)
                for (uint32_t q0 = 0; q0 < e - a; q0 += WAVE) {
                    const uint32_t q = q0 + lane;
                    uint32_t v = 0, nb = 0;
                    if (q < e - a) {
                        const uint32_t b = lit[e - 1 - q];
                        v = w.u.h.code[b], nb = w.u.h.len[b];
                    }
                    bw.put(v, 0u, nb, false);
                }
                bw.put1(1, 1);  // end mark
                const uint32_t end = bw.finish();
                if (k < 3 && lane < 2) dst[jt + 2 * k + lane] = (uint8_t)((end - o) >> (8 * lane));
                o = end;
            }
            return o;
        }
    }
    const uint32_t body = type == 1 ? 1u : nlit;
    if (o + rh + body >= limit) return limit;
    const uint32_t v = rh == 1 ? (type | (nlit << 3)) : (rh == 2 ? (type | 4u | (nlit << 4)) : (type | 12u | (nlit << 4)));
    if (lane < rh) dst[o + lane] = (uint8_t)(v >> (8 * lane));
    if (type == 1) {
        if (lane == 0) dst[o + rh] = lit[0];
    } else {
        ze_copy_bytes(dst + o + rh, lit, nlit);
    }
    return o + rh + body;
}

template <uint32_t FMT>
__device__ void ze_frame(const ZeShared& sh, ZeWave& w, const gc_item& it, const uint8_t* arena, const void* seqs,
                         const uint8_t* literals, const uint32_t* seq_counts, const uint32_t* lit_counts, uint8_t* out,
                         uint32_t* sizes, uint32_t max_frame, uint32_t max_seq) {
    const uint32_t lane = lane_id();
    const uint32_t n = it.n;
    const uint32_t nseq = min(uniform(seq_counts[it.count_pos]), max_seq);  // reads stay in the item's slots
    const uint32_t nlit = uniform(lit_counts[it.count_pos]);
    uint8_t* dst = out + (uint64_t)it.count_pos * max_frame;
    // frame header: magic, descriptor (single segment), content size
    const uint32_t hdr = n >= 256 ? 7u : 6u;
    if (lane < hdr) {
        const uint32_t fcs = n >= 256 ? n - 256 : n;
        const uint32_t b = lane < 4 ? ((0xFD2FB528u >> (8 * lane)) & 0xFF)
                                    : (lane == 4 ? (n >= 256 ? 0x60u : 0x20u) : ((fcs >> (8 * (lane - 5))) & 0xFF));
        dst[lane] = (uint8_t)b;
    }
    const uint32_t start = hdr + 3;    // after the block header
    const uint32_t limit = start + n;  // a compressed block must end before this
    // (raw also when the counts cannot be a sub-block's: literals alone would not be smaller, or more
    // sequences than 3-byte matches fit)
    bool raw = nlit + 2 >= n || nseq > ZE_MAXCHUNK * WAVE;
    uint32_t end = 0;
    ZeBits bw{w.bits, dst, 0, 0, limit, false};
    if (!raw) {
        uint32_t o = ze_literals(w, bw, dst, start, limit, literals + it.lit_pos, nlit);
        raw = bw.over || o + 4 > limit;
        if (!raw) {
            const uint32_t shd = nseq < 128 ? 1u : 2u;
            if (lane < shd) {
                const uint32_t v = shd == 1 ? nseq : (lane == 0 ? (nseq >> 8) + 0x80 : nseq & 0xFF);
                dst[o + lane] = (uint8_t)v;
            }
            o += shd;
            end = o;
        }
        if (!raw && nseq) {
            // forward: the offset history (kept at every 64th sequence) and the codes' histograms
            for (uint32_t k = lane; k < ZE_NSYM; k += WAVE) w.u.f.hist[k] = 0;
            __asm__ volatile("" ::: "memory");
            uint32_t r0 = 1, r1 = 4, r2 = 8;
            for (uint32_t i0 = 0; i0 < nseq; i0 += WAVE) {
                const uint32_t cnt = min(nseq - i0, (uint32_t)WAVE);
                if (lane == 0) w.reps[i0 / WAVE][0] = r0, w.reps[i0 / WAVE][1] = r1, w.reps[i0 / WAVE][2] = r2;
                uint32_t off = 0, ll = 0, ml = 3;
                if (lane < cnt) ze_load_seq<FMT>(seqs, it.seq_pos + i0 + lane, off, ll, ml);
                const uint32_t ov = ze_rep_chunk(w, off, ll, cnt, r0, r1, r2);
                if (lane < cnt) {
                    atomicAdd(&w.u.f.hist[ze_code(ze_ll_base, 36, ll)], 1u);
                    atomicAdd(&w.u.f.hist[36 + ze_code(ze_ml_base, 53, ml)], 1u);
                    atomicAdd(&w.u.f.hist[89 + ze_highbit(ov)], 1u);
                }
            }
            __asm__ volatile("" ::: "memory");
            // per-dimension modes and tables (one lane; the descriptions into LDS)
            uint32_t mode[3] = {0, 0, 0}, lg[3] = {6, 6, 5}, rs[3] = {0, 0, 0}, dl[3] = {0, 0, 0};
            if (lane == 0) {
                for (uint32_t t = 0; t < 3; t++) {
                    mode[t] = ze_choose_serial(w, t, nseq, &lg[t], &rs[t], &dl[t]);
                    if (mode[t] == 2)
                        ze_ctab_serial(w.u.f.nc + ZE_SYM0(t), t == 0 ? 36u : (t == 1 ? 53u : 29u), lg[t],
                                       w.u.f.st + ZE_ST0(t), w.u.f.dnb + ZE_SYM0(t), w.u.f.dfs + ZE_SYM0(t),
                                       w.u.f.cell);
                }
            }
            for (uint32_t t = 0; t < 3; t++) {
                mode[t] = uniform(mode[t]), rs[t] = uniform(rs[t]), dl[t] = uniform(dl[t]);
                lg[t] = mode[t] == 2 ? uniform(lg[t]) : (t == 2 ? 5u : 6u);
            }
            __asm__ volatile("" ::: "memory");
            // modes byte, then LL, OF, ML: the RLE byte or the table description
            if (lane == 0) dst[o] = (uint8_t)((mode[0] << 6) | (mode[2] << 4) | (mode[1] << 2));
            o += 1;
            const uint32_t order[3] = {0, 2, 1};
            for (uint32_t q = 0; q < 3; q++) {
                const uint32_t t = order[q];
                if (mode[t] == 1) {
                    if (lane == 0 && o < limit) dst[o] = (uint8_t)rs[t];
                    o += 1;
                } else if (mode[t] == 2) {
                    if (o + dl[t] < limit)
                        for (uint32_t k = lane; k < dl[t]; k += WAVE) dst[o + k] = w.u.f.desc[t][k];
                    o += dl[t];
                }
            }
            // each dimension's table: its own or the predefined one (LDS offsets)
            const uint16_t* stp[3];
            const uint32_t* dnbp[3];
            const int32_t* dfsp[3];
            for (uint32_t t = 0; t < 3; t++) {
                const bool own = mode[t] == 2;
                stp[t] = own ? w.u.f.st + ZE_ST0(t) : sh.st + (t == 0 ? 0u : (t == 1 ? 64u : 128u));
                dnbp[t] = own ? w.u.f.dnb + ZE_SYM0(t) : sh.dnb + ZE_SYM0(t);
                dfsp[t] = own ? w.u.f.dfs + ZE_SYM0(t) : sh.dfs + ZE_SYM0(t);
            }
            // the bitstream, 64 sequences at a time from the last.  The three FSE
            // state chains (each state follows from the next sequence's) run in
            // lanes 0-2, table t in lane t, with its sequences' (dnb, dfs) staged
            // in LDS by every lane: one LDS read of the next-state table per
            // sequence and lane, the three chains side by side.
            bw.reset(o);
            const uint16_t* mst = lane == 0 ? stp[0] : (lane == 1 ? stp[1] : stp[2]);
            const bool mchain = lane < 3 && mode[lane < 3 ? lane : 0] != 1;
            uint32_t cst = 0;  // lane t: table t's state
            const uint32_t nchunk = (nseq + WAVE - 1) / WAVE;
            for (int32_t c = (int32_t)nchunk - 1; c >= 0 && !raw && !bw.over; c--) {
                const uint32_t i0 = (uint32_t)c * WAVE, cnt = min(nseq - i0, (uint32_t)WAVE);
                uint32_t off = 0, ll = 0, ml = 3;
                if (lane < cnt) ze_load_seq<FMT>(seqs, it.seq_pos + i0 + lane, off, ll, ml);
                uint32_t q0 = w.reps[c][0], q1 = w.reps[c][1], q2 = w.reps[c][2];
                q0 = uniform(q0), q1 = uniform(q1), q2 = uniform(q2);
                const uint32_t ov = ze_rep_chunk(w, off, ll, cnt, q0, q1, q2);
                const uint32_t lc = ze_code(ze_ll_base, 36, ll), mc = ze_code(ze_ml_base, 53, ml);
                const uint32_t oc = lane < cnt ? ze_highbit(ov) : 0u;
                // each lane's symbols' dnb and dfs, read in parallel before the chains
                w.sym[0][lane] = make_uint2(dnbp[0][lc], (uint32_t)dfsp[0][lc]);
                w.sym[1][lane] = make_uint2(dnbp[1][mc], (uint32_t)dfsp[1][mc]);
                w.sym[2][lane] = make_uint2(dnbp[2][oc], (uint32_t)dfsp[2][oc]);
                __asm__ volatile("" ::: "memory");  // (one wavefront: its LDS operations run in order)
                if (mchain) {
                    uint32_t st = cst;
                    for (int32_t j = (int32_t)cnt - 1; j >= 0; j--) {
                        const uint2 sd = w.sym[lane][j];
                        uint32_t e = 0;
                        if (i0 + (uint32_t)j == nseq - 1) {  // the last sequence: the initial state, no bits
                            const uint32_t nb = (sd.x + (1u << 15)) >> 16;
                            st = mst[(int32_t)(((nb << 16) - sd.x) >> nb) + (int32_t)sd.y];
                        } else {
                            const uint32_t nb = (st + sd.x) >> 16;
                            e = (st & ((1u << nb) - 1)) | (nb << 12);
                            st = mst[(int32_t)(st >> nb) + (int32_t)sd.y];
                        }
                        w.ebits[lane][j] = e;
                    }
                    cst = st;
                }
                __asm__ volatile("" ::: "memory");
                // lane j's state bits (value | count << 12) per table; none for an RLE table
                const uint32_t bl = mode[0] != 1 ? w.ebits[0][lane] : 0u;
                const uint32_t bm = mode[1] != 1 ? w.ebits[1][lane] : 0u;
                const uint32_t bo = mode[2] != 1 ? w.ebits[2][lane] : 0u;
                // this lane's fields in write order: state bits OF, ML, LL, extra bits LL, ML, OF
                uint64_t lo = 0;
                uint32_t hi = 0, nbit = 0;
                auto field = [&](uint32_t v, uint32_t k) {
                    if (nbit < 64) lo |= (uint64_t)v << nbit;
                    if (nbit + k > 64) hi |= nbit >= 64 ? v << (nbit - 64) : v >> (64 - nbit);
                    nbit += k;
                };
                if (lane < cnt) {
                    field(bo & 0xFFFu, bo >> 12);
                    field(bm & 0xFFFu, bm >> 12);
                    field(bl & 0xFFFu, bl >> 12);
                    field(ll - ze_ll_base[lc], ze_ll_bits[lc]);
                    field(ml - ze_ml_base[mc], ze_ml_bits[mc]);
                    field(ov - (1u << oc), oc);
                }
                bw.put(lo, hi, nbit, true);  // the last sequence first
            }
            if (!raw && !bw.over) {
                // initial states, read as LL, OF, ML; then the end mark
                const uint32_t sl = bcast(cst, 0), sm = bcast(cst, 1), so = bcast(cst, 2);
                if (mode[1] != 1) bw.put1(sm & ((1u << lg[1]) - 1), lg[1]);
                if (mode[2] != 1) bw.put1(so & ((1u << lg[2]) - 1), lg[2]);
                if (mode[0] != 1) bw.put1(sl & ((1u << lg[0]) - 1), lg[0]);
                bw.put1(1, 1);
                end = bw.finish();
            }
            if (bw.over || end >= limit) raw = true;
        }
        if (!raw && end >= limit) raw = true;
        if (raw) __threadfence();  // the raw block below overwrites the bytes written so far
    }
    if (raw) {
        ze_copy_bytes(dst + start, at(arena, it.src_off), n);
        end = start + n;
    }
    if (lane < 3) {
        const uint32_t bh = raw ? (1u | (n << 3)) : (1u | (2u << 1) | ((end - start) << 3));
        dst[hdr + lane] = (uint8_t)(bh >> (8 * lane));
    }
    if (lane == 0) sizes[it.count_pos] = end;
}

// (scalar registers capped at 96: above that the hardware runs fewer waves per
// SIMD than the compiler counts)
extern "C" __global__ void __launch_bounds__(ZE_THREADS) __attribute__((amdgpu_num_sgpr(96)))
gc_zstd_encode_kernel(const uint8_t* arena, const gc_item* __restrict__ items, uint32_t n_items, const void* seqs,
                      uint32_t seq_fmt, const uint8_t* literals, const uint32_t* seq_counts,
                      const uint32_t* lit_counts, uint8_t* out, uint32_t* sizes, uint32_t max_frame,
                      uint32_t max_seq, const uint32_t* order) {
    __shared__ ZeShared sh;
    __shared__ ZeWave ws[ZE_WAVES];
    // the predefined tables, built with the first wavefront's workspace as scratch
    ze_build(sh, ws[0].u.f.nc, ws[0].u.f.cell);
    const uint32_t wave = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
    // (order, optional: the items with the most sequences first, so that the
    // longest frames do not start last)
    for (uint32_t i = blockIdx.x * ZE_WAVES + wave; i < n_items; i += gridDim.x * ZE_WAVES) {
        const gc_item it = load_item(items, order ? uniform(order[i]) : i);
        if (seq_fmt == GC_SEQ_REF)
            ze_frame<GC_SEQ_REF>(sh, ws[wave], it, arena, seqs, literals, seq_counts, lit_counts, out, sizes,
                                 max_frame, max_seq);
        else
            ze_frame<GC_SEQ_ZSTD>(sh, ws[wave], it, arena, seqs, literals, seq_counts, lit_counts, out, sizes,
                                  max_frame, max_seq);
    }
}

// Packs frames: out[offsets[s] + i] = in[s * slot + i] for i < sizes[s].
extern "C" __global__ void __launch_bounds__(256)
gc_gather_bytes(const uint8_t* in, const uint32_t* sizes, const uint64_t* offsets, uint8_t* out, uint32_t slot,
                uint32_t nsub) {
    for (uint32_t s = blockIdx.x; s < nsub; s += gridDim.x) {
        const uint32_t c = sizes[s];
        const uint8_t* src = in + (uint64_t)s * slot;
        uint8_t* dst = out + offsets[s];
        for (uint32_t i = threadIdx.x; i < c; i += blockDim.x) dst[i] = src[i];
    }
}

// Work items of a buffer tiled in sub-blocks of `sub` bytes from its start,
// scratch positions at a pitch of sub rounded up to 8 (the host's make_items
// for the kernel-level Zstd entry point, built on the device so that a call
// needs no host-to-device copy).
extern "C" __global__ void __launch_bounds__(256)
gc_make_items(gc_item* items, uint64_t total, uint32_t sub, uint32_t nsub, uint32_t max_seq, uint32_t depth) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsub) return;
    const uint64_t start = (uint64_t)s * sub;
    const uint64_t pitch = ((uint64_t)sub + 7) & ~(uint64_t)7;
    gc_item it;
    it.src_off = start;
    it.avail = total - start;
    it.cpos = (uint64_t)s * pitch;
    it.opos = (uint64_t)s * pitch;
    it.seq_pos = (uint64_t)s * max_seq;
    it.lit_pos = start;
    it.count_pos = s;
    it.n = (uint32_t)(total - start < sub ? total - start : sub);
    it.depth = depth;
    it.out_span = sub;
    items[s] = it;
}

// Served form: a tick's work list, written by the host into pinned memory
// (hdr: item counts and first items per kernel; src: the items), copied into
// the slot's device queue and item array by the first kernel of the slot's
// captured pass.  The work counters of q are not touched (they are zero
// between launches).  A count above the slot's capacity sets error bit 2 and
// runs nothing.
extern "C" __global__ void __launch_bounds__(256)
gc_tick_setup(const gc_queue* hdr, const gc_item* src, gc_queue* q, gc_item* items, uint32_t cap) {
    const uint32_t n_all = hdr->n[GC_Q_CAND];
    const bool ok = n_all <= cap;
    if (blockIdx.x == 0 && threadIdx.x < 4) {
        q->n[threadIdx.x] = ok ? hdr->n[threadIdx.x] : 0u;
        q->base[threadIdx.x] = hdr->base[threadIdx.x];
        if (!ok && threadIdx.x == 0) atomicOr(&q->err, 4u);
    }
    if (!ok) return;
    const uint4* s4 = (const uint4*)src;
    uint4* d4 = (uint4*)items;
    const uint32_t words = n_all * (uint32_t)(sizeof(gc_item) / 16);
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < words; k += gridDim.x * blockDim.x) d4[k] = s4[k];
}

// Small fill kernel.
extern "C" __global__ void gc_fill_u32(uint32_t* p, uint32_t v, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// n u32 from src to dst (dst may be pinned host memory: the copy then runs
// on the CUs over PCIe instead of on a DMA engine); vec4: both 16-byte
// aligned.
extern "C" __global__ void __launch_bounds__(256) gc_copy_u32(const uint32_t* src, uint32_t* dst, uint64_t n,
                                                               uint32_t vec4) {
    const uint64_t tid = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    uint64_t done = 0;
    if (vec4) {
        const uint64_t n4 = n / 4;
        for (uint64_t i = tid; i < n4; i += stride) ((uint4*)dst)[i] = ((const uint4*)src)[i];
        done = n4 * 4;
    }
    for (uint64_t i = done + tid; i < n; i += stride) dst[i] = src[i];
}

// Packs the used sequence slots of each sub-block back to back:
// out[offsets[s] + i] = in[s * max_seq + i] for i < counts[s].
extern "C" __global__ void __launch_bounds__(256)
gc_gather_seqs(const uint4* in, const uint32_t* counts, const uint64_t* offsets, uint4* out, uint32_t max_seq,
               uint32_t nsub) {
    for (uint32_t s = blockIdx.x; s < nsub; s += gridDim.x) {
        const uint32_t c = counts[s];
        const uint4* src = in + (uint64_t)s * max_seq;
        uint4* dst = out + offsets[s];
        for (uint32_t i = threadIdx.x; i < c; i += blockDim.x) dst[i] = src[i];
    }
}
#endif  // GC_BUILD_MAIN

#if GC_BUILD_MAIN
// ---------------------------------------------------------------------------
// LZMA2 chunks on the GPU: the coder of gc_lzma_enc.cpp (one LZMA chunk per
// sub-block with properties lc 3, lp 0, pb 2; greedy parse with one lazy step
// over the match finder's packed matches and the four repeat distances;
// compress_lzma2_custom.rs:344-465 and :601-664 for what it replaces),
// producing the same bytes.  One wavefront per sub-block, its probability
// model (7319 11-bit probabilities, 14.6 KB) in LDS: eleven wavefronts per CU.
//
// The parse walks 64-position windows.  Entering a window at p loads, one
// position per lane, the byte at p + lane, the byte one repeat distance back
// for each of the four distances and the packed match; a ballot per distance
// is a 64-bit mask of the positions whose byte differs from the byte that
// distance back (or that lie before the distance or past the sub-block), so
// the repeat lengths at every position of the window are counted from the
// masks; a length that reaches the window's end goes on 64 bytes per step.
// Literals and repeat-0 matches keep the distances, so the window serves the
// positions after them; a match or another repeat moves the distances and
// the next window starts where that symbol ends.
//
// A symbol is coded as a list of at most 32 binary decisions, one per lane
// (a probability index and the bit, or a direct bit): every lane reads its
// probability and writes back its update at once (the decisions of one
// symbol use distinct probabilities), then the range coder takes the
// decisions in order with the wavefront's uniform state.  Finished bytes are
// gathered four to a word, the words of a 256-byte block in a vector register
// (one per lane), which is written out when the block is whole.
//
// Output: slot s (LE_SLOT bytes, item count_pos s) holds the chunk from byte
// LE_HEAD (6 header bytes, the coded data from byte LE_HEAD + 6 = 8 so that
// the buffer's dword stores are aligned, the end marker), or the raw chunk
// (3 header bytes, the bytes, the end marker) when the coded chunk would not
// be smaller; sizes[s] is the chunk's length.
// ---------------------------------------------------------------------------
#define LE_IS_MATCH 0u     // [state][pos state]
#define LE_IS_REP 48u      // [state]
#define LE_IS_REP_G0 60u
#define LE_IS_REP_G1 72u
#define LE_IS_REP_G2 84u
#define LE_REP0_LONG 96u   // [state][pos state]
#define LE_POS_SLOT 144u   // [length state][64]
#define LE_POS_SPECIAL 400u
#define LE_ALIGN 515u
#define LE_LEN 531u        // choice, choice2, low[4][8], mid[4][8], high[256]
#define LE_REP_LEN 853u
#define LE_LIT 1175u       // [previous byte >> 5][0x300]
#define LE_NPROBS 7320u    // 7319 and one to make whole dwords
#define LE_DIRECT 0xFFFFu
#define LE_HEAD 2u
#define LE_SLOT 65552u

struct LeRc {  // range coder state (wavefront-uniform but for obv)
    uint64_t low;
    // add: what the decisions since the last byte step added to low (the
    // additions never exceed the range at that step, so 32 bits hold them;
    // the carry into low is taken once, at the next byte step)
    uint32_t add, range, cache, pending;
    // output: nout bytes so far (counted on past cap: more than cap means
    // the coded chunk does not fit and the raw one is written); the bytes of
    // the current word in acc from bit 0, sh = the next byte's shift; the
    // words of the current 256-byte block in obv, word j in lane j
    uint32_t nout, cap, acc, sh, obv;
};

// Decisions of one symbol, one per lane: idx = probability index or
// LE_DIRECT, bit; e = the count so far, d0 / dn = the run of direct bits
// (at most one per symbol) -- all three uniform.
struct LeList {
    uint32_t lane, idx, bit, e, d0, dn;
    __device__ __forceinline__ explicit LeList(uint32_t l) : lane(l), idx(LE_DIRECT), bit(0), e(0), d0(0), dn(0) {}
    __device__ __forceinline__ void one(uint32_t i, uint32_t b) {
        if (lane == e) idx = i, bit = b;
        e++;
    }
    // most significant bit first, node index from 1
    __device__ __forceinline__ void tree(uint32_t base, uint32_t nb, uint32_t v) {
        const uint32_t i = lane - e;
        if (i < nb) idx = base + ((1u << i) | (v >> (nb - i))), bit = (v >> (nb - 1 - i)) & 1u;
        e += nb;
    }
    // least significant bit first: node i is 1 followed by bits 0 .. i-1
    __device__ __forceinline__ void tree_reverse(uint32_t base, uint32_t nb, uint32_t v) {
        const uint32_t i = lane - e;
        if (i < nb)
            idx = base + ((1u << i) | (i ? __builtin_bitreverse32(v) >> (32 - i) : 0u)), bit = (v >> i) & 1u;
        e += nb;
    }
    __device__ __forceinline__ void direct(uint32_t v, uint32_t nb) {
        const uint32_t i = lane - e;
        if (i < nb) idx = LE_DIRECT, bit = (v >> (nb - 1 - i)) & 1u;
        d0 = e;
        dn = nb;
        e += nb;
    }
    __device__ __forceinline__ void length(uint32_t base, uint32_t len, uint32_t ps) {
        const uint32_t l = len - 2;
        if (l < 8) {
            one(base, 0);
            tree(base + 2 + ps * 8, 3, l);
        } else if (l < 16) {
            one(base, 1);
            one(base + 1, 0);
            tree(base + 34 + ps * 8, 3, l - 8);
        } else {
            one(base, 1);
            one(base + 1, 1);
            tree(base + 66, 8, l - 16);
        }
    }
};

// lane `lane` of v set to the uniform value x (the lane index in M0: two
// scalar operands would exceed the constant bus)
__device__ __forceinline__ void write_lane(uint32_t& v, uint32_t x, uint32_t lane) {
    __asm__("v_writelane_b32 %0, %1, %2" : "+v"(v) : "s"(uniform(x)), "{m0}"(lane));
}

// One finished byte: into acc; a whole word into its lane of obv; a whole
// block of 256 bytes to the slot as 64 aligned dwords (only blocks inside
// the largest chunk: past it the raw chunk is written anyway).
__device__ __forceinline__ void le_put(LeRc& rc, uint32_t b, uint8_t* data) {
    rc.acc |= (b & 0xFFu) << rc.sh;
    rc.sh += 8;
    rc.nout++;
    if (rc.sh == 32) {
        write_lane(rc.obv, rc.acc, ((rc.nout >> 2) - 1) & (WAVE - 1));
        rc.acc = 0;
        rc.sh = 0;
        if ((rc.nout & 255) == 0 && rc.nout <= GC_SUB_BLOCK)
            ((uint32_t*)(data + rc.nout - 256))[lane_id()] = rc.obv;
    }
}

// The range coder's byte step (after a decision leaves range < 2^24): the
// top byte of low is final unless a later carry can still reach it, which
// is the case while it is 0xFF -- such bytes wait as a count (pending)
// behind the last byte that can take a carry (cache).
__device__ __forceinline__ void le_shift_low(LeRc& rc, uint8_t* data) {
    rc.low += rc.add;
    rc.add = 0;
    const uint32_t lo = (uint32_t)rc.low, hi = (uint32_t)(rc.low >> 32);
    if (lo < 0xFF000000u || hi != 0) {
        le_put(rc, rc.cache + hi, data);
        for (uint32_t j = 1; j < rc.pending; j++) le_put(rc, 0xFFu + hi, data);
        rc.pending = 0;
        rc.cache = lo >> 24;
    }
    rc.pending++;
    rc.low = (uint64_t)((lo & 0x00FFFFFFu) << 8);
}

// One decision: probability pv, bit = bit k of v (all uniform).  bound =
// (range >> 11) * pv; a 1 adds bound to low and leaves range - bound, a 0
// leaves bound.  Written out because the compiler spends 11 scalar
// instructions on it (a bit-field extract, two ANDs and a compare for the
// bit) where a bit test and two selects do: 7, plus the normalisation test.
// One normalisation step always restores range >= 2^24: a probability stays
// within [31, 2017], so range >= 2^13 * 31 after a decision.
__device__ __forceinline__ void le_bit(LeRc& rc, uint32_t pv, uint32_t v, uint32_t k, uint8_t* data) {
    uint32_t t, u;
    __asm__("s_lshr_b32 %[t], %[r], 11\n\t"
            "s_mul_i32 %[t], %[t], %[p]\n\t"
            "s_sub_u32 %[u], %[r], %[t]\n\t"
            "s_bitcmp1_b32 %[v], %[k]\n\t"
            "s_cselect_b32 %[r], %[u], %[t]\n\t"
            "s_cselect_b32 %[t], %[t], 0\n\t"
            "s_add_u32 %[a], %[a], %[t]"
            : [r] "+s"(rc.range), [a] "+s"(rc.add), [t] "=&s"(t), [u] "=&s"(u)
            : [p] "s"(pv), [v] "s"(v), [k] "s"(k)
            : "scc");
    if (__builtin_expect(rc.range < (1u << 24), 0)) {
        rc.range <<= 8;
        le_shift_low(rc, data);
    }
}

// The probability update of a decision's lane (branch-free).
__device__ __forceinline__ uint32_t le_update(uint32_t pv, uint32_t bit) {
    const uint32_t dn = pv - (pv >> 5), up = pv + ((2048u - pv) >> 5);
    return bit ? dn : up;
}

// Decisions [k0, k1) of a list coded with probabilities pv (bit k of bits
// = decision k's bit).
__device__ __forceinline__ void le_run(LeRc& rc, uint32_t pv, uint32_t bits, uint32_t k0, uint32_t k1,
                                       uint8_t* data) {
    for (uint32_t k = k0; k < k1; k++) le_bit(rc, bcast(pv, k), bits, k, data);
}

// Codes the list: probabilities read and updated by the lanes at once (lanes
// without one use the spare slot), the range coder stepping through the
// decisions in order (the direct bits, when there are any, as their own run).
__device__ __forceinline__ void le_code(LeRc& rc, const LeList& L, uint16_t* model, uint8_t* data) {
    const uint32_t e = uniform(L.e), dn = uniform(L.dn);
    const uint32_t d0 = dn ? uniform(L.d0) : e;
    const uint32_t idx = L.idx != LE_DIRECT ? L.idx : LE_NPROBS - 1;  // lanes past the list hold LE_DIRECT
    const uint32_t pv = model[idx];
    model[idx] = (uint16_t)le_update(pv, L.bit);
    const uint32_t bits = uniform((uint32_t)ballot(L.bit != 0));  // at most 32 decisions; lanes past the list hold 0
    le_run(rc, pv, bits, 0, d0, data);
    if (dn) {
        for (uint32_t k = d0; k < d0 + dn; k++) {
            rc.range >>= 1;
            rc.add += ((bits >> k) & 1u) ? rc.range : 0u;
            if (__builtin_expect(rc.range < (1u << 24), 0)) {
                rc.range <<= 8;
                le_shift_low(rc, data);
            }
        }
        le_run(rc, pv, bits, d0 + dn, e, data);
    }
}

// A literal: is_match 0, then the byte's 8 bits msb first through the
// literal coder of context prev >> 5, in matched mode (state >= 7) steered
// by mbyte until the first bit where the two differ.  Lane 0 reads and
// updates is_match, lanes 1-8 the tree nodes; the 9 decisions are unrolled.
__device__ __forceinline__ void le_literal(LeRc& rc, uint32_t state, uint32_t p, uint32_t byte, uint32_t prev,
                                           uint32_t mbyte, uint16_t* model, uint8_t* data, uint32_t lane) {
    const uint32_t i = lane - 1;
    const uint32_t base = LE_LIT + (prev >> 5) * 0x300u;
    const uint32_t sym = (1u << i) | (byte >> (8 - i));
    uint32_t idx = base + sym;
    if (state >= 7) {
        const bool same = (byte >> (8 - i)) == (mbyte >> (8 - i));
        idx += same ? (1u + ((mbyte >> (7 - i)) & 1u)) << 8 : 0u;
    }
    if (lane == 0) idx = LE_IS_MATCH + state * 4 + (p & 3u);
    if (lane >= 9) idx = LE_NPROBS - 1;  // the spare slot: every lane reads and writes, none branches
    const uint32_t bit = lane ? (byte >> (8 - lane)) & 1u : 0u;
    const uint32_t pv = model[idx];
    model[idx] = (uint16_t)le_update(pv, bit);
    uint32_t q[9];  // every probability in a scalar register before the chain starts
#pragma unroll
    for (uint32_t k = 0; k < 9; k++) q[k] = bcast(pv, k);
    const uint32_t v = uniform(byte);
    le_bit(rc, q[0], 0u, 0, data);
#pragma unroll
    for (uint32_t k = 1; k < 9; k++) le_bit(rc, q[k], v, 8 - k, data);
}

__device__ __forceinline__ uint32_t le_after_literal(uint32_t s) { return s < 4 ? 0 : (s < 10 ? s - 3 : s - 6); }

// Repeat length at the window position k with the mismatch mask mm (at most
// limit), continued past the window 64 bytes at a time.
__device__ __forceinline__ uint32_t le_rep_len(uint64_t mm, uint32_t k, uint32_t limit, gbytes d, uint32_t wend,
                                               uint32_t r, uint32_t end, uint32_t lane) {
    const uint64_t x = mm >> k;
    if (x) return min((uint32_t)__builtin_ctzll(x), limit);
    uint32_t l = WAVE - k;
    for (uint32_t q = wend; l < limit; q += WAVE) {
        const uint32_t pos = q + lane;
        const bool ok = pos < end && d[pos] == d[pos - r];
        const uint64_t m = ballot(!ok);
        if (m) return min(l + (uint32_t)__builtin_ctzll(m), limit);
        l += WAVE;
    }
    return limit;
}

__device__ void le_item(const gc_item& it, const uint8_t* arena, const uint32_t* packed, uint16_t* model,
                        uint8_t* out, uint32_t* sizes) {
    const uint32_t lane = lane_id();
    const uint32_t n = it.n;
    gbytes d = (gbytes)at(arena, it.src_off);
    const uint32_t GC_GLOBAL* pk = (const uint32_t GC_GLOBAL*)(packed + it.opos);
    uint8_t* slot = out + (uint64_t)it.count_pos * LE_SLOT;
    uint8_t* data = slot + LE_HEAD + 6;
    uint32_t* m32 = (uint32_t*)model;
    for (uint32_t i = lane; i < LE_NPROBS / 2; i += WAVE) m32[i] = 0x04000400u;  // every probability 1024

    LeRc rc;
    rc.low = 0;
    rc.add = 0;
    rc.range = 0xFFFFFFFFu;
    rc.cache = 0;
    rc.pending = 1;
    rc.nout = 0;
    rc.acc = 0;
    rc.sh = 0;
    rc.obv = 0;
    rc.cap = min(n + 5, GC_SUB_BLOCK);  // the host coder's room; past it the raw chunk is smaller anyway
    uint32_t state = 0, rep0 = 1, rep1 = 1, rep2 = 1, rep3 = 1;
    uint32_t p = 0, wb = 0;
    bool fresh = false;
    uint32_t cur = 0, src0 = 0, pkw = 0, pk64 = 0, prev0 = 0;
    uint64_t mm0 = 0, mm1 = 0, mm2 = 0, mm3 = 0, quiet = 0;
    while (p < n && rc.nout <= rc.cap) {
        if (!fresh || p >= wb + WAVE) {  // a window at p
            wb = p;
            const uint32_t pos = wb + lane;
            const bool in = pos < n;
            cur = in ? d[pos] : 0u;
            const uint32_t s0 = in && rep0 <= pos ? d[pos - rep0] : 0u;
            const uint32_t s1 = in && rep1 <= pos ? d[pos - rep1] : 0u;
            const uint32_t s2 = in && rep2 <= pos ? d[pos - rep2] : 0u;
            const uint32_t s3 = in && rep3 <= pos ? d[pos - rep3] : 0u;
            pkw = in ? pk[pos] : 0u;
            pk64 = wb + WAVE < n ? pk[wb + WAVE] : 0u;
            prev0 = wb ? d[wb - 1] : 0u;
            mm0 = ballot(!(in && rep0 <= pos && s0 == cur));
            mm1 = ballot(!(in && rep1 <= pos && s1 == cur));
            mm2 = ballot(!(in && rep2 <= pos && s2 == cur));
            mm3 = ballot(!(in && rep3 <= pos && s3 == cur));
            src0 = s0;
            // quiet positions: no packed match of 2 bytes or more and every
            // repeat length below 2 (a mismatch here or at the next
            // position; at the window's last position only a mismatch here
            // counts), so the step is a literal or a short repeat
            const uint32_t pl = pkw & 0xFFFFu, pd = pkw >> 16;
            quiet = ~ballot(in && pl >= 2 && pd != 0 && pd <= pos) & (mm0 | (mm0 >> 1)) & (mm1 | (mm1 >> 1)) &
                    (mm2 | (mm2 >> 1)) & (mm3 | (mm3 >> 1));
            fresh = true;
        }
        uint32_t k = p - wb;
        if (state < 7 && ((quiet >> k) & 1u)) {
            // a run of plain literals: the walk's tests are settled by the
            // masks (a full output buffer stops the walk after the run: the
            // coder writes nothing past it)
            const uint64_t x = ~(quiet >> k);
            const uint32_t run = min(x ? (uint32_t)__builtin_ctzll(x) : (uint32_t)WAVE, n - p);
            uint32_t prev = k ? bcast(cur, k - 1) : prev0;
            for (uint32_t j = 0; j < run; j++) {
                const uint32_t byte = bcast(cur, k + j);
                le_literal(rc, state, p + j, byte, prev, 0, model, data, lane);
                state = state < 4 ? 0u : state - 3;  // le_after_literal below 7
                prev = byte;
            }
            p += run;
            continue;
        }
        const uint32_t avail = min(n - p, 273u);
        const uint32_t ps = p & 3u;
        // the packed match here (gc_lzma_enc.cpp gpu_match)
        const uint32_t w = bcast(pkw, k);
        const uint32_t mdist = w >> 16;
        uint32_t mlen = min(min(w & 0xFFFFu, 273u), n - p);
        if (mlen < 2 || mdist == 0 || mdist > p) mlen = 0;
        uint32_t rlen = 0, ridx = 0;
        if (avail >= 2) {
            const uint32_t wend = wb + WAVE, end = p + avail;
            uint32_t l = le_rep_len(mm0, k, avail, d, wend, rep0, end, lane);
            if (l > rlen) rlen = l, ridx = 0;
            l = le_rep_len(mm1, k, avail, d, wend, rep1, end, lane);
            if (l > rlen) rlen = l, ridx = 1;
            l = le_rep_len(mm2, k, avail, d, wend, rep2, end, lane);
            if (l > rlen) rlen = l, ridx = 2;
            l = le_rep_len(mm3, k, avail, d, wend, rep3, end, lane);
            if (l > rlen) rlen = l, ridx = 3;
        }
        // a match at a repeat distance is coded as that repeat
        if (mlen) {
            const uint32_t i = rep0 == mdist ? 0u : rep1 == mdist ? 1u : rep2 == mdist ? 2u : rep3 == mdist ? 3u : 4u;
            if (i < 4 && mlen >= rlen) rlen = mlen, ridx = i, mlen = 0;
        }
        bool take_rep = false, take_match = false;
        if (rlen >= 2 && (rlen >= 64 || rlen + 1 >= mlen || (rlen + 2 >= mlen && mdist >= (1u << 9)) ||
                          (rlen + 3 >= mlen && mdist >= (1u << 15))))
            take_rep = true;
        else if (mlen >= 2 && !(mlen == 2 && mdist >= 128))
            take_match = true;
        if (take_match && mlen < 64 && p + 1 < n) {  // one lazy step
            const uint32_t w1 = k + 1 < WAVE ? bcast(pkw, k + 1) : pk64;
            const uint32_t ndist = w1 >> 16;
            uint32_t nlen = min(min(w1 & 0xFFFFu, 273u), n - p - 1);
            if (nlen < 2 || ndist == 0 || ndist > p + 1) nlen = 0;
            if (nlen >= mlen + 2 || (nlen == mlen + 1 && ndist < mdist)) take_match = false;
        }
        LeList L(lane);
        if (take_rep) {
            L.one(LE_IS_MATCH + state * 4 + ps, 1);
            L.one(LE_IS_REP + state, 1);
            if (ridx == 0) {
                L.one(LE_IS_REP_G0 + state, 0);
                L.one(LE_REP0_LONG + state * 4 + ps, 1);
            } else {
                L.one(LE_IS_REP_G0 + state, 1);
                if (ridx == 1) {
                    L.one(LE_IS_REP_G1 + state, 0);
                } else {
                    L.one(LE_IS_REP_G1 + state, 1);
                    L.one(LE_IS_REP_G2 + state, ridx == 3 ? 1u : 0u);
                }
                const uint32_t dist = ridx == 1 ? rep1 : ridx == 2 ? rep2 : rep3;
                if (ridx == 3) rep3 = rep2;
                if (ridx >= 2) rep2 = rep1;
                rep1 = rep0;
                rep0 = dist;
                fresh = false;
            }
            L.length(LE_REP_LEN, rlen, ps);
            state = state < 7 ? 8u : 11u;
            p += rlen;
        } else if (take_match) {
            L.one(LE_IS_MATCH + state * 4 + ps, 1);
            L.one(LE_IS_REP + state, 0);
            L.length(LE_LEN, mlen, ps);
            const uint32_t dd = mdist - 1;
            uint32_t slot_ = dd;
            if (dd >= 4) {
                const uint32_t top = 31u - (uint32_t)__builtin_clz(dd);
                slot_ = 2 * top + ((dd >> (top - 1)) & 1u);
            }
            L.tree(LE_POS_SLOT + min(mlen - 2, 3u) * 64, 6, slot_);
            if (slot_ >= 4) {
                const uint32_t nbits = (slot_ >> 1) - 1;
                const uint32_t base = (2u | (slot_ & 1u)) << nbits;
                const uint32_t reduced = dd - base;
                if (slot_ < 14) {
                    L.tree_reverse(LE_POS_SPECIAL + base - slot_, nbits, reduced);
                } else {
                    L.direct(reduced >> 4, nbits - 4);
                    L.tree_reverse(LE_ALIGN, 4, reduced & 15u);
                }
            }
            rep3 = rep2;
            rep2 = rep1;
            rep1 = rep0;
            rep0 = mdist;
            fresh = false;
            state = state < 7 ? 7u : 10u;
            p += mlen;
        } else if (state >= 7 && !((mm0 >> k) & 1u)) {
            // the short repeat: the byte at distance rep0 (what a literal here
            // would be coded against)
            L.one(LE_IS_MATCH + state * 4 + ps, 1);
            L.one(LE_IS_REP + state, 1);
            L.one(LE_IS_REP_G0 + state, 0);
            L.one(LE_REP0_LONG + state * 4 + ps, 0);
            state = state < 7 ? 9u : 11u;
            p += 1;
        } else {
            le_literal(rc, state, p, bcast(cur, k), k ? bcast(cur, k - 1) : prev0, bcast(src0, k), model, data, lane);
            state = le_after_literal(state);
            p += 1;
            continue;
        }
        le_code(rc, L, model, data);
    }
    for (int i = 0; i < 5; i++) le_shift_low(rc, data);
    const uint32_t c = rc.nout;
    if (c <= rc.cap && 7 + c < n + 4) {
        // the last block's whole words are in obv, its last 0-3 bytes in acc
        // (the rest of acc is 0): acc's word also carries the end marker
        write_lane(rc.obv, rc.acc, (c >> 2) & (WAVE - 1));
        if (lane <= ((c & 255u) >> 2)) ((uint32_t*)(data + (c & ~255u)))[lane] = rc.obv;
        if (lane < 6) {
            const uint32_t h = lane == 0   ? 0xE0u | ((n - 1) >> 16)
                               : lane == 1 ? ((n - 1) >> 8) & 0xFFu
                               : lane == 2 ? (n - 1) & 0xFFu
                               : lane == 3 ? ((c - 1) >> 8) & 0xFFu
                               : lane == 4 ? (c - 1) & 0xFFu
                                           : 2u * 5u * 9u + 3u;  // (pb * 5 + lp) * 9 + lc
            slot[LE_HEAD + lane] = (uint8_t)h;
        }
        if (lane == 0) sizes[it.count_pos] = 7 + c;
    } else {
        // one uncompressed chunk with a dictionary reset
        uint8_t* r = slot + LE_HEAD;
        for (uint32_t j = lane; j < n; j += WAVE) r[3 + j] = d[j];
        if (lane == 0) {
            r[0] = 0x01;
            r[1] = (uint8_t)((n - 1) >> 8);
            r[2] = (uint8_t)(n - 1);
            r[3 + n] = 0;
            sizes[it.count_pos] = n + 4;
        }
    }
}

// order (optional): the items in the order to code them, costliest first
// (gc_order_items over gc_lzma2_cost_kernel's counts).  A sub-block's walk
// is long and serial, and one full of literals takes about three times the
// average: started last, such walks would set the end of the launch.
extern "C" __global__ void __launch_bounds__(WAVE)
gc_lzma2_encode_kernel(const uint8_t* arena, const gc_item* __restrict__ items, uint32_t n_items,
                       const uint32_t* packed, uint8_t* out, uint32_t* sizes, const uint32_t* order) {
    __shared__ uint16_t model[LE_NPROBS];
    for (uint32_t i = blockIdx.x; i < n_items; i += gridDim.x)
        le_item(load_item(items, order ? uniform(order[i]) : i), arena, packed, model, out, sizes);
}

// The coder's cost of each item, estimated as its positions without a packed
// match of 2 bytes or more (where the walk codes literals one at a time).
// One 256-thread workgroup per item.
extern "C" __global__ void __launch_bounds__(256)
gc_lzma2_cost_kernel(const gc_item* __restrict__ items, uint32_t n_items, const uint32_t* packed, uint32_t* cost) {
    __shared__ uint32_t part[256 / WAVE];
    for (uint32_t i = blockIdx.x; i < n_items; i += gridDim.x) {
        const gc_item it = load_item(items, i);
        const uint32_t* pk = packed + it.opos;
        uint32_t c = 0;
        for (uint32_t p = threadIdx.x; p < it.n; p += 256) c += (pk[p] & 0xFFFFu) < 2 ? 1u : 0u;
        c = wave_sum(c);
        if (lane_id() == 0) part[threadIdx.x / WAVE] = c;
        __syncthreads();
        if (threadIdx.x == 0) cost[i] = part[0] + part[1] + part[2] + part[3];
        __syncthreads();
    }
}
#endif  // GC_BUILD_MAIN
