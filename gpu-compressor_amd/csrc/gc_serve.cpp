// gc_serve.cpp -- served form: match-finding requests admitted under an HBM
// budget and batched continuously into the persistent kernels.
//
// Clients submit requests (Zstd L1/L2 or LZMA2, any size) from any thread:
// from host memory (the input is copied in, the results go back to host
// memory) or already in device memory (gc_server_submit_dev: the input is
// read where it lies and the results stay on the device).  One server
// thread:
//   1. admits pending requests in arrival order while their device memory
//      (candidates, Zstd info words and sequence slots or LZMA2 matches, and
//      a copy of a host input) fits the budget; a host input is copied in on
//      the context's second stream;
//   2. whenever one of kSlots tick slots is free and requests are ready,
//      starts a tick with every ready request (up to max_batch_items
//      sub-blocks): it writes the tick's work list into the slot's pinned
//      staging memory and launches the slot's pass on its stream -- a setup
//      kernel that loads the list into the slot's device work queue, then
//      the persistent kernels (candidates, LZMA2 lengths, Zstd info, parse
//      order, parse), which take the tick's sub-blocks from the queue, with
//      an event between every two kernels (per-kernel device time is part of
//      the server's statistics).  Each slot has its own stream, so
//      consecutive ticks overlap on the device: a request that arrives while
//      others run starts at the next free slot, beside them, instead of
//      waiting for them to end;
//   3. when a tick ends: device-result requests are done at once (their
//      sequence counts were written by the parse into pinned memory); the
//      others have their results brought back on a fetch stream, straight
//      into the caller's buffers when it gave pinned ones, and are done when
//      that copy ends.
// A request's results do not depend on what it was batched with (sub-blocks
// are independent).  No reference counterpart: the reference processes one
// chunk at a time (pipeline.rs:138-175).
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/gcompress.h"
#include "gc_runtime.h"

using namespace gc;

namespace {

const uint32_t kMaxSeq = GC_SUB_BLOCK / 3;
// A device-result Zstd request's sub-block region: its candidates (2 B per
// position), then its info words (4 B), with its sequence slot
// (GC_SERVE_ZSTD_SLOT records of 16 B) laid over both from the start.
const uint64_t kOverlay = (uint64_t)GC_SERVE_ZSTD_SLOT * sizeof(gc_zstd_sequence);
static_assert(kOverlay == 6ull * GC_SUB_BLOCK, "slot = candidates + info words of one sub-block");
const uint64_t kAlign = 256;
const int kSlots = 3;
// Under a backlog (requests waiting for admission, i.e. the HBM budget is
// full, or more than kBacklogItems sub-blocks admitted and waiting) at most
// kBacklogTicks ticks run at once: each tick then holds as much as the
// budget admits, and large ticks leave fewer kernel tails than many
// overlapping ones (saturation throughput on the 6000-request mix: 1 tick
// 149.4-150.3 GB/s, 2 ticks 150.6, 3 ticks about 141); one tick at a time
// also keeps each kernel alone on the GPU, so its measured duration is its
// own.  Under a light load every slot takes work as soon as it is free
// (latency).
const int kBacklogTicks = 1;
const uint64_t kBacklogItems = 16384;

uint64_t up(uint64_t v, uint64_t a) { return (v + a - 1) / a * a; }

// Device address of pinned host memory from hipHostMalloc.
template <class T>
T* dev_addr(T* host) {
    void* d = nullptr;
    if (host && hipHostGetDevicePointer(&d, host, 0) == hipSuccess && d) return (T*)d;
    return host;
}

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// First-fit allocator over [0, size).
struct PoolAlloc {
    uint64_t size = 0;
    std::map<uint64_t, uint64_t> free_;  // offset -> length
    uint64_t used = 0;
    void reset(uint64_t s) {
        size = s;
        free_.clear();
        if (s) free_[0] = s;
        used = 0;
    }
    bool alloc(uint64_t len, uint64_t* off) {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < len) continue;
            *off = it->first;
            uint64_t rest = it->second - len;
            uint64_t start = it->first + len;
            free_.erase(it);
            if (rest) free_[start] = rest;
            used += len;
            return true;
        }
        return false;
    }
    void release(uint64_t off, uint64_t len) {
        if (!len) return;
        used -= len;
        auto it = free_.emplace(off, len).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

enum State { PENDING, COPYING, READY, RUNNING, FETCHING, DONE, FAILED };

struct Req {
    uint64_t id = 0;
    const uint8_t* data = nullptr;  // host input
    const uint8_t* d_in = nullptr;  // device input (the caller's, or the copy of a host input)
    uint64_t n = 0;
    uint32_t algo = 0, level = 0, nsub = 0;
    bool dev_result = false;  // results stay on the device (gc_server_submit_dev)
    uint64_t dev_off = 0, dev_len = 0;  // the request's region of the pool
    uint64_t in_off = 0, cand_off = 0, info_off = 0, out_off = 0;  // byte offsets in the region
    bool overlay = false;  // the sequence slots lie over the candidates and info words (kOverlay per sub-block)
    uint64_t cnt_off = 0;  // first entry of the request's sequence counts in the counts zone
    uint32_t cnt_len = 0;
    void* dst = nullptr;      // caller's result buffer (gc_server_submit_into), else results kept here
    void* dst_dev = nullptr;  // the same buffer as the device sees it, when it is pinned host memory
    uint64_t dst_cap = 0;
    uint32_t* dst_counts = nullptr;
    hipEvent_t copied = nullptr;
    State state = PENDING;
    int error = 0;
    double t_submit = 0, t_admit = 0, t_start = 0, t_done = 0;
    std::vector<uint32_t> counts;
    std::vector<gc_zstd_sequence> seqs;
    std::vector<uint32_t> packed;
    uint64_t fetch_total = 0;  // Zstd: sequences found
};

// Device bytes a request needs (every part 256-byte aligned).  For a Zstd
// request whose results go to the host the candidate and info regions are
// contiguous and together hold at least 21845 x 16 bytes per sub-block:
// packed sequences for a pageable destination are gathered there at the
// end.  A Zstd request whose results stay on the device has one region of
// kOverlay bytes per sub-block holding its candidates, then its info words,
// and, from its start, its sequence slot: the candidates are dead once the
// info pass has read them, and the parse writes its records 64 at a time
// behind the info words it still reads (after j sequences it has passed
// position 3j, so record j + 63 ends at byte 16 (j + 64), below the first
// info word it may still read, 2 * 65536 + 4 * (3 (j + 64) - 64): j + 64 <=
// 21845 + 63 keeps it there), so a request holds 6 bytes per input byte
// instead of 11.3 and the HBM budget admits larger ticks.
void layout(Req& r) {
    const uint64_t ns = r.nsub, slots = ns * GC_SUB_BLOCK;
    uint64_t o = 0;
    if (!r.d_in) {  // a host input is copied here (with room past its end for the kernels' reads)
        r.in_off = o;
        o += up(r.n + 64, kAlign);
    }
    if (r.dev_result && r.algo == GC_ALGO_ZSTD) {
        r.overlay = true;
        r.cand_off = r.info_off = r.out_off = o;
        o += ns * kOverlay;
        r.dev_len = o;
        return;
    }
    r.cand_off = o;
    o += up(slots * 2, kAlign);
    if (r.algo == GC_ALGO_ZSTD) {
        r.info_off = o;
        o += up(slots * 4, kAlign);
        r.out_off = o;
        o += up(ns * kMaxSeq * sizeof(gc_zstd_sequence), kAlign);
    } else {
        r.out_off = o;
        o += up(slots * 4, kAlign);
    }
    r.dev_len = o;
}

// Kernels of a tick, in launch order (events between them).
enum { K_SETUP, K_CAND, K_LZMA2, K_INFO, K_ORDER, K_PARSE, K_COUNT };

// One tick slot: its stream, pinned work list and device queue.
struct Slot {
    hipStream_t st = nullptr;
    hipEvent_t ev[K_COUNT + 1] = {};  // before each kernel, and after the last
    hipEvent_t ev_done = nullptr;
    gc::DevBuf items, cost, order, queue;
    gc_queue* hdr = nullptr;    // pinned: the tick's counts and first items
    gc_item* stage = nullptr;   // pinned: the tick's items
    uint32_t* err = nullptr;    // pinned: the queue's error word, copied after each tick
    std::vector<Req*> reqs;
    bool busy = false;
    uint32_t n_l = 0, n_z = 0;  // the tick's LZMA2 and Zstd items
};

}  // namespace

struct gc_server {
    gc_ctx* ctx = nullptr;
    uint8_t* pool = nullptr;
    uint64_t budget = 0;
    PoolAlloc alloc;
    uint32_t* counts_zone = nullptr;  // pinned: Zstd sequence counts, written by the parse
    uint64_t zone_len = 0;
    PoolAlloc zone;
    std::thread th;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    bool stop = false;
    uint64_t next_id = 1;
    std::deque<uint64_t> pending;  // arrival order
    std::unordered_map<uint64_t, std::unique_ptr<Req>> reqs;
    // server thread only (requests held by address: a request in flight is
    // not released)
    std::vector<Req*> copying;    // host inputs on their way in
    std::deque<Req*> ready;       // admitted, input on the device, in arrival order
    uint64_t ready_items = 0;     // their sub-blocks
    Slot slot[kSlots];
    std::vector<Req*> fetching;
    std::deque<std::vector<Req*>> finished;  // ticks whose results still go back to the host
    hipStream_t fetch_stream = nullptr;
    hipEvent_t fetch_start = nullptr, fetch_done = nullptr;
    uint8_t* meta_host = nullptr;  // pinned: counts and offsets for the gathers
    size_t meta_cap = 0;
    uint32_t* zero_host = nullptr;  // pinned 0, to clear a queue's error word
    gc_server_stats stats{};
    uint32_t max_batch_items = 0;

    int setup_slot(Slot& sl);
    int launch_pass(Slot& sl, uint32_t n_all);
    void loop();
    void admit_locked();
    bool launch_tick(Slot& sl);
    void tick_ended(Slot& sl);
    void finish_fetch(std::vector<Req*>& done);
    void complete_fetch();
    void fail_req_locked(Req& r, int code);
};

int gc_server::setup_slot(Slot& sl) {
    const uint32_t cap = max_batch_items;
    hipError_t e = hipStreamCreateWithFlags(&sl.st, hipStreamNonBlocking);
    for (auto& ev : sl.ev)
        if (e == hipSuccess) e = hipEventCreate(&ev);
    if (e == hipSuccess) e = hipEventCreate(&sl.ev_done);
    if (e == hipSuccess) e = hipHostMalloc((void**)&sl.hdr, sizeof(gc_queue), 0);
    if (e == hipSuccess) e = hipHostMalloc((void**)&sl.stage, (size_t)cap * sizeof(gc_item), 0);
    if (e == hipSuccess) e = hipHostMalloc((void**)&sl.err, 256, 0);
    if (e != hipSuccess) return hip_fail(e, "gc_server_start: slot");
    memset(sl.hdr, 0, sizeof(gc_queue));
    *sl.err = 0;
    int rc = sl.items.ensure((size_t)cap * sizeof(gc_item));
    if (!rc) rc = sl.cost.ensure((size_t)cap * 4);
    if (!rc) rc = sl.order.ensure((size_t)cap * 4);
    if (!rc) {
        gc_queue q0{};
        rc = upload(ctx, sl.queue, &q0, sizeof(q0));
    }
    return rc;
}

// The tick's pass on the slot's stream: the setup kernel and the persistent
// kernels with resident grids (4 candidate workgroups per CU, 2 of the
// 1024-thread per-position kernels, 8 four-wavefront parse workgroups; fewer
// when the tick has fewer items).  Inputs are addressed absolutely (a null
// arena), everything else in the pool; the queue gives each kernel its items.
int gc_server::launch_pass(Slot& sl, uint32_t n_all) {
    gc_queue* q = (gc_queue*)sl.queue.p;
    gc_item* items = (gc_item*)sl.items.p;
    const uint32_t cus = (uint32_t)ctx->cu_count;
    auto grid = [&](uint32_t per_cu, uint32_t n) { return std::max(1u, std::min(per_cu * cus, n)); };
    hipStream_t st = sl.st;
    int rc = GC_OK;
    hipError_t e = hipEventRecord(sl.ev[K_SETUP], st);
    if (e == hipSuccess) rc = launch_tick_setup(ctx, st, dev_addr(sl.hdr), dev_addr(sl.stage), q, items, max_batch_items);
    if (!rc && e == hipSuccess) e = hipEventRecord(sl.ev[K_CAND], st);
    if (!rc && e == hipSuccess) rc = launch_cand_q(ctx, st, nullptr, items, q, grid(4, n_all), (uint16_t*)pool, true);
    if (!rc && e == hipSuccess) e = hipEventRecord(sl.ev[K_LZMA2], st);
    if (!rc && e == hipSuccess && sl.n_l)
        rc = launch_lzma2_len_q(ctx, st, nullptr, items, q, grid(2, sl.n_l), (const uint16_t*)pool, (uint32_t*)pool,
                                true);
    if (!rc && e == hipSuccess) e = hipEventRecord(sl.ev[K_INFO], st);
    if (!rc && e == hipSuccess && sl.n_z)
        rc = launch_zstd_info_q(ctx, st, nullptr, items, q, grid(2, sl.n_z), (const uint16_t*)pool, (uint32_t*)pool, 8,
                                (uint32_t*)sl.cost.p, true);
    if (!rc && e == hipSuccess) e = hipEventRecord(sl.ev[K_ORDER], st);
    if (!rc && e == hipSuccess && sl.n_z)
        rc = launch_order(ctx, st, (const uint32_t*)sl.cost.p, 0, (uint32_t*)sl.order.p, q);
    if (!rc && e == hipSuccess) e = hipEventRecord(sl.ev[K_PARSE], st);
    if (!rc && e == hipSuccess && sl.n_z)
        rc = launch_zstd_parse_q(ctx, st, nullptr, items, q, grid(8, (sl.n_z + 3) / 4), (const uint32_t*)pool, pool,
                                 dev_addr(counts_zone), kMaxSeq, (const uint32_t*)sl.order.p, true);
    if (!rc && e == hipSuccess) e = hipEventRecord(sl.ev[K_COUNT], st);
    if (!rc && e == hipSuccess)
        e = hipMemcpyAsync(sl.err, (uint8_t*)q + offsetof(gc_queue, err), 4, hipMemcpyDeviceToHost, st);
    if (!rc && e == hipSuccess) e = hipEventRecord(sl.ev_done, st);
    if (rc) return rc;
    return e == hipSuccess ? GC_OK : hip_fail(e, "gc_server: tick launch");
}

void gc_server::fail_req_locked(Req& r, int code) {
    r.state = FAILED;
    r.error = code;
    alloc.release(r.dev_off, r.dev_len);
    r.dev_len = 0;
    zone.release(r.cnt_off, r.cnt_len);
    r.cnt_len = 0;
    if (r.copied) (void)hipEventDestroy(r.copied);
    r.copied = nullptr;
    r.t_done = now_s();
    cv_done.notify_all();
}

void gc_server::admit_locked() {
    while (!pending.empty()) {
        Req& r = *reqs[pending.front()];
        if (!alloc.alloc(r.dev_len, &r.dev_off)) break;  // in arrival order: wait for memory
        if (r.algo == GC_ALGO_ZSTD) {
            uint64_t z = 0;
            if (!zone.alloc(r.nsub, &z)) {
                alloc.release(r.dev_off, r.dev_len);
                break;
            }
            r.cnt_off = z;
            r.cnt_len = r.nsub;
        }
        pending.pop_front();
        r.t_admit = now_s();
        if (r.d_in) {
            r.state = READY;
            ready.push_back(&r);
            ready_items += r.nsub;
            continue;
        }
        r.state = COPYING;
        uint8_t* dst = pool + r.dev_off + r.in_off;
        r.d_in = dst;
        hipError_t e = hipMemcpyAsync(dst, r.data, r.n, hipMemcpyHostToDevice, ctx->stream2);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&r.copied, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventRecord(r.copied, ctx->stream2);
        if (e != hipSuccess) {
            fail_req_locked(r, GC_EHIP);
            continue;
        }
        copying.push_back(&r);
    }
}

bool gc_server::launch_tick(Slot& sl) {
    std::vector<Req*> batch;
    {
        std::lock_guard<std::mutex> g(mu);
        uint32_t n_items = 0;
        while (!ready.empty()) {
            Req* r = ready.front();
            if (n_items + r->nsub > max_batch_items && !batch.empty()) break;
            ready.pop_front();
            ready_items -= r->nsub;
            batch.push_back(r);
            n_items += r->nsub;
        }
        const double t = now_s();
        for (Req* r : batch) {
            r->state = RUNNING;
            r->t_start = t;
        }
    }
    if (batch.empty()) return false;
    // items: LZMA2 first, then Zstd (the queue gives each kernel its range)
    uint32_t k = 0, n_l = 0;
    for (int pass = 0; pass < 2; pass++)
        for (Req* r : batch) {
            if ((pass == 0) != (r->algo == GC_ALGO_LZMA2)) continue;
            const uint64_t base = r->dev_off;
            const uint64_t readable = r->data ? up(r->n + 64, kAlign) : r->n;  // a copy has room past its end
            for (uint32_t s = 0; s < r->nsub; s++) {
                gc_item& it = sl.stage[k++];
                const uint64_t start = (uint64_t)s * GC_SUB_BLOCK;
                it.src_off = (uint64_t)(uintptr_t)(r->d_in + start);  // absolute (the pass's arena is null)
                it.avail = readable - start;
                it.cpos = (base + r->cand_off) / 2 + start;
                it.lit_pos = 0;
                it.n = (uint32_t)std::min<uint64_t>(GC_SUB_BLOCK, r->n - start);
                it.out_span = GC_SUB_BLOCK;
                if (r->overlay) {
                    const uint64_t reg = base + r->out_off + (uint64_t)s * kOverlay;
                    it.cpos = reg / 2;
                    it.opos = (reg + 2ull * GC_SUB_BLOCK) / 4;
                    it.seq_pos = reg / sizeof(gc_zstd_sequence);
                    it.count_pos = (uint32_t)(r->cnt_off + s);
                    it.depth = zstd_depth_for_level(r->level);
                } else if (r->algo == GC_ALGO_ZSTD) {
                    it.opos = (base + r->info_off) / 4 + start;
                    it.seq_pos = (base + r->out_off) / sizeof(gc_zstd_sequence) + (uint64_t)s * kMaxSeq;
                    it.count_pos = (uint32_t)(r->cnt_off + s);
                    it.depth = zstd_depth_for_level(r->level);
                } else {
                    it.opos = (base + r->out_off) / 4 + start;
                    it.seq_pos = 0;
                    it.count_pos = 0;
                    it.depth = GC_LZMA2_DEPTH;
                }
            }
            if (pass == 0) n_l += r->nsub;
        }
    const uint32_t n_all = k, n_z = n_all - n_l;
    gc_queue* h = sl.hdr;
    h->n[GC_Q_CAND] = n_all;
    h->n[GC_Q_LZMA2] = n_l;
    h->n[GC_Q_INFO] = n_z;
    h->n[GC_Q_PARSE] = n_z;
    h->base[GC_Q_CAND] = h->base[GC_Q_LZMA2] = 0;
    h->base[GC_Q_INFO] = h->base[GC_Q_PARSE] = n_l;
    sl.reqs = batch;
    sl.n_l = n_l;
    sl.n_z = n_z;
    if (launch_pass(sl, n_all) != GC_OK) {
        (void)hipStreamSynchronize(sl.st);
        std::lock_guard<std::mutex> g(mu);
        for (Req* r : batch) fail_req_locked(*r, GC_EHIP);
        sl.reqs.clear();
        return false;
    }
    sl.busy = true;
    stats.batches++;
    stats.items += n_all;
    return true;
}

// A tick ended on the device: device-result requests are done; the others
// wait for their results to be brought back.
void gc_server::tick_ended(Slot& sl) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, sl.ev[K_SETUP], sl.ev[K_COUNT]) == hipSuccess) stats.t_batches += ms * 1e-3;
    const bool ran[K_COUNT] = {true, true, sl.n_l > 0, sl.n_z > 0, sl.n_z > 0, sl.n_z > 0};
    for (int k = 0; k < K_COUNT; k++)
        if (ran[k] && hipEventElapsedTime(&ms, sl.ev[k], sl.ev[k + 1]) == hipSuccess) {
            stats.t_kernel[k] += ms * 1e-3;
            stats.launches[k]++;
        }
    const uint32_t err = *sl.err;
    if (err) {  // the queue protocol reported an error: this tick's results are not trusted
        *sl.err = 0;
        (void)hipMemcpyAsync((uint8_t*)sl.queue.p + offsetof(gc_queue, err), zero_host, 4, hipMemcpyHostToDevice,
                             sl.st);
        (void)hipStreamSynchronize(sl.st);
    }
    std::vector<Req*> back;
    {
        std::lock_guard<std::mutex> g(mu);
        const double t = now_s();
        for (Req* rp : sl.reqs) {
            Req& r = *rp;
            if (err) {
                fail_req_locked(r, GC_EHIP);
                continue;
            }
            if (r.algo == GC_ALGO_ZSTD) {
                r.counts.assign(counts_zone + r.cnt_off, counts_zone + r.cnt_off + r.nsub);
                uint64_t tot = 0;
                for (uint32_t c : r.counts) tot += c;
                r.fetch_total = tot;
                zone.release(r.cnt_off, r.cnt_len);
                r.cnt_len = 0;
                stats.zstd_bytes += r.n;
                stats.zstd_sub_blocks += r.nsub;
                stats.zstd_sequences += tot;
            } else {
                stats.lzma2_bytes += r.n;
            }
            if (r.dev_result) {
                r.state = DONE;
                r.t_done = t;
                stats.completed++;
                stats.bytes += r.n;
            } else {
                back.push_back(&r);
            }
        }
        cv_done.notify_all();
    }
    sl.reqs.clear();
    sl.busy = false;
    if (!back.empty()) finished.push_back(std::move(back));
}

// Bring a finished tick's host-bound results back on the fetch stream (no
// other fetch is in flight).  Zstd: each request's used sequence slots are
// packed by gc_gather_seqs (counts and offsets staged in pinned `meta`).
void gc_server::finish_fetch(std::vector<Req*>& done) {
    fetching.swap(done);
    done.clear();
    hipStream_t cs = fetch_stream;
    (void)hipEventRecord(fetch_start, cs);
    uint32_t n_z = 0;
    for (Req* r : fetching)
        if (r->algo == GC_ALGO_ZSTD) n_z += r->nsub;
    const size_t off_bytes = up((size_t)n_z * 4, 8);
    if (off_bytes + (size_t)n_z * 8 > meta_cap) {
        (void)hipHostFree(meta_host);
        meta_cap = off_bytes + (size_t)n_z * 8;
        if (hipHostMalloc((void**)&meta_host, meta_cap, 0) != hipSuccess) {
            meta_host = nullptr;
            meta_cap = 0;
            for (Req* r : fetching) r->error = GC_ENOMEM;
        }
    }
    uint32_t* mc = (uint32_t*)meta_host;
    uint64_t* mo = (uint64_t*)(meta_host + off_bytes);
    {
        std::lock_guard<std::mutex> g(mu);
        for (Req* r : fetching) r->state = FETCHING;
    }
    uint32_t at = 0;
    std::vector<uint32_t> first(fetching.size(), 0);
    for (size_t j = 0; j < fetching.size(); j++) {
        Req& r = *fetching[j];
        if (r.algo != GC_ALGO_ZSTD || !meta_host) continue;
        first[j] = at;
        uint64_t tot = 0;
        for (uint32_t s2 = 0; s2 < r.nsub; s2++) {
            mc[at + s2] = r.counts[s2];
            mo[at + s2] = tot;
            tot += r.counts[s2];
        }
        at += r.nsub;
        if (r.dst) {
            memcpy(r.dst_counts, r.counts.data(), (size_t)r.nsub * 4);
            if (tot * sizeof(gc_zstd_sequence) > r.dst_cap) r.error = GC_ESIZE;
        }
    }
    uint8_t* meta_dev = dev_addr(meta_host);
    for (size_t j = 0; j < fetching.size(); j++) {
        Req& r = *fetching[j];
        if (r.error) continue;
        uint8_t* base = pool + r.dev_off;
        if (r.algo == GC_ALGO_ZSTD) {
            if (!r.dst) r.seqs.resize(r.fetch_total);
            if (!r.fetch_total) continue;
            const uint32_t* mcounts = (const uint32_t*)meta_dev + first[j];
            const uint64_t* moffs = (const uint64_t*)(meta_dev + off_bytes) + first[j];
            if (r.dst_dev) {  // pinned caller memory: written by the gather kernel over PCIe
                (void)launch_gather(ctx, cs, base + r.out_off, mcounts, moffs, r.dst_dev, r.nsub, kMaxSeq);
                continue;
            }
            void* to = r.dst ? r.dst : (void*)r.seqs.data();
            (void)launch_gather(ctx, cs, base + r.out_off, mcounts, moffs, base + r.cand_off, r.nsub, kMaxSeq);
            (void)hipMemcpyAsync(to, base + r.cand_off, r.fetch_total * sizeof(gc_zstd_sequence),
                                 hipMemcpyDeviceToHost, cs);
        } else {
            if (r.dst_dev) {
                (void)launch_copy_u32(ctx, cs, (const uint32_t*)(base + r.out_off), (uint32_t*)r.dst_dev, r.n);
                continue;
            }
            if (!r.dst) r.packed.resize(r.n);
            void* to = r.dst ? r.dst : (void*)r.packed.data();
            (void)hipMemcpyAsync(to, base + r.out_off, r.n * 4, hipMemcpyDeviceToHost, cs);
        }
    }
    (void)hipEventRecord(fetch_done, cs);
}

void gc_server::complete_fetch() {
    float ms = 0;
    if (hipEventElapsedTime(&ms, fetch_start, fetch_done) == hipSuccess) stats.t_fetch += ms * 1e-3;
    std::lock_guard<std::mutex> g(mu);
    const double t = now_s();
    for (Req* rp : fetching) {
        Req& r = *rp;
        r.state = r.error ? FAILED : DONE;
        r.t_done = t;
        alloc.release(r.dev_off, r.dev_len);
        r.dev_len = 0;
        if (r.copied) (void)hipEventDestroy(r.copied);
        r.copied = nullptr;
        stats.completed++;
        stats.bytes += r.n;
    }
    fetching.clear();
    cv_done.notify_all();
}

void gc_server::loop() {
    (void)hipSetDevice(ctx->device);
    for (;;) {
        bool any_busy = false;
        for (auto& sl : slot) any_busy |= sl.busy;
        {
            std::unique_lock<std::mutex> lk(mu);
            admit_locked();
            const bool idle = !any_busy && fetching.empty() && finished.empty() && pending.empty() &&
                              copying.empty() && ready.empty();
            if (stop && idle) break;
            if (idle) {
                cv_work.wait_for(lk, std::chrono::milliseconds(2));
                continue;
            }
        }
        bool progressed = false;
        const double t0 = now_s();
        // host inputs that have arrived on the device
        if (!copying.empty()) {
            std::lock_guard<std::mutex> g(mu);
            for (size_t j = 0; j < copying.size();) {
                Req* r = copying[j];
                if (hipEventQuery(r->copied) == hipSuccess) {
                    r->state = READY;
                    ready.push_back(r);
                    ready_items += r->nsub;
                    copying[j] = copying.back();
                    copying.pop_back();
                    progressed = true;
                } else {
                    j++;
                }
            }
        }
        for (auto& sl : slot)
            if (sl.busy && hipEventQuery(sl.ev_done) == hipSuccess) {
                tick_ended(sl);
                progressed = true;
            }
        {
            int running = 0;
            for (auto& sl : slot) running += sl.busy ? 1 : 0;
            for (auto& sl : slot) {
                if (sl.busy) continue;
                if (running >= kBacklogTicks) {
                    std::lock_guard<std::mutex> g(mu);
                    if (!pending.empty() || ready_items > kBacklogItems) break;
                }
                if (launch_tick(sl)) {
                    progressed = true;
                    running++;
                }
            }
        }
        if (!fetching.empty() && hipEventQuery(fetch_done) == hipSuccess) {
            complete_fetch();
            progressed = true;
        }
        if (fetching.empty() && !finished.empty()) {
            finish_fetch(finished.front());
            finished.pop_front();
            progressed = true;
        }
        if (progressed)
            stats.t_host += now_s() - t0;
        else
            std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
}

namespace {

void destroy_server(gc_server* s) {
    for (auto& sl : s->slot) {
        if (sl.st) (void)hipStreamSynchronize(sl.st);
        for (auto& ev : sl.ev)
            if (ev) (void)hipEventDestroy(ev);
        if (sl.ev_done) (void)hipEventDestroy(sl.ev_done);
        if (sl.hdr) (void)hipHostFree(sl.hdr);
        if (sl.stage) (void)hipHostFree(sl.stage);
        if (sl.err) (void)hipHostFree(sl.err);
        sl.items.release();
        sl.cost.release();
        sl.order.release();
        sl.queue.release();
        if (sl.st) (void)hipStreamDestroy(sl.st);
    }
    if (s->fetch_stream) {
        (void)hipStreamSynchronize(s->fetch_stream);
        (void)hipStreamDestroy(s->fetch_stream);
    }
    (void)hipStreamSynchronize(s->ctx->stream2);
    for (auto& kv : s->reqs)
        if (kv.second->copied) (void)hipEventDestroy(kv.second->copied);
    if (s->fetch_start) (void)hipEventDestroy(s->fetch_start);
    if (s->fetch_done) (void)hipEventDestroy(s->fetch_done);
    if (s->meta_host) (void)hipHostFree(s->meta_host);
    if (s->zero_host) (void)hipHostFree(s->zero_host);
    if (s->counts_zone) (void)hipHostFree(s->counts_zone);
    if (s->pool) (void)hipFree(s->pool);
    delete s;
}

int64_t submit(gc_server* s, std::unique_ptr<Req> r) {
    r->nsub = (uint32_t)((r->n + GC_SUB_BLOCK - 1) / GC_SUB_BLOCK);
    layout(*r);
    if (r->dev_len > s->budget) return fail(GC_EBUDGET, "request needs more device memory than the whole HBM budget");
    if (r->nsub > s->max_batch_items) return fail(GC_EBUDGET, "request has more sub-blocks than a tick holds");
    r->t_submit = now_s();
    std::lock_guard<std::mutex> g(s->mu);
    if (s->stop) return fail(GC_EINVAL, "server stopped");
    r->id = s->next_id++;
    const uint64_t id = r->id;
    if (r->n == 0) {
        r->state = DONE;
        r->t_admit = r->t_start = r->t_done = r->t_submit;
        s->reqs[id] = std::move(r);
        s->cv_done.notify_all();
        return (int64_t)id;
    }
    s->pending.push_back(id);
    s->reqs[id] = std::move(r);
    s->cv_work.notify_one();
    return (int64_t)id;
}

}  // namespace

extern "C" {

gc_server* gc_server_start(gc_ctx* ctx, uint64_t hbm_budget, uint32_t max_batch_items) {
    if (!ctx || hbm_budget < (1u << 20)) {
        set_error("gc_server_start: null context or budget below 1 MiB");
        return nullptr;
    }
    (void)hipSetDevice(ctx->device);
    gc_server* s = new gc_server();
    s->ctx = ctx;
    s->budget = up(hbm_budget, kAlign);
    s->max_batch_items = max_batch_items ? max_batch_items : 65536;
    // every Zstd sub-block in flight holds more than 6 bytes per position of the budget
    s->zone_len = s->budget / (6ull * GC_SUB_BLOCK) + s->max_batch_items;
    hipError_t e = hipMalloc((void**)&s->pool, s->budget + 64);
    if (e == hipSuccess) e = hipHostMalloc((void**)&s->counts_zone, s->zone_len * 4, 0);
    if (e == hipSuccess) e = hipHostMalloc((void**)&s->zero_host, 256, 0);
    if (e == hipSuccess) e = hipEventCreate(&s->fetch_start);
    if (e == hipSuccess) e = hipEventCreate(&s->fetch_done);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->fetch_stream, hipStreamNonBlocking);
    const size_t meta_bytes = (size_t)s->max_batch_items * 12 + 64;
    s->meta_cap = meta_bytes;
    if (e == hipSuccess) e = hipHostMalloc((void**)&s->meta_host, meta_bytes, 0);
    if (e != hipSuccess) {
        hip_fail(e, "gc_server_start");
        destroy_server(s);
        return nullptr;
    }
    memset(s->zero_host, 0, 256);
    memset(s->counts_zone, 0, s->zone_len * 4);
    s->alloc.reset(s->budget);
    s->zone.reset(s->zone_len);
    for (auto& sl : s->slot) {
        int rc = s->setup_slot(sl);
        if (rc) {
            std::string keep = gc_last_error();
            destroy_server(s);
            set_error(keep);
            return nullptr;
        }
    }
    s->th = std::thread([s] { s->loop(); });
    return s;
}

int64_t gc_server_submit(gc_server* s, const uint8_t* data, uint64_t n, uint32_t algo, uint32_t level) {
    return gc_server_submit_into(s, data, n, algo, level, nullptr, 0, nullptr);
}

int64_t gc_server_submit_into(gc_server* s, const uint8_t* data, uint64_t n, uint32_t algo, uint32_t level, void* out,
                              uint64_t out_bytes, uint32_t* seq_counts) {
    if (!s || (!data && n) || algo > GC_ALGO_LZMA2) return fail(GC_EINVAL, "gc_server_submit: bad argument");
    if (out && algo == GC_ALGO_ZSTD && !seq_counts)
        return fail(GC_EINVAL, "gc_server_submit_into: Zstd results need seq_counts");
    if (out && algo == GC_ALGO_LZMA2 && out_bytes < n * 4)
        return fail(GC_ESIZE, "gc_server_submit_into: LZMA2 results need 4 bytes per input byte");
    auto r = std::make_unique<Req>();
    r->dst = out;
    r->dst_cap = out_bytes;
    if (out) {
        // pinned host memory (hipHostMalloc) has a device address: results are written there by
        // the kernels; other memory gets DMA copies
        hipPointerAttribute_t at;
        if (hipPointerGetAttributes(&at, out) == hipSuccess && at.type == hipMemoryTypeHost && at.devicePointer)
            r->dst_dev = at.hostPointer ? (uint8_t*)at.devicePointer + ((uint8_t*)out - (uint8_t*)at.hostPointer)
                                        : at.devicePointer;
        (void)hipGetLastError();  // pageable memory reports an error here: not one of ours
    }
    r->dst_counts = seq_counts;
    r->data = data;
    r->n = n;
    r->algo = algo;
    r->level = level;
    return submit(s, std::move(r));
}

int64_t gc_server_submit_dev(gc_server* s, const uint8_t* d_input, uint64_t n, uint32_t algo, uint32_t level) {
    if (!s || (!d_input && n) || algo > GC_ALGO_LZMA2) return fail(GC_EINVAL, "gc_server_submit_dev: bad argument");
    if ((uintptr_t)d_input & 3u) return fail(GC_EINVAL, "gc_server_submit_dev: d_input must be 4-byte aligned");
    auto r = std::make_unique<Req>();
    r->d_in = d_input;
    r->dev_result = true;
    r->n = n;
    r->algo = algo;
    r->level = level;
    return submit(s, std::move(r));
}

int gc_server_submit_dev_many(gc_server* s, const uint8_t* const* d_inputs, const uint64_t* sizes,
                              const uint32_t* algos, const uint32_t* levels, uint32_t count, int64_t* ids) {
    if (!s || (count && (!d_inputs || !sizes || !algos || !levels || !ids)))
        return fail(GC_EINVAL, "gc_server_submit_dev_many: null argument");
    // checked and laid out first, then queued under one lock, in order
    std::vector<std::unique_ptr<Req>> rs(count);
    const double t = now_s();
    for (uint32_t i = 0; i < count; i++) {
        if ((!d_inputs[i] && sizes[i]) || algos[i] > GC_ALGO_LZMA2)
            return fail(GC_EINVAL, "gc_server_submit_dev_many: bad request");
        if ((uintptr_t)d_inputs[i] & 3u) return fail(GC_EINVAL, "gc_server_submit_dev_many: d_input must be 4-byte aligned");
        auto r = std::make_unique<Req>();
        r->d_in = d_inputs[i];
        r->dev_result = true;
        r->n = sizes[i];
        r->algo = algos[i];
        r->level = levels[i];
        r->nsub = (uint32_t)((r->n + GC_SUB_BLOCK - 1) / GC_SUB_BLOCK);
        layout(*r);
        if (r->dev_len > s->budget) return fail(GC_EBUDGET, "request needs more device memory than the whole HBM budget");
        if (r->nsub > s->max_batch_items) return fail(GC_EBUDGET, "request has more sub-blocks than a tick holds");
        r->t_submit = t;
        rs[i] = std::move(r);
    }
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (s->stop) return fail(GC_EINVAL, "server stopped");
        bool any_done = false;
        for (uint32_t i = 0; i < count; i++) {
            Req* r = rs[i].get();
            r->id = s->next_id++;
            ids[i] = (int64_t)r->id;
            if (r->n == 0) {
                r->state = DONE;
                r->t_admit = r->t_start = r->t_done = r->t_submit;
                any_done = true;
            } else {
                s->pending.push_back(r->id);
            }
            s->reqs[r->id] = std::move(rs[i]);
        }
        if (any_done) s->cv_done.notify_all();
    }
    s->cv_work.notify_one();
    return GC_OK;
}

int gc_server_wait(gc_server* s, uint64_t id, int timeout_ms, gc_request_info* info) {
    if (!s) return fail(GC_EINVAL, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return fail(GC_EINVAL, "unknown request id");
    Req& r = *it->second;
    auto done = [&] { return r.state == DONE || r.state == FAILED; };
    if (timeout_ms < 0)
        s->cv_done.wait(lk, done);
    else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), done))
        return 1;  // still running
    if (info) {
        info->size = r.n;
        info->algo = r.algo;
        info->n_sub = r.nsub;
        info->n_sequences = r.fetch_total;
        info->t_submit = r.t_submit;
        info->t_admit = r.t_admit;
        info->t_start = r.t_start;
        info->t_done = r.t_done;
    }
    if (r.state == FAILED && r.error == GC_ESIZE)
        return fail(GC_ESIZE, "request failed: its sequences exceed the result buffer given at submit");
    if (r.state == FAILED) return fail(r.error ? r.error : GC_EHIP, "request failed");
    return GC_OK;
}

int gc_server_result_zstd(gc_server* s, uint64_t id, gc_zstd_sequence* seqs, uint64_t cap, uint32_t* seq_counts) {
    if (!s || !seq_counts) return fail(GC_EINVAL, "null argument");
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end() || it->second->state != DONE || it->second->algo != GC_ALGO_ZSTD)
        return fail(GC_EINVAL, "no finished Zstd request with that id");
    Req& r = *it->second;
    if (r.dst) return fail(GC_EINVAL, "the request's results went to the buffers given at submit");
    if (r.dev_result) return fail(GC_EINVAL, "the request's results stay on the device (gc_server_result_dev)");
    if (r.seqs.size() > cap) return fail(GC_ESIZE, "sequence capacity below the request's sequence count");
    memcpy(seq_counts, r.counts.data(), r.counts.size() * 4);
    if (!r.seqs.empty()) memcpy(seqs, r.seqs.data(), r.seqs.size() * sizeof(gc_zstd_sequence));
    return GC_OK;
}

int gc_server_result_lzma2(gc_server* s, uint64_t id, uint32_t* packed) {
    if (!s || !packed) return fail(GC_EINVAL, "null argument");
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end() || it->second->state != DONE || it->second->algo != GC_ALGO_LZMA2)
        return fail(GC_EINVAL, "no finished LZMA2 request with that id");
    if (it->second->dst) return fail(GC_EINVAL, "the request's results went to the buffers given at submit");
    if (it->second->dev_result)
        return fail(GC_EINVAL, "the request's results stay on the device (gc_server_result_dev)");
    if (it->second->n) memcpy(packed, it->second->packed.data(), it->second->n * 4);
    return GC_OK;
}

int gc_server_result_dev(gc_server* s, uint64_t id, void** d_results, const uint32_t** seq_counts) {
    if (!s || !d_results) return fail(GC_EINVAL, "null argument");
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end() || it->second->state != DONE || !it->second->dev_result)
        return fail(GC_EINVAL, "no finished device-result request with that id");
    Req& r = *it->second;
    *d_results = r.n ? (void*)(s->pool + r.dev_off + r.out_off) : nullptr;
    if (seq_counts) *seq_counts = r.algo == GC_ALGO_ZSTD ? r.counts.data() : nullptr;
    return GC_OK;
}

int gc_server_release(gc_server* s, uint64_t id) {
    if (!s) return fail(GC_EINVAL, "null server");
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return fail(GC_EINVAL, "unknown request id");
    Req& r = *it->second;
    if (r.state != DONE && r.state != FAILED) return fail(GC_EINVAL, "request still running");
    if (r.dev_len) {  // a device-result request holds its region until now
        s->alloc.release(r.dev_off, r.dev_len);
        r.dev_len = 0;
        s->cv_work.notify_one();
    }
    s->reqs.erase(it);
    return GC_OK;
}

int gc_server_wait_release_many(gc_server* s, const uint64_t* ids, uint32_t count, gc_request_info* infos,
                                int* status) {
    if (!s || (count && (!ids || !infos || !status))) return fail(GC_EINVAL, "null argument");
    int first_err = GC_OK;
    for (uint32_t i = 0; i < count; i++) {
        const int rc = gc_server_wait(s, ids[i], -1, &infos[i]);
        status[i] = rc;
        if (rc < 0 && first_err == GC_OK) first_err = rc;
        if (rc == GC_EINVAL) continue;  // unknown id: nothing to release
        const int rr = gc_server_release(s, ids[i]);
        if (rr < 0 && first_err == GC_OK) first_err = rr;
    }
    return first_err;
}

int gc_server_get_stats(gc_server* s, gc_server_stats* out) {
    if (!s || !out) return fail(GC_EINVAL, "null argument");
    std::lock_guard<std::mutex> g(s->mu);
    *out = s->stats;
    out->hbm_in_use = s->alloc.used;
    out->pending = s->pending.size();
    return GC_OK;
}

int gc_server_stop(gc_server* s) {
    if (!s) return GC_OK;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->th.joinable()) s->th.join();
    destroy_server(s);
    return GC_OK;
}

}  // extern "C"
