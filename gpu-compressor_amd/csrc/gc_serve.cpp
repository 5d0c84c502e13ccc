// gc_serve.cpp -- served form: a request loop with continuous batching.
//
// Clients submit match-finding requests (Zstd L1/L2 or LZMA2, any size) from
// any thread.  One server thread:
//   1. admits pending requests in arrival order while their device memory
//      (input, candidates, Zstd info, outputs) fits the HBM budget,
//   2. copies each admitted input to the device on the copy stream,
//   3. whenever no batch is running, launches every request whose input has
//      arrived as one batch (one launch per kernel over all their sub-blocks),
//   4. when a batch ends, starts the next one at once and meanwhile brings the
//      finished requests' results back on a second copy stream (sequences
//      packed by gc_gather_seqs first), straight into the caller's buffers
//      when it gave some (gc_server_submit_into), frees their memory and marks
//      them done.
// Nothing on the compute stream uses a DMA engine (those carry the inputs in):
// the batch's work items are read from pinned memory by a copy kernel, and
// the parse writes its sequence counts straight into pinned memory, so the
// host reads them when the batch's event completes.  The server thread only polls events: a batch that ends
// while results of an earlier one are still being copied waits in `finished`
// and the device goes on with the next batch.
// So requests that arrive while others run are admitted as memory frees up
// and join the next batch; a result does not depend on what it was batched
// with (sub-blocks are independent).  No reference counterpart: the
// reference processes one chunk at a time (pipeline.rs:138-175).
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
const uint64_t kAlign = 256;

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

// First-fit allocator over [0, size) of one device pool.
struct PoolAlloc {
    uint64_t size = 0;
    std::map<uint64_t, uint64_t> free_;  // offset -> length
    uint64_t used = 0;
    void reset(uint64_t s) {
        size = s;
        free_.clear();
        free_[0] = s;
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
    uint64_t id;
    const uint8_t* data;
    uint64_t n;
    uint32_t algo, level, nsub;
    uint64_t dev_off = 0, dev_len = 0;            // the request's region of the pool
    uint64_t in_off = 0, cand_off = 0, info_off = 0, out_off = 0;  // byte offsets in the region
    uint32_t item0 = 0;        // first Zstd count index of the request in its batch
    void* dst = nullptr;       // caller's result buffer (gc_server_submit_into), else results kept here
    void* dst_dev = nullptr;   // the same buffer as the device sees it, when it is pinned host memory
    uint64_t dst_cap = 0;      // its size in bytes
    uint32_t* dst_counts = nullptr;
    hipEvent_t copied = nullptr;
    State state = PENDING;
    int error = 0;
    double t_submit = 0, t_admit = 0, t_start = 0, t_done = 0;
    std::vector<uint32_t> counts;
    std::vector<gc_zstd_sequence> seqs;
    std::vector<uint32_t> packed;
    uint64_t fetch_total = 0;  // sequences packed for the copy back
};

// Device bytes a request needs (every part 256-byte aligned).  For Zstd the
// candidate and info regions are contiguous and together hold at least
// 21845 x 16 bytes per sub-block: the packed sequences go there at the end.
void layout(Req& r) {
    const uint64_t ns = r.nsub, slots = ns * GC_SUB_BLOCK;
    uint64_t o = 0;
    r.in_off = o;
    o += up(r.n + 16, kAlign);
    r.cand_off = o;
    o += up(slots * 2, kAlign);
    if (r.algo == GC_ALGO_ZSTD) {
        r.info_off = o;
        o += up(slots * 4, kAlign);
        r.out_off = o;  // sequence slots, later packed in place
        o += up(ns * kMaxSeq * sizeof(gc_zstd_sequence), kAlign);
    } else {
        r.out_off = o;  // one packed match per position
        o += up(slots * 4, kAlign);
    }
    r.dev_len = o;
}

}  // namespace

struct gc_server {
    gc_ctx* ctx = nullptr;
    uint8_t* pool = nullptr;
    uint64_t budget = 0;
    PoolAlloc alloc;
    std::thread th;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    bool stop = false;
    uint64_t next_id = 1;
    std::deque<uint64_t> pending;                        // arrival order
    std::unordered_map<uint64_t, std::unique_ptr<Req>> reqs;
    // batch state (server thread only)
    // Requests of the running, finished and fetching batches.  Held by address: `reqs` is
    // changed by client threads (under `mu`), and a request in a batch cannot be released
    // until it is done.
    std::vector<Req*> running, fetching;
    std::deque<std::vector<Req*>> finished;  // batches done on the device, results not yet copied
    hipStream_t fetch_stream = nullptr;  // results back to the host (the context's stream2 brings inputs in)
    hipEvent_t batch_start = nullptr, batch_done = nullptr, fetch_start = nullptr, fetch_done = nullptr;
    DevBuf items;                     // the running batch's work items
    DevBuf cost, order;               // Zstd items: parse cost and the parse's order (costliest first)
    uint32_t* counts_host = nullptr;  // pinned: per-batch sequence counts, written by the parse
    size_t counts_host_cap = 0;       // in u32
    gc_item* items_host = nullptr;    // pinned staging of `items`
    size_t items_host_cap = 0;        // in items
    uint8_t* meta_host = nullptr; // pinned: counts and offsets for the gathers
    size_t meta_cap = 0;
    std::vector<gc_item> host_items;
    gc_server_stats stats{};
    uint32_t max_batch_items = 0;

    void loop();
    void admit_locked();
    bool launch_batch();
    void batch_ended();
    void finish_batch(std::vector<Req*>& done);
    void complete_fetch();
};

void gc_server::admit_locked() {
    while (!pending.empty()) {
        Req& r = *reqs[pending.front()];
        if (!alloc.alloc(r.dev_len, &r.dev_off)) break;  // in arrival order: wait for memory
        pending.pop_front();
        r.t_admit = now_s();
        r.state = COPYING;
        hipError_t e = hipMemcpyAsync(pool + r.dev_off + r.in_off, r.data, r.n, hipMemcpyHostToDevice, ctx->stream2);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&r.copied, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventRecord(r.copied, ctx->stream2);
        if (e != hipSuccess) {
            r.state = FAILED;
            r.error = GC_EHIP;
            alloc.release(r.dev_off, r.dev_len);
            r.t_done = now_s();
            cv_done.notify_all();
        }
    }
}

bool gc_server::launch_batch() {
    // every request whose input has arrived
    std::vector<Req*> batch;
    {
        std::lock_guard<std::mutex> g(mu);
        for (auto& kv : reqs) {
            Req& r = *kv.second;
            if (r.state == COPYING && hipEventQuery(r.copied) == hipSuccess) r.state = READY;
        }
        uint32_t n_items = 0;
        for (auto& kv : reqs) {
            Req& r = *kv.second;
            if (r.state != READY) continue;
            if (n_items + r.nsub > max_batch_items && !batch.empty()) break;
            batch.push_back(&r);
            n_items += r.nsub;
        }
    }
    if (batch.empty()) return false;
    // items: LZMA2 first, then Zstd (each kernel gets a contiguous slice)
    host_items.clear();
    uint32_t n_l = 0;
    for (int pass = 0; pass < 2; pass++)
        for (Req* r : batch) {
            if ((pass == 0) != (r->algo == GC_ALGO_LZMA2)) continue;
            const uint64_t base = r->dev_off;
            for (uint32_t s = 0; s < r->nsub; s++) {
                gc_item it{};
                const uint64_t start = (uint64_t)s * GC_SUB_BLOCK;
                it.src_off = base + r->in_off + start;
                it.avail = budget + 64 - it.src_off;  // the pool is readable to its end
                it.n = (uint32_t)std::min<uint64_t>(GC_SUB_BLOCK, r->n - start);
                it.cpos = (base + r->cand_off) / 2 + start;
                it.out_span = GC_SUB_BLOCK;
                if (r->algo == GC_ALGO_ZSTD) {
                    it.opos = (base + r->info_off) / 4 + start;
                    it.seq_pos = (base + r->out_off) / sizeof(gc_zstd_sequence) + (uint64_t)s * kMaxSeq;
                    it.count_pos = 0;  // set below: counts are per request in the pool
                    it.depth = zstd_depth_for_level(r->level);
                } else {
                    it.opos = (base + r->out_off) / 4 + start;
                    it.depth = GC_LZMA2_DEPTH;
                }
                host_items.push_back(it);
            }
            if (pass == 0) n_l += r->nsub;
        }
    // Zstd counts: one u32 per sub-block in the server's counts array, in
    // batch order
    {
        uint32_t k = n_l, c = 0;
        for (Req* r : batch) {
            if (r->algo != GC_ALGO_ZSTD) continue;
            r->item0 = c;
            for (uint32_t s2 = 0; s2 < r->nsub; s2++) host_items[k++].count_pos = c++;
        }
    }
    const uint32_t n_all = (uint32_t)host_items.size(), n_z = n_all - n_l;
    std::lock_guard<std::mutex> g(ctx->mu);
    int rc = items.ensure(n_all * sizeof(gc_item));
    if (!rc) rc = cost.ensure((size_t)n_z * 4 + 4);
    if (!rc) rc = order.ensure((size_t)n_z * 4 + 4);
    // pinned staging (no batch runs now, so the last batch's reads and writes are done)
    if (!rc && n_z > counts_host_cap) {
        (void)hipHostFree(counts_host);
        counts_host_cap = 0;
        if (hipHostMalloc((void**)&counts_host, (size_t)n_z * 4, 0) != hipSuccess) {
            counts_host = nullptr;
            rc = fail(GC_ENOMEM, "gc_server: pinned counts buffer");
        } else {
            counts_host_cap = n_z;
        }
    }
    if (!rc && n_all > items_host_cap) {
        (void)hipHostFree(items_host);
        items_host_cap = 0;
        if (hipHostMalloc((void**)&items_host, (size_t)n_all * sizeof(gc_item), 0) != hipSuccess) {
            items_host = nullptr;
            rc = fail(GC_ENOMEM, "gc_server: pinned item buffer");
        } else {
            items_host_cap = n_all;
        }
    }
    hipError_t e = hipSuccess;
    if (!rc) {
        memcpy(items_host, host_items.data(), (size_t)n_all * sizeof(gc_item));
        e = hipEventRecord(batch_start, ctx->stream);
        if (e == hipSuccess)
            rc = launch_copy_u32(ctx, ctx->stream, (const uint32_t*)dev_addr(items_host), (uint32_t*)items.p,
                                 (uint64_t)n_all * sizeof(gc_item) / 4);
    }
    const gc_item* all = (const gc_item*)items.p;
    {
        std::lock_guard<std::mutex> g2(mu);  // clients read the state in gc_server_wait
        const double t = now_s();
        for (Req* r : batch) {
            r->state = RUNNING;
            r->t_start = t;
        }
    }
    if (!rc && e == hipSuccess) {
        rc = launch_cand(ctx, ctx->stream, pool, all, n_all, n_all, (uint16_t*)pool);
        if (!rc && n_l)
            rc = launch_lzma2_len(ctx, ctx->stream, pool, all, n_l, n_l, (const uint16_t*)pool, (uint32_t*)pool,
                                  nullptr, nullptr, 1);
        if (!rc && n_z)
            rc = launch_zstd_info(ctx, ctx->stream, pool, all + n_l, n_z, n_z, (const uint16_t*)pool, (uint32_t*)pool,
                                  4, (uint32_t*)cost.p);
        if (!rc && n_z) rc = launch_order(ctx, ctx->stream, (const uint32_t*)cost.p, n_z, (uint32_t*)order.p);
        if (!rc && n_z)
            rc = launch_zstd_parse(ctx, ctx->stream, pool, all + n_l, n_z, (n_z + 3) / 4, (const uint32_t*)pool,
                                   pool, GC_SEQ_FORMAT_ZSTD, nullptr, dev_addr(counts_host), nullptr, kMaxSeq,
                                   (const uint32_t*)order.p);
        if (!rc && e == hipSuccess) e = hipEventRecord(batch_done, ctx->stream);
    }
    running = batch;
    if (rc || e != hipSuccess) {
        std::lock_guard<std::mutex> g2(mu);
        for (Req* r : batch) {
            r->state = FAILED;
            r->error = rc ? rc : GC_EHIP;
            r->t_done = now_s();
            alloc.release(r->dev_off, r->dev_len);
            if (r->copied) (void)hipEventDestroy(r->copied);
            r->copied = nullptr;
        }
        running.clear();
        cv_done.notify_all();
        return false;
    }
    stats.batches++;
    stats.items += n_all;
    return true;
}

// The running batch ended on the device: take its sequence counts (already
// in pinned memory) and queue it for the copy back.
void gc_server::batch_ended() {
    float ms = 0;
    if (hipEventElapsedTime(&ms, batch_start, batch_done) == hipSuccess) stats.t_batches += ms * 1e-3;
    for (Req* rp : running) {
        Req& r = *rp;
        if (r.algo == GC_ALGO_ZSTD) r.counts.assign(counts_host + r.item0, counts_host + r.item0 + r.nsub);
    }
    finished.emplace_back();
    finished.back().swap(running);
}

// Bring a finished batch's results back on the fetch stream (no other fetch
// is in flight).  Zstd: each request's used sequence slots are packed by
// gc_gather_seqs (counts and offsets staged in `meta`) and copied.
void gc_server::finish_batch(std::vector<Req*>& done) {
    fetching.swap(done);
    done.clear();
    std::lock_guard<std::mutex> g(ctx->mu);
    hipStream_t cs = fetch_stream;
    (void)hipEventRecord(fetch_start, cs);
    uint32_t n_z = 0;
    for (Req* r : fetching)
        if (r->algo == GC_ALGO_ZSTD) n_z += r->nsub;
    // meta: counts (u32) for every Zstd item, then offsets (u64)
    const size_t off_bytes = up((size_t)n_z * 4, 8);
    if (off_bytes + (size_t)n_z * 8 > meta_cap) {  // a batch larger than any before
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
        std::lock_guard<std::mutex> g2(mu);
        for (Req* r : fetching) r->state = FETCHING;
    }
    for (Req* rp : fetching) {
        Req& r = *rp;
        if (r.algo != GC_ALGO_ZSTD || !meta_host) continue;
        uint64_t tot = 0;
        for (uint32_t s2 = 0; s2 < r.nsub; s2++) {
            mc[r.item0 + s2] = r.counts[s2];
            mo[r.item0 + s2] = tot;
            tot += r.counts[s2];
        }
        r.fetch_total = tot;
        if (r.dst) {
            memcpy(r.dst_counts, r.counts.data(), (size_t)r.nsub * 4);
            if (tot * sizeof(gc_zstd_sequence) > r.dst_cap) r.error = GC_ESIZE;
        }
    }
    // the gathers read the counts and offsets from pinned memory (no earlier fetch is still reading it)
    uint8_t* meta_dev = dev_addr(meta_host);
    for (Req* rp : fetching) {
        Req& r = *rp;
        if (r.error) continue;
        uint8_t* base = pool + r.dev_off;
        // pinned caller memory is written by the gather / copy kernels directly (over PCIe, on a
        // few hundred wavefronts); anything else is packed on the device and copied by DMA
        if (r.algo == GC_ALGO_ZSTD) {
            if (!r.dst) r.seqs.resize(r.fetch_total);
            if (!r.fetch_total) continue;
            const uint32_t* mcounts = (const uint32_t*)meta_dev + r.item0;
            const uint64_t* moffs = (const uint64_t*)(meta_dev + off_bytes) + r.item0;
            if (r.dst_dev) {
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
        {
            std::unique_lock<std::mutex> lk(mu);
            admit_locked();
            bool idle = running.empty() && fetching.empty() && finished.empty() && pending.empty();
            bool any_copying = false;
            for (auto& kv : reqs)
                if (kv.second->state == COPYING || kv.second->state == READY) any_copying = true;
            if (stop && idle && !any_copying) break;
            if (idle && !any_copying) {
                cv_work.wait_for(lk, std::chrono::milliseconds(2));
                continue;
            }
        }
        bool progressed = false;
        double t0 = now_s();
        if (!running.empty() && hipEventQuery(batch_done) == hipSuccess) {
            batch_ended();
            progressed = true;
        }
        if (running.empty()) progressed |= launch_batch();  // the device goes on at once
        if (!fetching.empty() && hipEventQuery(fetch_done) == hipSuccess) {
            complete_fetch();
            progressed = true;
        }
        if (fetching.empty() && !finished.empty()) {
            finish_batch(finished.front());
            finished.pop_front();
            progressed = true;
        }
        if (progressed)
            stats.t_host += now_s() - t0;
        else
            std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
}

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
    hipError_t e = hipMalloc((void**)&s->pool, s->budget + 64);
    if (e == hipSuccess) e = hipEventCreate(&s->batch_start);
    if (e == hipSuccess) e = hipEventCreate(&s->batch_done);
    if (e == hipSuccess) e = hipEventCreate(&s->fetch_start);
    if (e == hipSuccess) e = hipEventCreate(&s->fetch_done);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->fetch_stream, hipStreamNonBlocking);
    const size_t meta_bytes = (size_t)s->max_batch_items * 12 + 64;
    s->meta_cap = meta_bytes;
    if (e == hipSuccess) e = hipHostMalloc((void**)&s->meta_host, meta_bytes, 0);
    if (e != hipSuccess) {
        hip_fail(e, "gc_server_start");
        if (s->pool) (void)hipFree(s->pool);
        if (s->fetch_stream) (void)hipStreamDestroy(s->fetch_stream);
        delete s;
        return nullptr;
    }
    s->alloc.reset(s->budget);
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
    r->nsub = (uint32_t)((n + GC_SUB_BLOCK - 1) / GC_SUB_BLOCK);
    layout(*r);
    if (r->dev_len > s->budget)
        return fail(GC_EBUDGET, "request needs more device memory than the whole HBM budget");
    r->t_submit = now_s();
    std::lock_guard<std::mutex> g(s->mu);
    if (s->stop) return fail(GC_EINVAL, "server stopped");
    r->id = s->next_id++;
    const uint64_t id = r->id;
    if (n == 0) {
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
    memcpy(packed, it->second->packed.data(), it->second->n * 4);
    return GC_OK;
}

int gc_server_release(gc_server* s, uint64_t id) {
    if (!s) return fail(GC_EINVAL, "null server");
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return fail(GC_EINVAL, "unknown request id");
    if (it->second->state != DONE && it->second->state != FAILED) return fail(GC_EINVAL, "request still running");
    s->reqs.erase(it);
    return GC_OK;
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
    (void)hipStreamSynchronize(s->ctx->stream);
    (void)hipStreamSynchronize(s->ctx->stream2);
    (void)hipStreamSynchronize(s->fetch_stream);
    (void)hipStreamDestroy(s->fetch_stream);
    if (s->counts_host) (void)hipHostFree(s->counts_host);
    if (s->items_host) (void)hipHostFree(s->items_host);
    if (s->batch_start) (void)hipEventDestroy(s->batch_start);
    if (s->fetch_start) (void)hipEventDestroy(s->fetch_start);
    for (auto& kv : s->reqs)
        if (kv.second->copied) (void)hipEventDestroy(kv.second->copied);
    if (s->batch_done) (void)hipEventDestroy(s->batch_done);
    if (s->fetch_done) (void)hipEventDestroy(s->fetch_done);
    s->items.release();
    s->cost.release();
    s->order.release();
    if (s->meta_host) (void)hipHostFree(s->meta_host);
    if (s->pool) (void)hipFree(s->pool);
    delete s;
    return GC_OK;
}

}  // extern "C"
