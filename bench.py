"""Match-find throughput and latency on one MI355X: the served and batched forms against the plain
per-chunk GPU form, at identical results.

Default (no --workload): the serving workload.  Requests arrive open-loop (Poisson at 500, 2000 and
8000 req/s, then all at once for the saturation passes that `value` is measured on), sizes log-uniform
1 KiB-64 MiB, content from the four-kind mix (a 1 GiB pool), 50% Zstd L1 / 25% Zstd L2 / 25% LZMA2 L1,
under a 64 GiB HBM budget.  Inputs are resident in HBM before the timed region (each request is a window
of the pool, read in place) and results stay in HBM until the client releases them; latency runs from
each request's scheduled arrival.  A step is one saturation pass over the same request set (K timed after
W warm-up), from the requests' scheduled arrival to the last completion.  The plain form is one synchronous call per request on the same device-resident inputs.
Every served result of a verification pass is compared with the plain form's, and a sample with the CPU
oracle.  The batched configurations (1 GiB Zstd L1+L2 in 16 MiB chunks; 640 MiB LZMA2 in 128 MiB chunks;
4096 inputs, persistent device-queue form replayed from a hipGraph) are extra keys, each with its own
roofline.

    python bench.py [--steps K] [--warmup W] [--workload serve|zstd|lzma2|mixed|lzma2_coder|containers]
                    [--no-extras] [--no-cpu]

`--workload lzma2_coder` times the GPU LZMA2 coder (the 640 MiB of the LZMA2 configuration, streams byte-equal
to the host coder's) and NVLZ end to end with the GPU and the host coder.  `--workload containers` times the
NVZS and NVLZ containers written on the GPU from host memory, in a process that runs nothing else.

Prints one JSON line; --gpus N > 1 exits with a message: this path runs on one GPU.
"""
import argparse
import ctypes as C
import hashlib
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gpu-compressor_amd"))

import gcompress  # noqa: E402
from gcompress import workloads as W  # noqa: E402

KiB, MiB, GiB = 1 << 10, 1 << 20, 1 << 30
HBM_PEAK_GBS = 8000.0
MAX_SEQ = 65536 // 3
METRIC = "match-find input GB/s at equal ratio, batched/served vs plain; p50/p99 latency"
ZSTD, LZMA2 = gcompress.ALGO_ZSTD, gcompress.ALGO_LZMA2
LEVEL_DEPTH = {1: 16, 2: 64}


class Hip:
    """HIP events on an arbitrary stream (the library's own), via the runtime's C API."""

    def __init__(self):
        self.L = C.CDLL("libamdhip64.so")
        self.L.hipEventCreate.argtypes = [C.c_void_p]
        self.L.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.L.hipEventSynchronize.argtypes = [C.c_void_p]
        self.L.hipEventElapsedTime.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        self.L.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]

    def event(self):
        e = C.c_void_p()
        assert self.L.hipEventCreate(C.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.L.hipEventRecord(e, stream) == 0

    def elapsed_ms(self, a, b):
        assert self.L.hipEventSynchronize(b) == 0
        ms = C.c_float()
        assert self.L.hipEventElapsedTime(C.byref(ms), a, b) == 0
        return ms.value


def host_cores():
    return {"os_cpu_count": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}


def threads_all():
    return len(os.sched_getaffinity(0))


def hbm_traffic_per_input_byte(workload: str, kernel: str):
    """HBM bytes per input byte of one launch of `kernel`, from the committed rocprofv3 counter summary
    profiles/<workload>_hbm_traffic.md (FETCH_SIZE + WRITE_SIZE per dispatch, each scaled by what the
    counter reads for a device copy of known size), or None when there is none."""
    path = os.path.join(ROOT, "profiles", f"{workload}_hbm_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) == 3 and parts[0] == "bytes_per_input_byte:" and parts[1] == kernel:
            return float(parts[2])
    return None


def digest(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def zstd_digest_from_slots(slots, counts):
    mask = np.arange(slots.shape[1])[None, :] < counts[:, None]
    return digest(counts.astype(np.uint32), np.ascontiguousarray(slots[mask][:, :3]))


def oracle_digest(x, algo, level, threads):
    sys.path.insert(0, ROOT)
    import oracle

    if algo == LZMA2:
        return digest(oracle.lzma2_packed(x, threads=threads))
    seq, sc, _, _ = oracle.zstd_match_find(x, depth=LEVEL_DEPTH.get(level, 1), threads=threads)
    mask = np.arange(seq.shape[1])[None, :] < sc[:, None]
    r = seq[mask]
    return digest(sc.astype(np.uint32), np.stack([r[:, 2], r[:, 0], r[:, 1] + 3], axis=1).astype(np.uint32))


# ---------------------------------------------------------------------------------------------------
# Serving workload
# ---------------------------------------------------------------------------------------------------
KINDS = [(ZSTD, 1), (ZSTD, 2), (LZMA2, 1)]
KIND_P = [0.5, 0.25, 0.25]


def draw_requests(rng, nreq, rate, pool_size):
    """(scheduled arrival offsets in s, [(offset, size, algo, level)]): Poisson arrivals (rate 0: all at
    once), sizes log-uniform 1 KiB-64 MiB, windows of the pool at 4-byte aligned offsets."""
    gaps = rng.exponential(1.0 / rate, nreq) if rate else np.zeros(nreq)
    sizes = np.exp(rng.uniform(np.log(KiB), np.log(64 * MiB), nreq)).astype(np.int64)
    offs = (rng.integers(0, pool_size - sizes + 1) // 4) * 4
    kinds = rng.choice(3, nreq, p=KIND_P)
    reqs = [(int(offs[i]), int(sizes[i]), *KINDS[kinds[i]]) for i in range(nreq)]
    return np.cumsum(gaps), reqs


def serve_run(srv, d_pool, arrivals, reqs, fetch=None):
    """One open-loop run: a producer submits at the scheduled times, a consumer waits in order, records
    completion and releases (after fetch(i, rid) when given).  Latency from the scheduled arrival.  When
    every request arrives at once and nothing is fetched, the client submits them in one call and reaps
    them in one (each released as soon as it and the requests before it are done), as a native client
    would: a Python call per request costs about 20 us."""
    n = len(reqs)
    if fetch is None and n and not np.any(arrivals):
        batch = gcompress.Server.dev_batch([(d_pool + o, s, algo, lv) for (o, s, algo, lv) in reqs])
        t0 = time.monotonic()
        ids = srv.submit_dev_many(batch)
        infos, status = srv.wait_release_many(ids)
        done = np.array([x.t_done for x in infos[:n]])
        start_t = np.array([x.t_start for x in infos[:n]])
        failed = [(i, f"gc_server_wait returned {int(st)}") for i, st in enumerate(status) if st != 0]
        return {"sched": np.full(n, t0), "done": done, "start": start_t, "t0": t0, "failed": failed}
    ids = [None] * n
    sched = np.zeros(n)
    done = np.zeros(n)
    start_t = np.zeros(n)
    failed = []
    ready = threading.Semaphore(0)
    t0 = time.monotonic() + 0.02

    def producer():
        for i, (o, s, algo, lv) in enumerate(reqs):
            t = t0 + arrivals[i]
            sched[i] = t
            d = t - time.monotonic()
            if d > 0:
                time.sleep(d)
            ids[i] = srv.submit_dev(d_pool + o, s, algo, lv)
            ready.release()

    def consumer():
        for i in range(n):
            ready.acquire()
            try:
                info = srv.wait(ids[i])
                done[i] = info.t_done
                start_t[i] = info.t_start
                if fetch:
                    fetch(i, ids[i])
            except gcompress.GcError as e:
                failed.append((i, str(e)))
                done[i] = time.monotonic()
            srv.release(ids[i])

    tp, tc = threading.Thread(target=producer), threading.Thread(target=consumer)
    tp.start()
    tc.start()
    tp.join()
    tc.join()
    return {"sched": sched, "done": done, "start": start_t, "t0": t0, "failed": failed}


def served_digest(ctx, srv, rid, n, algo):
    nsub = -(-n // 65536)
    p, counts = srv.result_dev(rid, nsub)
    if algo == LZMA2:
        out = np.empty(n, np.uint32)
        ctx.dtoh(out, p)
        return digest(out)
    slots = np.empty((nsub, gcompress.SERVE_ZSTD_SLOT, 4), np.uint32)
    ctx.dtoh(slots, p)
    return zstd_digest_from_slots(slots, counts)


class Plain:
    """The plain per-request GPU form: one synchronous kernel-ABI call per request on its device-resident
    window (gc_zstd_match_find_dev / gc_lzma2_match_find_packed_dev), results into device scratch."""

    def __init__(self, ctx, max_size):
        self.ctx, self.L = ctx, gcompress.lib()
        nsub = -(-max_size // 65536)
        self.d_seq = ctx.alloc(nsub * MAX_SEQ * 16)
        self.d_cnt = ctx.alloc(nsub * 4)
        self.d_pk = ctx.alloc(max_size * 4 + 64)

    def run(self, d_in, s, algo, lv):
        if algo == ZSTD:
            rc = self.L.gc_zstd_match_find_dev(self.ctx.h, d_in, self.d_seq, 0, None, self.d_cnt, None, 65536,
                                               -(-s // 65536), s, LEVEL_DEPTH.get(lv, 1), MAX_SEQ)
        else:
            rc = self.L.gc_lzma2_match_find_packed_dev(self.ctx.h, d_in, self.d_pk, 65536, s, 32)
        assert rc == 0, self.L.gc_last_error()
        self.ctx.sync()

    def digest(self, s, algo):
        if algo == LZMA2:
            out = np.empty(s, np.uint32)
            self.ctx.dtoh(out, self.d_pk)
            return digest(out)
        nsub = -(-s // 65536)
        counts = np.empty(nsub, np.uint32)
        self.ctx.dtoh(counts, self.d_cnt)
        slots = np.empty((nsub, MAX_SEQ, 4), np.uint32)
        self.ctx.dtoh(slots, self.d_seq)
        return zstd_digest_from_slots(slots, counts)

    def close(self):
        for p in (self.d_seq, self.d_cnt, self.d_pk):
            self.ctx.free(p)


def lat_summary(r, reqs):
    lat = (r["done"] - r["sched"]) * 1e3
    nbytes = float(sum(s for (_, s, _, _) in reqs))
    span = r["done"].max() - r["t0"]
    arrive = r["sched"][-1] - r["t0"]
    return {"requests": len(reqs), "failed": len(r["failed"]), "input_bytes": int(nbytes),
            "offered_gbs": round(nbytes / arrive / 1e9, 2) if arrive > 1e-3 else None,  # None: all at once
            "sustained_gbs": round(nbytes / span / 1e9, 3),
            "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3)}


def kernel_roofline(stats0, stats1, wall_s):
    """Top-level roofline of the served form: the dominant kernel (most device time over the ticks) with
    the algorithmic bytes it is responsible for (Zstd parse: N_in + 16 n_seq + 4 n_sub; LZMA2 lengths:
    5 N_in; candidates: N_in; Zstd info: N_in), over its average launch (HIP events around every launch
    of it in the served runs)."""
    names = ["gc_tick_setup", "gc_cand_sq_kernel", "gc_lzma2_len_sq_kernel", "gc_zstd_info_sq_kernel",
             "gc_order_items", "gc_zstd_parse_sq_kernel"]
    d = {k: (np.array(stats1[k]) - np.array(stats0[k])) if isinstance(stats1[k], list) else stats1[k] - stats0[k]
         for k in ("t_kernel", "launches", "zstd_bytes", "zstd_sub_blocks", "zstd_sequences", "lzma2_bytes")}
    zb, lb = int(d["zstd_bytes"]), int(d["lzma2_bytes"])
    share = {1: zb + lb, 2: 5 * lb, 3: zb, 5: zb + 16 * int(d["zstd_sequences"]) + 4 * int(d["zstd_sub_blocks"])}
    inputs_of = {1: zb + lb, 2: lb, 3: zb, 5: zb}  # the input bytes each kernel reads
    dom = max(share, key=lambda k: d["t_kernel"][k])
    launches = int(d["launches"][dom])
    avg_s = float(d["t_kernel"][dom]) / max(launches, 1)
    per_launch = share[dom] / max(launches, 1)
    achieved = per_launch / avg_s / 1e9 if avg_s > 0 else 0.0
    per_byte = hbm_traffic_per_input_byte("serve", names[dom])
    return {"bound": "hbm", "kernel": names[dom], "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": int(per_byte * inputs_of[dom] / max(launches, 1)) if per_byte is not None else None,
            "algorithmic_bytes": int(per_launch), "launches": launches, "avg_launch_ms": round(avg_s * 1e3, 4),
            "kernel_s": {names[k]: round(float(d["t_kernel"][k]), 4) for k in range(6)},
            "kernel_launches": {names[k]: int(d["launches"][k]) for k in range(6)},
            # every served kernel together: algorithmic bytes of the timed passes over their wall time
            "path_algorithmic_gbs": round((share[2] + share[5]) / max(wall_s, 1e-9) / 1e9, 2),
            "device_busy_frac_of_wall": round(float(sum(d["t_kernel"])) / max(wall_s, 1e-9), 3)}


def cpu_baseline_serve(pool, reqs, budget_s):
    """The oracle (CPU restatement of the reference kernels) on the serving stream's first requests, all
    host cores, until about budget_s of work; then one core on a smaller share.  Returns the baseline and
    the oracle digests of the requests it covered."""
    t_all = threads_all()
    digs = {}
    done = 0
    t = time.perf_counter()
    for i, (o, s, algo, lv) in enumerate(reqs):
        digs[i] = oracle_digest(pool[o:o + s], algo, lv, t_all)
        done += s
        if time.perf_counter() - t > budget_s:
            break
    dt = time.perf_counter() - t
    one, one_b = 0.0, 0
    t = time.perf_counter()
    for (o, s, algo, lv) in reqs:
        if one_b + s > 64 * MiB and one_b:
            continue
        oracle_digest(pool[o:o + s], algo, lv, 1)
        one_b += s
        if time.perf_counter() - t > budget_s / 4:
            break
    one = one_b / (time.perf_counter() - t) / 1e9
    base = {"value": round(done / dt / 1e9, 4), "unit": "GB/s", "cores": t_all, "kind": "port",
            "sample": f"the serving stream's first {len(digs)} requests ({done / 1e9:.2f} GB of match-find input, "
                      f"the same request mix), oracle on {t_all} threads",
            "one_core_gbs": round(one, 4), "one_core_sample_bytes": int(one_b), "host": host_cores()}
    return base, digs


def bench_serve(a, hip, ctx, extras):
    L = gcompress.lib()
    print("[bench] serving: pool", file=sys.stderr, flush=True)
    pool = W.config5_pool()
    d_pool = ctx.alloc(pool.size + 256)
    ctx.htod(d_pool, pool)
    rng = np.random.default_rng(5)
    budget = a.budget_gib << 30
    # the saturation request set (all arrive at once); the rates' own draws
    _, sat = draw_requests(rng, a.sat_requests, 0, pool.size)
    counts = {500: 1500, 2000: 4000, 8000: 8000}  # about 3, 2 and 1 s of arrivals
    rate_sets = {r: draw_requests(rng, counts.get(r, 4000), r, pool.size)
                 for r in [int(x) for x in a.rates.split(",") if x]}
    sat_bytes = sum(s for (_, s, _, _) in sat)

    # plain per-request form on the first requests of the saturation set
    plain = Plain(ctx, 64 * MiB)
    k = min(a.plain_requests, len(sat))
    for (o, s, algo, lv) in sat[:8]:
        plain.run(d_pool + o, s, algo, lv)  # warm-up
    t = time.perf_counter()
    for (o, s, algo, lv) in sat[:k]:
        plain.run(d_pool + o, s, algo, lv)
    plain_s = time.perf_counter() - t
    plain_bytes = sum(s for (_, s, _, _) in sat[:k])
    plain_gbs = plain_bytes / plain_s / 1e9
    print(f"[bench] plain per request: {plain_gbs:.3f} GB/s over {k} requests", file=sys.stderr, flush=True)

    srv = gcompress.Server(ctx, budget)
    # verification pass (untimed): every result of the first V requests against the plain form's
    v = min(a.verify_requests, len(sat))
    want = []
    for (o, s, algo, lv) in sat[:v]:
        plain.run(d_pool + o, s, algo, lv)
        want.append(plain.digest(s, algo))
    got = [None] * v
    serve_run(srv, d_pool, np.zeros(v), sat[:v],
              fetch=lambda i, rid: got.__setitem__(i, served_digest(ctx, srv, rid, sat[i][1], sat[i][2])))
    plain.close()
    mism = [i for i in range(v) if got[i] != want[i]]
    verify = {"requests": v, "bytes": int(sum(s for (_, s, _, _) in sat[:v])), "identical_to_plain": not mism}
    assert not mism, f"served results differ from the plain form for requests {mism[:8]}"

    # latency at the stated rates
    rates = {}
    for rate, (arr, reqs) in rate_sets.items():
        print(f"[bench] serving at {rate} req/s", file=sys.stderr, flush=True)
        r = serve_run(srv, d_pool, arr, reqs)
        rates[str(rate)] = lat_summary(r, reqs)
        print(f"[bench]   {json.dumps(rates[str(rate)])}", file=sys.stderr, flush=True)

    # saturation passes: W warm-up, K timed (events on the host clock: the passes span three streams)
    for _ in range(a.warmup):
        serve_run(srv, d_pool, np.zeros(len(sat)), sat)
    st0 = srv.stats()
    ctx.sync()
    t = time.perf_counter()
    runs = [serve_run(srv, d_pool, np.zeros(len(sat)), sat) for _ in range(a.steps)]
    ctx.sync()
    wall_host = time.perf_counter() - t
    st1 = srv.stats()
    # a pass lasts from its requests' scheduled arrival to its last completion (serve_run schedules the
    # first arrival 20 ms after it starts, so that no open-loop arrival lies in the past: the host clock
    # around the passes also counts those idle leads, and is reported beside)
    wall = sum(float(r["done"].max() - r["t0"]) for r in runs)
    value = sat_bytes * a.steps / wall / 1e9
    sat_lat = lat_summary(runs[-1], sat)
    roof = kernel_roofline(st0, st1, wall)
    srv.stop()

    cpu, odig = (None, {}) if a.no_cpu else cpu_baseline_serve(pool, sat, a.cpu_budget_s)
    parity = None
    if odig:
        common = [i for i in odig if i < v]
        parity = {"requests": len(common), "bytes": int(sum(sat[i][1] for i in common)),
                  "identical": all(odig[i] == want[i] for i in common)}
        assert parity["identical"], "served/plain results differ from the oracle"
    ctx.free(d_pool)

    sizes = np.array([s for (_, s, _, _) in sat])
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(wall * 1e3 / a.steps, 3),
        "ms_per_step_host_clock": round(wall_host * 1e3 / a.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic (seeded four-kind mix: random, word text, back-references, structured binary; 1 GiB pool)",
        "config": {
            "workload": "serving: open-loop requests, sizes log-uniform 1 KiB-64 MiB, 50% Zstd L1 / 25% Zstd L2 / "
                        "25% LZMA2 L1, 64 GiB HBM budget; inputs resident in HBM (windows of the pool), results "
                        "left in HBM; value = saturation passes (all requests arrive at once), each from its "
                        "requests' arrival to its last completion",
            "form": "served: continuous batching into the persistent device-queue kernels, 3 tick slots on "
                    "their own streams",
            "requests_per_step": len(sat),
            "input_bytes_per_step": int(sat_bytes),
            "mean_request_bytes": int(sizes.mean()),
            "hbm_budget_gib": a.budget_gib,
            "sub_block": 65536,
        },
        "served_vs_plain": round(value / plain_gbs, 2),
        "plain_per_request_gbs": round(plain_gbs, 3),
        "plain": {"requests": k, "bytes": int(plain_bytes),
                  "form": "one synchronous kernel-ABI call per request, same device-resident inputs"},
        "latency_at_rates": rates,
        "saturation": sat_lat,
        "verify_vs_plain": verify,
        "parity_vs_oracle": parity,
        "server": {"ticks": int(st1["batches"] - st0["batches"]),
                   "sub_blocks": int(st1["items"] - st0["items"]),
                   # every served byte of this run (verification, rates, warm-up and timed passes): what the
                   # served kernels' HBM counters are divided by
                   "totals": {"zstd_bytes": int(st1["zstd_bytes"]), "lzma2_bytes": int(st1["lzma2_bytes"]),
                              "launches": [int(x) for x in st1["launches"]]}},
        "roofline": roof,
        "cpu_baseline": cpu,
        "host": host_cores(),
    }
    out.update(extras)
    return out


# ---------------------------------------------------------------------------------------------------
# Batched configurations
# ---------------------------------------------------------------------------------------------------
def batched_config(a, hip, ctx, name, data, inputs, modes=("static",), plain_pass=True, steps=None):
    """One batched configuration: K timed passes per mode (HIP events on the library stream), per-kernel
    durations (one event pair per kernel), the plain per-chunk/per-input form, the dominant kernel's
    roofline, and the batched results against the plain form on a sample."""
    L = gcompress.lib()
    stream = L.gc_stream(ctx.h)
    steps = steps or a.steps
    n = data.size
    d_in = ctx.alloc(n + 256)
    ctx.htod(d_in, data)
    work = sum(s for (_, s, _, _) in inputs)
    zbytes = sum(s for (_, s, al, _) in inputs if al == ZSTD)
    lbytes = work - zbytes
    out = {"inputs": len(inputs), "input_bytes_per_step": int(work), "zstd_bytes": int(zbytes),
           "lzma2_bytes": int(lbytes)}
    plain = None
    if plain_pass:
        big = max(s for (_, s, _, _) in inputs)
        p = Plain(ctx, big)
        for (o, s, algo, lv) in inputs[:2]:
            p.run(d_in + o, s, algo, lv)
        t = time.perf_counter()
        for (o, s, algo, lv) in inputs:
            p.run(d_in + o, s, algo, lv)
        out["plain_gbs"] = round(work / (time.perf_counter() - t) / 1e9, 3)
        plain = p
    b = gcompress.Batch(ctx, d_in, n + 256, inputs)
    knames = ["candidates", "lzma2_lengths", "zstd_info", "zstd_parse"]
    for mode in modes:
        b.set_mode(gcompress.BATCH_PERSISTENT if mode == "persistent" else gcompress.BATCH_STATIC)
        for _ in range(max(a.warmup, 1)):
            b.run()
        ctx.sync()
        e0, e1 = hip.event(), hip.event()
        hip.record(e0, stream)
        for _ in range(steps):
            b.run()
        hip.record(e1, stream)
        ms = hip.elapsed_ms(e0, e1) / steps
        b.check()
        kt = [b.run_timed() for _ in range(3)]
        k_ms = [float(np.mean([x[i] for x in kt])) for i in range(4)]
        out[mode] = {"gbs": round(work / (ms / 1e3) / 1e9, 3), "ms_per_pass": round(ms, 3),
                     "kernel_ms": {knames[i]: round(k_ms[i], 3) for i in range(4)}}
        if plain:
            out[mode]["vs_plain"] = round(out[mode]["gbs"] / out["plain_gbs"], 2)
    # results (outside the timed region): sequence totals, the roofline, equality with plain on a sample
    n_seq = n_sub_z = 0
    for i, (_, s, algo, _) in enumerate(inputs):
        if algo == ZSTD:
            _, counts = b.zstd(i, with_seqs=False)
            n_seq += int(counts.sum())
            n_sub_z += counts.size
    share = {0: work, 1: 5 * lbytes, 2: zbytes, 3: zbytes + 16 * n_seq + 4 * n_sub_z}
    names = {"static": ["gc_cand_kernel", "gc_lzma2_len_kernel", "gc_zstd_info_kernel", "gc_zstd_parse_kernel"],
             "persistent": ["gc_cand_q_kernel", "gc_lzma2_len_q_kernel", "gc_zstd_info_q_kernel",
                            "gc_zstd_parse_q_kernel"]}
    for mode in modes:
        k_ms = [out[mode]["kernel_ms"][kn] for kn in knames]
        dom = max((i for i in share if share[i]), key=lambda i: k_ms[i])
        achieved = share[dom] / (k_ms[dom] / 1e3) / 1e9
        kname = names[mode][dom]
        per_byte = hbm_traffic_per_input_byte(name, kname)
        out[mode]["roofline"] = {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 2),
                                 "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                                 "traffic": int(per_byte * work) if per_byte is not None else None,
                                 "algorithmic_bytes": int(share[dom])}
    out["results"] = {"n_sequences": n_seq, "zstd_sub_blocks": n_sub_z}
    if plain:
        same = True
        picked = list(range(0, len(inputs), max(1, len(inputs) // 16)))
        cmp_bytes = 0
        for i in picked:
            o, s, algo, lv = inputs[i]
            plain.run(d_in + o, s, algo, lv)
            if algo == LZMA2:
                same &= digest(b.lzma2(i)) == plain.digest(s, algo)
            else:
                seqs, counts = b.zstd(i)
                same &= zstd_digest_from_slots(seqs, counts) == plain.digest(s, algo)
            cmp_bytes += s
        out["identical_to_plain"] = {"inputs": len(picked), "bytes": int(cmp_bytes), "identical": bool(same)}
        assert same, f"{name}: batched results differ from the plain form"
        plain.close()
    b.close()
    ctx.free(d_in)
    return out


def lzma2_end_to_end(a):
    """LZMA2 L1 end to end on configuration 3 (the whole 640 MiB): the reference's shape -- GPU match
    finding in 128 MiB chunks, device-side triage, liblzma preset 9 per 64 KiB sub-block on a pool of
    min(cores, 8), NVLZ -- beside the range coder driven by the GPU's matches (on the host, and on the GPU:
    the same bytes) and level 0 (liblzma on whole chunks, the reference's CPU path); every container decoded
    back and compared with the input."""
    data, chunks = W.config3()
    pool_n = min(threads_all(), 8)

    def by_chunk(fn, name):
        # the container of the whole input, made chunk by chunk (sub-blocks never cross a 128 MiB chunk,
        # so the entries are the same) with a line after each chunk
        sizes, payload = [], []
        for k, (o, s, _, _) in enumerate(chunks):
            blob = fn(data[o:o + s])
            nch = int.from_bytes(blob[20:28], "little")
            sizes += [int.from_bytes(blob[28 + 8 * j: 36 + 8 * j], "little") for j in range(nch)]
            payload.append(blob[28 + 8 * nch:])
            print(f"[bench]   {name}: chunk {k + 1}/{len(chunks)}", file=sys.stderr, flush=True)
        head = b"NVLZ" + data.size.to_bytes(8, "little") + (65536).to_bytes(8, "little") + len(sizes).to_bytes(8, "little")
        return head + b"".join(z.to_bytes(8, "little") for z in sizes) + b"".join(payload)

    out = {"input_bytes": int(data.size), "encoder_threads": pool_n, "host": host_cores(),
           "reference_readme": {"input": "613 MB Windows Server 2003 ISO, RTX 3090 / 3090 Ti, host unstated",
                                "l1_s": 164.5, "l1_size_pct": 91.4, "l0_s": 12.3, "l0_size_pct": 91.2}}
    legs = (("reference_shape_triage_preset9",
             lambda: by_chunk(lambda x: gcompress.nvlz_compress(x, 1, 128, threads=pool_n), "preset 9")),
            ("gpu_matches_range_coder", lambda: gcompress.nvlz_compress_matches(data, 128, threads=pool_n)),
            ("gpu_matches_gpu_range_coder", lambda: gcompress.nvlz_compress_gpu(data, 128)),
            ("level0_liblzma", lambda: gcompress.nvlz_compress_l0(data, 128, 0, threads=pool_n)))
    gpu_blob = None
    for key, fn in legs:
        print(f"[bench] LZMA2 end to end: {key}", file=sys.stderr, flush=True)
        t = time.perf_counter()
        blob = fn()
        dt = time.perf_counter() - t
        same = gcompress.decompress(blob) == data.tobytes()
        assert same, f"{key}: NVLZ does not round-trip"
        out[key] = {"s": round(dt, 2), "gbs": round(data.size / dt / 1e9, 4), "ratio": round(data.size / len(blob), 4),
                    "size_pct": round(100.0 * len(blob) / data.size, 2), "round_trip": same}
        if key == "gpu_matches_gpu_range_coder":
            gpu_blob = blob
    out["gpu_matches_gpu_range_coder_reused_buffer"] = reused_buffer_leg(
        lambda buf: gcompress.nvlz_compress_gpu(data, 128, out=buf),
        gcompress.lib().gc_container_bound(data.size, gcompress.ALGO_LZMA2), data.size, gpu_blob)
    return out


def reused_buffer_leg(call, cap, n, want):
    """A container call through the C ABI into an output buffer the caller keeps across calls, as a host
    program holding its own buffer makes it (the first call untimed): no fresh pages and no copy into a
    Python bytes object, which the bytes-returning legs include.  Three timed calls: the median and the
    fastest.  Checked equal to `want`."""
    buf = np.ones(cap, np.uint8)
    call(buf)
    ts = []
    for _ in range(3):
        t = time.perf_counter()
        m = call(buf)
        ts.append(time.perf_counter() - t)
    same = buf[:m].tobytes() == want
    assert same, "the container written into the caller's buffer differs"
    dt = float(np.median(ts))
    return {"s": round(dt, 3), "gbs": round(n / dt / 1e9, 3), "s_min": round(min(ts), 3), "identical": same}


def gpu_frames(hip, ctx, data, inputs, sample_bytes=256 * MiB):
    """Zstd frames coded on the GPU (gc_zstd_encode_block_dev: per-dimension predefined/RLE/FSE tables,
    repeat-offset codes, Huffman or RLE literals) for the first Zstd L1 input's bytes onwards, up to
    sample_bytes, as one buffer of 64 KiB sub-chunks: the match finder with literals first (untimed), then
    the coding kernel alone timed with HIP events on the library stream (median of 3).  Ratio beside
    libzstd's (ZSTD_compressSequences on the same sequences, first 16 MiB); the frames themselves are
    checked against the oracle coder and libzstd in tests/test_gpu_zenc.py."""
    L = gcompress.lib()
    stream = L.gc_stream(ctx.h)
    o = next(x for x in inputs if x[2] == ZSTD and x[3] == 1)[0]
    n = min(sample_bytes, data.size - o)
    nsub = -(-n // 65536)
    max_seq = 65536 // 3
    slot = 65536 + 16
    d_in = ctx.alloc(n + 256)
    ctx.htod(d_in, np.ascontiguousarray(data[o:o + n]))
    d_s, d_l = ctx.alloc(nsub * max_seq * 16), ctx.alloc(nsub * 65536)
    d_sc, d_lc = ctx.alloc(nsub * 4), ctx.alloc(nsub * 4)
    d_o, d_os = ctx.alloc(nsub * slot), ctx.alloc(nsub * 4)
    rc = L.gc_zstd_match_find_dev(ctx.h, d_in, d_s, 0, d_l, d_sc, d_lc, 65536, nsub, n, 16, max_seq)
    assert rc == 0, L.gc_last_error()
    ctx.sync()
    ms = []
    for _ in range(3):
        e0, e1 = hip.event(), hip.event()
        hip.record(e0, stream)
        rc = L.gc_zstd_encode_block_dev(ctx.h, d_in, d_s, 0, d_l, d_sc, d_lc, d_o, d_os, 65536, nsub, n, max_seq,
                                        slot)
        assert rc == 0, L.gc_last_error()
        hip.record(e1, stream)
        ms.append(hip.elapsed_ms(e0, e1))
    sizes = np.zeros(nsub, np.uint32)
    ctx.dtoh(sizes, d_os)
    k = min(nsub, 256)
    seqs = np.zeros((k, max_seq, 4), np.uint32)
    counts = np.zeros(k, np.uint32)
    ctx.dtoh(seqs, d_s)
    ctx.dtoh(counts, d_sc)
    for p_ in (d_in, d_s, d_l, d_sc, d_lc, d_o, d_os):
        ctx.free(p_)
    x = np.ascontiguousarray(data[o:o + min(n, k * 65536)])
    rows = np.concatenate([seqs[b, :counts[b]] for b in range(k)] + [np.zeros((1, 4), np.uint32)])
    _, zs = gcompress.zstd_encode_sequences(x, rows, counts, threads=min(threads_all(), 16))
    kms = float(np.median(ms))
    # the NVZS container of the same bytes from host memory (match finding and frames on the GPU, the
    # frames copied back): returned as bytes, and written into a buffer the caller keeps
    xs = np.ascontiguousarray(data[o:o + n])
    t = time.perf_counter()
    blob = gcompress.nvzs_compress_gpu(xs, 1)
    dt = time.perf_counter() - t
    assert gcompress.decompress(blob) == xs.tobytes(), "NVZS does not round-trip"
    nvzs = {"bytes_returned": {"s": round(dt, 3), "gbs": round(n / dt / 1e9, 3), "ratio": round(n / len(blob), 4)},
            "reused_buffer": reused_buffer_leg(lambda buf: gcompress.nvzs_compress_gpu(xs, 1, out=buf),
                                               gcompress.lib().gc_nvzs_gpu_bound(n), n, blob)}
    return {"sample_bytes": int(n), "kernel": "gc_zstd_encode_kernel", "kernel_ms": round(kms, 3),
            "encode_input_gbs": round(n / (kms / 1e3) / 1e9, 2), "ratio": round(n / int(sizes.sum()), 4),
            "ratio_first_16mib": round(x.size / int(sizes[:k].sum()), 4),
            "libzstd_ratio_first_16mib": round(x.size / sum(zs), 4), "nvzs_end_to_end": nvzs}


def gpu_lzma2_coder(hip, ctx, data):
    """LZMA2 streams coded on the GPU (gc_lzma2_encode_dev: one stream per 64 KiB sub-block from the packed
    matches, the host coder's bytes) for the whole 640 MiB of configuration 3 as one buffer: the match finder
    first (untimed), then the coding kernel alone timed with HIP events on the library stream (median of 3).
    The streams are checked against the host coder in tests/test_gpu_lzma_enc.py."""
    L = gcompress.lib()
    stream = L.gc_stream(ctx.h)
    n = int(data.size)
    nsub = -(-n // 65536)
    d_in, d_pk = ctx.alloc(n + 256), ctx.alloc(n * 4)
    ctx.htod(d_in, data)
    d_o, d_os = ctx.alloc(nsub * gcompress.LZMA2_GPU_SLOT), ctx.alloc(nsub * 4)
    rc = L.gc_lzma2_match_find_packed_dev(ctx.h, d_in, d_pk, 65536, n, 32)
    assert rc == 0, L.gc_last_error()
    ctx.sync()
    ms = []
    for _ in range(3):
        e0, e1 = hip.event(), hip.event()
        hip.record(e0, stream)
        rc = L.gc_lzma2_encode_dev(ctx.h, d_in, d_pk, d_o, d_os, n)
        assert rc == 0, L.gc_last_error()
        hip.record(e1, stream)
        ms.append(hip.elapsed_ms(e0, e1))
    sizes = np.zeros(nsub, np.uint32)
    ctx.dtoh(sizes, d_os)
    for p_ in (d_in, d_pk, d_o, d_os):
        ctx.free(p_)
    kms = float(np.median(ms))
    raw = int((sizes == np.minimum(65536, n - 65536 * np.arange(nsub)) + 4).sum())
    return {"input_bytes": n, "kernel": "gc_lzma2_encode_kernel", "kernel_ms": round(kms, 3),
            "encode_input_gbs": round(n / (kms / 1e3) / 1e9, 3), "ratio": round(n / int(sizes.sum()), 4),
            "raw_sub_blocks": raw, "sub_blocks": nsub}


def bench_lzma2_coder(a, hip, ctx):
    """The GPU LZMA2 coder alone on configuration 3 (640 MiB, one launch over its 10240 sub-blocks, HIP
    events), then NVLZ end to end from host memory (GPU match finding and coding, streams gathered and copied
    back) beside the same container from the host coder on min(cores, 8) threads: byte-equal, decoded back."""
    data, _ = W.config3()
    r = gpu_lzma2_coder(hip, ctx, data)
    n = int(data.size)
    # per sub-block the coder reads its bytes and their packed matches (4 B per position) and writes the stream
    alg = int(n * 5 + n / r["ratio"])
    ms = r["kernel_ms"]
    e2e = {}
    blob = None
    for key, fn in (("gpu_coder", lambda: gcompress.nvlz_compress_gpu(data, 128)),
                    ("host_coder", lambda: gcompress.nvlz_compress_matches(data, 128, threads=min(threads_all(), 8)))):
        t = time.perf_counter()
        b = fn()
        dt = time.perf_counter() - t
        if blob is None:
            blob = b
            assert gcompress.decompress(b) == data.tobytes(), "NVLZ does not round-trip"
        e2e[key] = {"s": round(dt, 3), "gbs": round(n / dt / 1e9, 3), "ratio": round(n / len(b), 4),
                    "identical_to_gpu_coder": b == blob}
    e2e["gpu_coder_reused_buffer"] = reused_buffer_leg(lambda buf: gcompress.nvlz_compress_gpu(data, 128, out=buf),
                                                       gcompress.lib().gc_container_bound(n, gcompress.ALGO_LZMA2),
                                                       n, blob)
    return {"metric": "LZMA2 coding input GB/s (GPU coder; streams byte-equal to the host coder's)", "value":
            r["encode_input_gbs"], "unit": "GB/s", "n_gpus": 1, "steps": 3, "warmup": 0, "ms_per_step": ms,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8",
            "data": "synthetic (seeded disk-image-like mix)",
            "config": {"workload": "LZMA2 coding of the 640 MiB disk-image-like mix from its GPU matches",
                       "form": "one launch, one wavefront per 64 KiB sub-block"},
            "roofline": {"bound": "hbm", "kernel": "gc_lzma2_encode_kernel", "achieved": round(alg / ms / 1e6, 2),
                         "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(alg / ms / 1e6 / HBM_PEAK_GBS, 5),
                         "traffic": None, "algorithmic_bytes": alg},
            "cpu_baseline": None, "detail": r, "end_to_end_nvlz": e2e, "host": host_cores()}


def bench_containers(a, hip, ctx):
    """The GPU containers end to end from host memory, in a process that has run nothing else: NVZS of the
    first 256 MiB of the configuration-2 mix (Zstd L1, match finding and frames on the GPU) and NVLZ of the
    640 MiB of configuration 3 (match finding and range coding on the GPU).  Each leg returned as bytes (a
    fresh output buffer and a copy into a Python bytes object per call) and written into a buffer the caller
    keeps; decoded back and compared with the input."""
    z2, _ = W.config2()
    xs = np.ascontiguousarray(z2[:256 * MiB])
    z3, _ = W.config3()
    out = {}
    for name, x, call, cap in (
            ("nvzs_256MiB", xs, lambda x, **k: gcompress.nvzs_compress_gpu(x, 1, **k),
             gcompress.lib().gc_nvzs_gpu_bound(xs.size)),
            ("nvlz_640MiB", z3, lambda x, **k: gcompress.nvlz_compress_gpu(x, 128, **k),
             gcompress.lib().gc_container_bound(z3.size, gcompress.ALGO_LZMA2))):
        call(x)  # first call: the library's device buffers are allocated
        t = time.perf_counter()
        blob = call(x)
        dt = time.perf_counter() - t
        assert gcompress.decompress(blob) == x.tobytes(), f"{name}: does not round-trip"
        out[name] = {"input_bytes": int(x.size), "ratio": round(x.size / len(blob), 4),
                     "bytes_returned": {"s": round(dt, 3), "gbs": round(x.size / dt / 1e9, 3)},
                     "reused_buffer": reused_buffer_leg(lambda buf: call(x, out=buf), cap, x.size, blob)}
    top = out["nvlz_640MiB"]["reused_buffer"]
    return {"metric": "container input GB/s from host memory (GPU match finding and coding)",
            "value": top["gbs"], "unit": "GB/s", "n_gpus": 1, "steps": 3, "warmup": 1, "ms_per_step": top["s"] * 1e3,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8",
            "data": "synthetic (seeded mixes)",
            "config": {"workload": "NVLZ of the 640 MiB disk-image-like mix (value) and NVZS of 256 MiB of the "
                                   "four-kind mix, inputs and outputs in host memory"},
            "detail": out, "host": host_cores()}


def extras_all(a, hip, ctx):
    ex = {}
    if not a.no_lzma2_e2e:
        ex["lzma2_l1_end_to_end_config3"] = lzma2_end_to_end(a)
    print("[bench] config 2 (1 GiB Zstd L1+L2)", file=sys.stderr, flush=True)
    data, inputs = W.config2()
    ex["config2_zstd_1GiB_L1L2"] = batched_config(a, hip, ctx, "zstd", data, inputs)
    print("[bench] GPU frame coder (256 MiB of the config 2 mix)", file=sys.stderr, flush=True)
    ex["zstd_gpu_frames_config2"] = gpu_frames(hip, ctx, data, inputs)
    del data
    print("[bench] config 3 (640 MiB LZMA2)", file=sys.stderr, flush=True)
    data, inputs = W.config3()
    ex["config3_lzma2_640MiB"] = batched_config(a, hip, ctx, "lzma2", data, inputs)
    print("[bench] GPU LZMA2 coder (the 640 MiB of config 3)", file=sys.stderr, flush=True)
    ex["lzma2_gpu_coder_config3"] = gpu_lzma2_coder(hip, ctx, data)
    del data
    print("[bench] config 4 (4096 inputs)", file=sys.stderr, flush=True)
    data, inputs = W.config4(threads=min(threads_all(), 16))
    ex["config4_4096_inputs"] = batched_config(a, hip, ctx, "mixed", data, inputs, modes=("persistent", "static"),
                                               steps=max(1, a.steps // 2))
    del data
    return ex


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["serve", "zstd", "lzma2", "mixed", "lzma2_coder", "containers"], default="serve")
    ap.add_argument("--no-extras", action="store_true", help="serve: only the serving workload")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-lzma2-e2e", action="store_true", help="serve: skip the LZMA2 end-to-end leg")
    ap.add_argument("--rates", default="500,2000,8000", help="serve: Poisson arrival rates (req/s), comma-separated")
    ap.add_argument("--sat-requests", type=int, default=6000)
    ap.add_argument("--plain-requests", type=int, default=1000)
    ap.add_argument("--verify-requests", type=int, default=256)
    ap.add_argument("--budget-gib", type=int, default=64)
    ap.add_argument("--cpu-budget-s", type=float, default=12.0)
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="one device-to-device copy of this size first, to scale HBM byte counters")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this match-finding path runs on one GPU; multi-GPU runs are not part of it "
              "(sub-blocks shard by range or by request with no exchange between GPUs)", file=sys.stderr)
        sys.exit(2)
    hip = Hip()
    ctx = gcompress.context(0)
    if a.calib_copy_mib:
        # a plain copy of known size (the runtime's copy kernel): FETCH_SIZE / WRITE_SIZE read for it
        # against calib_copy_mib MiB give the counters' scale
        cb = a.calib_copy_mib * MiB
        c_src, c_dst = ctx.alloc(cb), ctx.alloc(cb)
        assert hip.L.hipMemcpyAsync(c_dst, c_src, cb, 3, gcompress.lib().gc_stream(ctx.h)) == 0
        ctx.sync()
        ctx.free(c_src)
        ctx.free(c_dst)
    if a.workload == "serve":
        extras = {} if a.no_extras else extras_all(a, hip, ctx)
        out = bench_serve(a, hip, ctx, extras)
    elif a.workload == "lzma2_coder":
        out = bench_lzma2_coder(a, hip, ctx)
    elif a.workload == "containers":
        out = bench_containers(a, hip, ctx)
    else:
        if a.workload == "zstd":
            data, inputs = W.config2()
            modes, desc = ("static", "persistent"), "drop-in Zstd L1+L2 match-find, 1 GiB mix in 16 MiB chunks"
        elif a.workload == "lzma2":
            data, inputs = W.config3()
            modes, desc = ("static", "persistent"), "LZMA2 L1 match-find, 640 MiB disk-image-like mix, 128 MiB chunks"
        else:
            data, inputs = W.config4(threads=min(threads_all(), 16))
            modes, desc = ("persistent", "static"), "4096 inputs, 4 KiB-32 MiB log-uniform, a third each Zstd L1/L2/LZMA2"
        r = batched_config(a, hip, ctx, a.workload, data, inputs, modes=modes)
        top = r[modes[0]]
        out = {"metric": METRIC, "value": top["gbs"], "unit": "GB/s", "n_gpus": 1, "steps": a.steps,
               "warmup": a.warmup, "ms_per_step": top["ms_per_pass"], "higher_is_better": True, "scaling": "weak",
               "vs_baseline": None, "dtype": "u8", "data": "synthetic (seeded four-kind mix)",
               "config": {"workload": desc, "form": modes[0]}, "roofline": top["roofline"], "cpu_baseline": None,
               "detail": r, "host": host_cores()}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
