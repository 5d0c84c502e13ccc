"""Served form: each request's result is the same alone and in a concurrent mix, equal to the
oracle; admission under a tight HBM budget; oversized requests are rejected."""
import ctypes as C

import numpy as np
import pytest

import gcompress
import oracle

pytestmark = pytest.mark.gpu
SUB = 65536


def _requests(seed, count, max_size=600000):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(np.exp(rng.uniform(np.log(1024), np.log(max_size))))
        data = gcompress.corpus(n, 1000 + seed * 100 + i, rng.dirichlet([1, 1, 1, 1]))
        algo, level = [(gcompress.ALGO_ZSTD, 1), (gcompress.ALGO_ZSTD, 1), (gcompress.ALGO_ZSTD, 2),
                       (gcompress.ALGO_LZMA2, 1)][i % 4]
        out.append((data, algo, level))
    return out


def _check(srv, rid, data, algo, level):
    info = srv.wait(rid)
    assert info.size == data.size
    if algo == gcompress.ALGO_LZMA2:
        got = srv.result_lzma2(rid, data.size)
        np.testing.assert_array_equal(got, oracle.lzma2_packed(data))
    else:
        seqs, counts = srv.result_zstd(rid, info.n_sub, info.n_sequences)
        oseq, osc, _, _ = oracle.zstd_match_find(data, depth={1: 16, 2: 64}[level])
        np.testing.assert_array_equal(counts, osc)
        want = oracle.zstd_as_zstd_sequences(oseq, osc)
        np.testing.assert_array_equal(seqs, np.concatenate(want) if want else np.zeros((0, 4), np.uint32))
    assert info.t_submit <= info.t_admit <= info.t_start <= info.t_done
    srv.release(rid)


def test_alone_and_in_mix(gctx):
    reqs = _requests(1, 24)
    srv = gcompress.Server(gctx, 2 << 30)
    try:
        print("[serve] alone", flush=True)
        for data, algo, level in reqs[:6]:
            rid = srv.submit(data, algo, level)
            _check(srv, rid, data, algo, level)
        print("[serve] mix", flush=True)
        ids = [srv.submit(data, algo, level) for data, algo, level in reqs]
        for rid, (data, algo, level) in zip(ids, reqs):
            _check(srv, rid, data, algo, level)
        st = srv.stats()
        assert st["completed"] == 30 and st["hbm_in_use"] == 0
    finally:
        srv.stop()


def test_tight_budget_and_rejection(gctx):
    reqs = _requests(2, 12, 300000)
    srv = gcompress.Server(gctx, 24 << 20)  # a few requests at a time
    try:
        print("[serve] tight budget", flush=True)
        ids = [srv.submit(data, algo, level) for data, algo, level in reqs]
        for rid, (data, algo, level) in zip(ids, reqs):
            _check(srv, rid, data, algo, level)
        assert srv.stats()["batches"] >= 2
        big = np.zeros(32 << 20, np.uint8)
        with pytest.raises(gcompress.GcError, match="budget"):
            srv.submit(big, gcompress.ALGO_ZSTD, 1)
        rid = srv.submit(np.zeros(0, np.uint8), gcompress.ALGO_LZMA2, 1)
        assert srv.wait(rid).size == 0
        srv.release(rid)
    finally:
        srv.stop()


def test_submit_into_caller_buffers(gctx):
    """Results written straight into the caller's buffers equal the oracle; a Zstd buffer too small
    for the sequences fails that request alone with GC_ESIZE."""
    reqs = _requests(3, 16, 400000)
    srv = gcompress.Server(gctx, 1 << 30)
    try:
        print("[serve] submit_into", flush=True)
        outs = []
        for data, algo, level in reqs:
            nsub = -(-data.size // SUB)
            if algo == gcompress.ALGO_LZMA2:
                out, counts = np.full(data.size, 0xDEADBEEF, np.uint32), None
            else:
                out, counts = np.zeros((data.size // 3 + 64, 4), np.uint32), np.zeros(nsub, np.uint32)
            outs.append((srv.submit_into(data, algo, level, out, counts), out, counts))
        for (rid, out, counts), (data, algo, level) in zip(outs, reqs):
            info = srv.wait(rid)
            if algo == gcompress.ALGO_LZMA2:
                np.testing.assert_array_equal(out, oracle.lzma2_packed(data))
            else:
                oseq, osc, _, _ = oracle.zstd_match_find(data, depth={1: 16, 2: 64}[level])
                np.testing.assert_array_equal(counts, osc)
                want = oracle.zstd_as_zstd_sequences(oseq, osc)
                want = np.concatenate(want) if want else np.zeros((0, 4), np.uint32)
                assert info.n_sequences == want.shape[0]
                np.testing.assert_array_equal(out[:want.shape[0]], want)
            with pytest.raises(gcompress.GcError, match="buffers given"):
                if algo == gcompress.ALGO_LZMA2:
                    srv.result_lzma2(rid, data.size)
                else:
                    srv.result_zstd(rid, info.n_sub, info.n_sequences)
            srv.release(rid)
        print("[serve] submit_into too small", flush=True)
        text = gcompress.corpus(200000, 77, [0, 1, 0, 0])
        small = np.zeros((4, 4), np.uint32)
        cnt = np.zeros(4, np.uint32)
        rid = srv.submit_into(text, gcompress.ALGO_ZSTD, 1, small, cnt)
        ok = srv.submit_into(text, gcompress.ALGO_LZMA2, 1, np.zeros(text.size, np.uint32))
        with pytest.raises(gcompress.GcError, match="exceed the result buffer"):
            srv.wait(rid)
        srv.wait(ok)
        srv.release(rid)
        srv.release(ok)
        with pytest.raises(gcompress.GcError, match="4 bytes per input byte"):
            srv.submit_into(text, gcompress.ALGO_LZMA2, 1, np.zeros(10, np.uint32))
    finally:
        srv.stop()


def _pinned(nbytes):
    """Pinned host memory (hipHostMalloc) as a uint8 array, and its pointer for hipHostFree."""
    hip = C.CDLL("libamdhip64.so")
    hip.hipHostMalloc.argtypes = [C.c_void_p, C.c_size_t, C.c_uint]
    p = C.c_void_p()
    assert hip.hipHostMalloc(C.byref(p), nbytes, 0) == 0
    return np.ctypeslib.as_array((C.c_uint8 * nbytes).from_address(p.value)), hip, p


def test_submit_into_pinned_buffers(gctx):
    """Pinned result buffers are written by the kernels (no DMA copy); same results as the oracle,
    for buffers at offsets inside one pinned allocation."""
    reqs = _requests(4, 12, 300000)
    need = sum(max(d.size * 4, (d.size // 3 + 64) * 16) + 256 for d, _, _ in reqs)
    ring, hip, ptr = _pinned(need)
    srv = gcompress.Server(gctx, 1 << 30)
    try:
        print("[serve] submit_into pinned", flush=True)
        outs, o = [], 16  # not 256-aligned: the copy kernel's unaligned path too
        for data, algo, level in reqs:
            nsub = -(-data.size // SUB)
            if algo == gcompress.ALGO_LZMA2:
                nb, counts = data.size * 4, None
            else:
                nb, counts = (data.size // 3 + 64) * 16, np.zeros(nsub, np.uint32)
            buf = ring[o: o + nb].view(np.uint32)
            buf[:] = 0xDEADBEEF
            outs.append((srv.submit_into(data, algo, level, buf, counts), buf, counts))
            o += nb + 256
        for (rid, buf, counts), (data, algo, level) in zip(outs, reqs):
            info = srv.wait(rid)
            if algo == gcompress.ALGO_LZMA2:
                np.testing.assert_array_equal(buf, oracle.lzma2_packed(data))
            else:
                oseq, osc, _, _ = oracle.zstd_match_find(data, depth={1: 16, 2: 64}[level])
                np.testing.assert_array_equal(counts, osc)
                want = oracle.zstd_as_zstd_sequences(oseq, osc)
                want = np.concatenate(want) if want else np.zeros((0, 4), np.uint32)
                assert info.n_sequences == want.shape[0]
                np.testing.assert_array_equal(buf.reshape(-1, 4)[:want.shape[0]], want)
            srv.release(rid)
    finally:
        srv.stop()
        hip.hipHostFree.argtypes = [C.c_void_p]
        hip.hipHostFree(ptr)


def test_concurrent_clients(gctx):
    """Clients on several threads submit, wait and release while the server thread batches (the server
    must not look requests up in the shared table while clients change it)."""
    import threading

    srv = gcompress.Server(gctx, 256 << 20)
    errors = []
    rng = np.random.default_rng(8)
    pool = gcompress.corpus(1 << 20, 88, [1, 1, 1, 1])
    plans = []
    for t in range(6):
        plan = []
        for i in range(150):
            n = int(rng.integers(1, 40000))
            o = int(rng.integers(0, pool.size - n)) & ~3
            algo, level = [(gcompress.ALGO_ZSTD, 1), (gcompress.ALGO_ZSTD, 2), (gcompress.ALGO_LZMA2, 1)][(t + i) % 3]
            plan.append((pool[o:o + n], algo, level))
        plans.append(plan)

    def client(plan, check_every):
        try:
            for i, (data, algo, level) in enumerate(plan):
                rid = srv.submit(data, algo, level)
                if i % check_every == 0:
                    _check(srv, rid, data, algo, level)  # waits, compares with the oracle, releases
                else:
                    srv.wait(rid)
                    srv.release(rid)
        except Exception as e:  # reported by the main thread
            errors.append(e)

    try:
        print("[serve] concurrent clients", flush=True)
        ths = [threading.Thread(target=client, args=(p, 10)) for p in plans]
        for t in ths:
            t.start()
        for t in ths:
            t.join()
        assert not errors, errors[0]
        assert srv.stats()["completed"] == 6 * 150
    finally:
        srv.stop()


def _dev_result(gctx, srv, rid, n, algo):
    """A device-result request's results copied to the host: (rows as ZSTD_Sequence, counts) or packed."""
    info = srv.wait(rid)
    nsub = -(-n // SUB)
    p, counts = srv.result_dev(rid, nsub)
    if algo == gcompress.ALGO_LZMA2:
        out = np.empty(n, np.uint32)
        gctx.dtoh(out, p)
        return out, info
    slots = np.empty((nsub, gcompress.SERVE_ZSTD_SLOT, 4), np.uint32)
    gctx.dtoh(slots, p)
    rows = np.concatenate([slots[k, : counts[k]] for k in range(nsub)] + [np.zeros((0, 4), np.uint32)])
    assert info.n_sequences == rows.shape[0]
    return (rows, counts), info


def _check_dev(gctx, srv, rid, data, algo, level):
    got, info = _dev_result(gctx, srv, rid, data.size, algo)
    if algo == gcompress.ALGO_LZMA2:
        np.testing.assert_array_equal(got, oracle.lzma2_packed(data))
    else:
        rows, counts = got
        oseq, osc, _, _ = oracle.zstd_match_find(data, depth={1: 16, 2: 64}[level])
        np.testing.assert_array_equal(counts, osc)
        want = oracle.zstd_as_zstd_sequences(oseq, osc)
        np.testing.assert_array_equal(rows, np.concatenate(want) if want else np.zeros((0, 4), np.uint32))
    assert info.t_submit <= info.t_admit <= info.t_start <= info.t_done
    srv.release(rid)


def test_device_resident_requests(gctx):
    """Requests whose input is already in HBM (windows of one device buffer, read in place), results left
    on the device: alone and in a mix, equal to the oracle; their memory is held until released."""
    pool = gcompress.corpus(8 << 20, 91, [1, 1, 1, 1], 65536, False)
    d = gctx.alloc(pool.size + 64)
    gctx.htod(d, pool)
    rng = np.random.default_rng(9)
    reqs = []
    for i in range(24):
        n = int(np.exp(rng.uniform(np.log(1024), np.log(3 << 20))))
        o = int(rng.integers(0, pool.size - n)) & ~3
        reqs.append((o, n) + [(gcompress.ALGO_ZSTD, 1), (gcompress.ALGO_ZSTD, 2), (gcompress.ALGO_LZMA2, 1)][i % 3])
    reqs.append((pool.size - 5000, 5000, gcompress.ALGO_ZSTD, 1))  # ends where the buffer ends
    reqs.append((pool.size - 70000, 70000, gcompress.ALGO_LZMA2, 1))
    srv = gcompress.Server(gctx, 2 << 30)
    try:
        print("[serve] device-resident alone", flush=True)
        for o, n, algo, level in reqs[:4] + reqs[-2:]:
            rid = srv.submit_dev(d + o, n, algo, level)
            _check_dev(gctx, srv, rid, pool[o:o + n], algo, level)
        print("[serve] device-resident mix", flush=True)
        ids = [srv.submit_dev(d + o, n, algo, level) for o, n, algo, level in reqs]
        for rid, (o, n, algo, level) in zip(ids, reqs):
            _check_dev(gctx, srv, rid, pool[o:o + n], algo, level)
        assert srv.stats()["hbm_in_use"] == 0
        with pytest.raises(gcompress.GcError, match="aligned"):
            srv.submit_dev(d + 2, 100, gcompress.ALGO_ZSTD, 1)
    finally:
        srv.stop()
        gctx.free(d)


def test_bulk_submit_and_reap(gctx):
    """gc_server_submit_dev_many / gc_server_wait_release_many: the same results as one call per request
    (equal to the oracle), every request released, and a batch with a bad request rejected whole."""
    pool = gcompress.corpus(8 << 20, 92, [1, 1, 1, 1], 65536, False)
    d = gctx.alloc(pool.size + 64)
    gctx.htod(d, pool)
    rng = np.random.default_rng(10)
    reqs = [(0, 0, gcompress.ALGO_ZSTD, 1)]  # an empty request is done at once
    for i in range(30):
        n = int(np.exp(rng.uniform(np.log(1024), np.log(3 << 20))))
        o = int(rng.integers(0, pool.size - n)) & ~3
        reqs.append((o, n) + [(gcompress.ALGO_ZSTD, 1), (gcompress.ALGO_ZSTD, 2), (gcompress.ALGO_LZMA2, 1)][i % 3])
    srv = gcompress.Server(gctx, 1 << 30)  # tight: admission waits for releases
    try:
        print("[serve] bulk submit", flush=True)
        ids = srv.submit_dev_many(srv.dev_batch([(d + o, n, a, lv) for o, n, a, lv in reqs]))
        assert len(ids) == len(reqs) and len(set(int(x) for x in ids)) == len(reqs)
        for rid, (o, n, algo, level) in list(zip(ids, reqs))[:12]:
            _check_dev(gctx, srv, int(rid), pool[o:o + n], algo, level)
        infos, status = srv.wait_release_many(ids[12:])
        assert list(status) == [0] * (len(reqs) - 12)
        for info, (o, n, algo, level) in zip(infos, reqs[12:]):
            assert info.size == n and info.t_submit <= info.t_admit <= info.t_start <= info.t_done
        st = srv.stats()
        assert st["hbm_in_use"] == 0 and st["pending"] == 0
        bad = [(d, 4096, gcompress.ALGO_ZSTD, 1), (d + 2, 100, gcompress.ALGO_ZSTD, 1)]
        with pytest.raises(gcompress.GcError, match="aligned"):
            srv.submit_dev_many(srv.dev_batch(bad))
        assert srv.stats()["pending"] == 0
    finally:
        srv.stop()
        gctx.free(d)


def test_dense_sequences_device_results(gctx):
    """Device-result Zstd requests whose sub-blocks hold about 10 000 sequences each (3-byte tokens in random
    order): their records reach far into the info words they are laid over while the parse still reads
    them, and every sequence equals the oracle's."""
    rng = np.random.default_rng(17)
    toks = rng.integers(0, 256, (16, 3), dtype=np.uint8)
    data = np.ascontiguousarray(toks[rng.integers(0, 16, 3 * SUB // 3 + 900)].reshape(-1)[: 3 * SUB + 2500])
    d = gctx.alloc(data.size + 64)
    gctx.htod(d, data)
    srv = gcompress.Server(gctx, 1 << 30)
    try:
        print("[serve] dense sequences", flush=True)
        for level in (1, 2):
            rid = srv.submit_dev(d, data.size, gcompress.ALGO_ZSTD, level)
            _check_dev(gctx, srv, rid, data, gcompress.ALGO_ZSTD, level)
        _, counts = oracle.zstd_match_find(data, depth=16)[:2]
        assert counts[:3].min() > 9000  # a whole sub-block's records cover more than 140 KB of its 384 KB
    finally:
        srv.stop()
        gctx.free(d)


def test_ticks_overlap(gctx):
    """Continuous batching: a request that arrives while a long tick runs starts in the next tick slot
    beside it (after that tick started, before it ended), and its result is the same as alone."""
    import time

    pool = gcompress.corpus(32 << 20, 92, [1, 1, 1, 1], 65536, False)
    d = gctx.alloc(pool.size + 64)
    gctx.htod(d, pool)
    srv = gcompress.Server(gctx, 64 << 30)
    try:
        print("[serve] overlapping ticks", flush=True)
        big = [srv.submit_dev(d, pool.size, gcompress.ALGO_ZSTD, 1) for _ in range(96)]
        time.sleep(0.005)
        small_data = pool[:50000]
        small = srv.submit_dev(d, small_data.size, gcompress.ALGO_LZMA2, 1)
        s_info = srv.wait(small)
        infos = [srv.wait(r) for r in big]
        first = min(i.t_start for i in infos)
        last_done = max(i.t_done for i in infos)
        print(f"[serve] big ticks {first:.4f}..{last_done:.4f}, small start {s_info.t_start:.4f} "
              f"done {s_info.t_done:.4f}", flush=True)
        assert first < s_info.t_start < last_done
        _check_dev(gctx, srv, small, small_data, gcompress.ALGO_LZMA2, 1)
        _check_dev(gctx, srv, big[0], pool, gcompress.ALGO_ZSTD, 1)
        for r in big[1:]:
            srv.release(r)
        assert srv.stats()["hbm_in_use"] == 0
    finally:
        srv.stop()
        gctx.free(d)
