"""The allocation-free Zstd entry point (gc_zstd_match_find_dev_noalloc) gives exactly what
gc_zstd_match_find_dev gives, through the C ABI, on calls in a row whose inputs shrink (the second
call finds the first call's contents in the shared scratch), and refuses inputs beyond its scratch."""
import ctypes as C

import numpy as np
import pytest

import gcompress

pytestmark = pytest.mark.gpu

SUB = 65536
MAX_SEQ = SUB // 3


def _find(L, ctx, fn, d_in, n, sub, fmt, lits, depth):
    nsub = -(-n // sub)
    max_seq = max(1, sub // 3)  # compress_zstd_custom.rs:55
    rec = 16 if fmt == 0 else 12
    d_s, d_c = ctx.alloc(nsub * max_seq * rec), ctx.alloc(nsub * 4)
    ctx.zero(d_s, nsub * max_seq * rec)
    ctx.zero(d_c, nsub * 4)
    d_l = ctx.alloc(nsub * sub) if lits else None
    d_lc = ctx.alloc(nsub * 4) if lits else None
    try:
        rc = fn(ctx.h, d_in, d_s, fmt, d_l, d_c, d_lc, sub, nsub, n, depth, max_seq)
        assert rc == 0, L.gc_last_error()
        ctx.sync()
        counts = np.empty(nsub, np.uint32)
        ctx.dtoh(counts, d_c)
        seqs = np.empty((nsub, max_seq, rec // 4), np.uint32)
        ctx.dtoh(seqs, d_s)
        out = {"counts": counts, "seqs": [seqs[k, : counts[k]].copy() for k in range(nsub)]}
        if lits:
            lc = np.empty(nsub, np.uint32)
            ctx.dtoh(lc, d_lc)
            lb = np.empty((nsub, sub), np.uint8)
            ctx.dtoh(lb, d_l)
            out["lit_counts"] = lc
            out["lits"] = [lb[k, : lc[k]].copy() for k in range(nsub)]
        return out
    finally:
        for p in (d_s, d_c, d_l, d_lc):
            if p:
                ctx.free(p)


def _same(a, b):
    assert np.array_equal(a["counts"], b["counts"])
    assert all(np.array_equal(x, y) for x, y in zip(a["seqs"], b["seqs"]))
    if "lits" in a:
        assert np.array_equal(a["lit_counts"], b["lit_counts"])
        assert all(np.array_equal(x, y) for x, y in zip(a["lits"], b["lits"]))


CASES = [
    # (name, bytes, seed, sub-chunk, record layout, literals, depth): the first fills the scratch to its
    # limit (8 MiB in 64 KiB sub-chunks), the next ones are smaller and other shapes
    ("8MiB_zstd_L1", 8 << 20, 11, SUB, 0, False, 16),
    ("1MiB+777_ref_lits_L2", (1 << 20) + 777, 12, SUB, 1, True, 64),
    ("28KiB_in_7B_subchunks", 4096 * 7, 13, 7, 0, True, 16),
]


def test_noalloc_equals_allocating_entry_point(gctx):
    L = gcompress.lib()
    for name, n, seed, sub, fmt, lits, depth in CASES:
        print(f"[noalloc] {name}", flush=True)
        data = gcompress.corpus(n, seed, (1, 1, 1, 1), 4096, False)
        d_in = gctx.alloc(n + 64)
        try:
            gctx.htod(d_in, data)
            want = _find(L, gctx, L.gc_zstd_match_find_dev, d_in, n, sub, fmt, lits, depth)
            got = _find(L, gctx, L.gc_zstd_match_find_dev_noalloc, d_in, n, sub, fmt, lits, depth)
            _same(got, want)
            assert int(want["counts"].sum()) > 0
        finally:
            gctx.free(d_in)


def test_noalloc_refuses_inputs_beyond_its_scratch(gctx):
    L = gcompress.lib()
    n = (8 << 20) + 1
    d_in = gctx.alloc(n + 64)
    try:
        nsub = -(-n // SUB)
        rc = L.gc_zstd_match_find_dev_noalloc(gctx.h, d_in, d_in, 0, None, d_in, None, SUB, nsub, n, 16, MAX_SEQ)
        assert rc == -1 and b"beyond its fixed scratch" in L.gc_last_error()
        rc = L.gc_zstd_match_find_dev_noalloc(gctx.h, d_in, d_in, 0, None, d_in, None, 8, 4097, 4097 * 8, 16, 3)
        assert rc == -1 and b"beyond its fixed scratch" in L.gc_last_error()
    finally:
        gctx.free(d_in)
