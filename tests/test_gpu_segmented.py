"""The segmented parse (gc_zstd_parse_seg_kernel: one workgroup per sub-chunk, its walk split over 16
wavefronts that meet in fix-up rounds; the per-call Zstd entry points use it for calls of at most
GC_SEG_DEFAULT_MAX_SUB sub-chunks) gives the serial walk's results and the oracle's, bit for bit:
sequences in both record layouts, sequence counts, literal counts and literals, through the kernel ABI
(gc_zstd_match_find_dev), on the data kinds, ragged and tiny sizes, sub-chunk sizes down to 1 byte,
sequence caps, long matches across segments, long chains of fix-ups, and launches that take more
sub-chunks than the grid has workgroups (grid-stride)."""
import numpy as np
import pytest

import datagen
import gcompress
import oracle

pytestmark = pytest.mark.gpu

FORMS = (("segmented", 1 << 30), ("serial", 0))


def _say(msg):
    print(f"[segmented] {msg}", flush=True)


def _find(ctx, data, chunk, depth, max_seq, fmt, lits):
    """gc_zstd_match_find_dev on a device copy: (records [nsub, max_seq, 4|3], literals, counts, lit counts)."""
    n = data.size
    nsub = -(-n // chunk)
    rec = 16 if fmt == 0 else 12
    slots = max(max_seq, 1)
    d_in = ctx.alloc(n + 64)
    d_s = ctx.alloc(nsub * slots * rec)
    d_l = ctx.alloc(nsub * chunk) if lits else None
    d_sc = ctx.alloc(nsub * 4)
    d_lc = ctx.alloc(nsub * 4)
    try:
        ctx.htod(d_in, data)
        ctx.zero(d_s, nsub * slots * rec)
        if lits:
            ctx.zero(d_l, nsub * chunk)
        rc = gcompress.lib().gc_zstd_match_find_dev(ctx.h, d_in, d_s, fmt, d_l, d_sc, d_lc, chunk, nsub, n, depth,
                                                    max_seq)
        assert rc == 0, gcompress.lib().gc_last_error()
        s = np.zeros((nsub, slots, rec // 4), np.uint32)
        lit = np.zeros(nsub * chunk, np.uint8)
        sc = np.zeros(nsub, np.uint32)
        lc = np.zeros(nsub, np.uint32)
        ctx.dtoh(s, d_s)
        if lits:
            ctx.dtoh(lit, d_l)
        ctx.dtoh(sc, d_sc)
        ctx.dtoh(lc, d_lc)
    finally:
        for p in (d_in, d_s, d_l, d_sc, d_lc):
            if p:
                ctx.free(p)
    return s, lit, sc, lc


def _check(ctx, data, chunk=65536, depth=16, max_seq=None, fmt=0, lits=True, forms=FORMS):
    if max_seq is None:
        max_seq = chunk // 3
    oseq, osc, olc, olit = oracle.zstd_match_find(data, chunk, depth, max_seq, literals=True)
    got = {}
    try:
        for name, limit in forms:
            ctx.set_segmented_parse(limit)
            s, lit, sc, lc = _find(ctx, data, chunk, depth, max_seq, fmt, lits)
            np.testing.assert_array_equal(sc, osc, err_msg=f"{name}: sequence counts")
            np.testing.assert_array_equal(lc, olc, err_msg=f"{name}: literal counts")
            for b in range(sc.size):
                k = int(osc[b])
                if fmt == 1:
                    want = oseq[b, :k]
                else:
                    want = np.stack([oseq[b, :k, 2], oseq[b, :k, 0], oseq[b, :k, 1] + 3, np.zeros(k, np.uint32)], 1)
                if not np.array_equal(s[b, :k], want):
                    d = int(np.nonzero((s[b, :k] != want).any(axis=1))[0][0])
                    raise AssertionError(f"{name}: sub-chunk {b} sequence {d}: got {s[b, d]} want {want[d]}")
                if lits:
                    np.testing.assert_array_equal(lit[b * chunk: b * chunk + olc[b]],
                                                  olit[b * chunk: b * chunk + olc[b]],
                                                  err_msg=f"{name}: literals of sub-chunk {b}")
            got[name] = (s, sc)
    finally:
        ctx.set_segmented_parse(gcompress.SEG_DEFAULT_MAX_SUB)
    return got


def _planted(n, seed):
    """Random bytes with one short repeat early on: the offsets it makes current travel through every
    later segment, which has no sequence of its own to meet its earlier walk at (a chain of fix-ups)."""
    d = datagen.random_bytes(n, seed).copy()
    d[3000:3010] = d[1000:1010]
    return d


def _cases():
    out = [("mix", datagen.mix(4 * 65536 + 99, 4)),
           ("text", datagen.text(2 * 65536 + 777, 1)),
           ("backrefs", datagen.backrefs(2 * 65536 + 3, 2)),
           ("structured", datagen.structured(2 * 65536 + 5, 3)),
           ("random", datagen.random_bytes(65536 * 2 + 1, 5)),
           ("zeros", np.zeros(65536 + 7, np.uint8)),
           ("chain", _planted(65536 * 2, 6)),
           # one match spans most segments: they have nothing to walk and pass its end state on
           ("longmatch", np.concatenate([datagen.random_bytes(5000, 7)] * 27)[:2 * 65536]),
           ("corpus", gcompress.corpus(3 * 65536, 8, (1, 1, 1, 1), 4096, False))]
    for k in (1, 3, 37, 200, 4097):
        out.append((f"period{k}", datagen.periodic(65536 + 17, k, k)))
    for n in (1, 3, 4, 5, 63, 64, 65, 127, 1000, 4095, 4096, 4097, 65535):
        out.append((f"text{n}", datagen.text(n, n)))
    return out


CASES = _cases()


@pytest.mark.parametrize("name,data", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("depth", [16, 1])
def test_segmented_equals_serial_and_oracle(gctx, name, data, depth):
    _say(f"{name} n={data.size} depth={depth}")
    got = _check(gctx, data, depth=depth)
    assert np.array_equal(got["segmented"][1], got["serial"][1])


@pytest.mark.parametrize("fmt", [0, 1])
@pytest.mark.parametrize("lits", [True, False])
def test_layouts(gctx, fmt, lits):
    _say(f"layout fmt={fmt} lits={lits}")
    _check(gctx, datagen.mix(3 * 65536 + 4321, 12), fmt=fmt, lits=lits)


@pytest.mark.parametrize("max_seq", [0, 1, 63, 64, 65, 500, 3000])
def test_sequence_cap(gctx, max_seq):
    _say(f"max_seq={max_seq}")
    _check(gctx, datagen.text(2 * 65536, 21), max_seq=max_seq)


@pytest.mark.parametrize("chunk", [1, 7, 64, 999, 4096, 40000])
def test_sub_chunk_sizes(gctx, chunk):
    _say(f"chunk={chunk}")
    n = 3 * 4096 + 11 if chunk < 64 else 3 * 65536 + 11
    _check(gctx, datagen.mix(n, 13, seg=4096), chunk=chunk)


def test_grid_stride(gctx):
    # more sub-chunks than the segmented grid's 2 x CU workgroups: each workgroup takes several
    data = gcompress.corpus(600 * 65536, 14, (1, 1, 1, 1), 4096, False)
    _say(f"grid stride: {data.size >> 16} sub-chunks")
    _check(gctx, data, lits=False)
