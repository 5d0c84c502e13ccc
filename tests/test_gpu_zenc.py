"""GPU Zstd frame coding (gc_zstd_encode_block_dev, gc_zstd_frames_gpu; the capability of
zstd_compress.cu:756-1038, widened: per-dimension predefined/RLE/FSE_Compressed tables, repeat-offset
codes, Huffman and RLE literals): every frame equals the oracle's frame for the same sequences byte for
byte (oracle/gc_oracle.c orc_zstd_frame, itself checked by libzstd) and decodes with libzstd; the ratio
on the Zstd mix is at least 2.25."""
import numpy as np
import pytest

import datagen
import gcompress
import oracle

pytestmark = pytest.mark.gpu

SUB = 65536


def _cases():
    rng = np.random.default_rng(77)
    out = [(f"mix{n}", datagen.mix(n, n % 97)) for n in (1, 3, 100, 255, 256, 4095, 4096, 65535, 65536,
                                                          3 * 65536 + 777)]
    out += [("text", datagen.text(2 * SUB + 5, 1)), ("backrefs", datagen.backrefs(2 * SUB, 2)),
            ("structured", datagen.structured(2 * SUB, 3)), ("random", datagen.random_bytes(SUB, 4)),
            ("zero", np.zeros(SUB + 3, np.uint8)), ("period5", datagen.periodic(SUB, 5, 5))]
    # sub-blocks of every kind side by side
    out.append(("kinds", np.concatenate([datagen.random_bytes(SUB, 8), np.zeros(SUB, np.uint8),
                                         datagen.text(SUB, 9), rng.integers(0, 4, SUB, dtype=np.uint8)])))
    return out


CASES = _cases()


def oracle_frames(data, level):
    depth = {1: 16, 2: 64}.get(level, 1)
    seq, sc, _, _ = oracle.zstd_match_find(data, depth=depth)
    rows = oracle.zstd_as_zstd_sequences(seq, sc)
    return [oracle.zstd_frame(data[s * SUB:(s + 1) * SUB], rows[s]) for s in range(len(rows))]


@pytest.mark.parametrize("level", [1, 2, 3])
@pytest.mark.parametrize("name,data", CASES, ids=[c[0] for c in CASES])
def test_frames_equal_oracle_and_decode(gctx, name, data, level):
    print(f"[zenc] {name} n={data.size} level={level}", flush=True)
    frames, sizes = gcompress.zstd_frames_gpu(data, level)
    want = oracle_frames(data, level)
    assert len(frames) == len(want)
    for s, (f, w) in enumerate(zip(frames, want)):
        sub = data[s * SUB:(s + 1) * SUB].tobytes()
        assert len(f) == sizes[s]
        assert f == w, f"sub-block {s}: frame differs from the oracle's ({len(f)} vs {len(w)} bytes)"
        assert gcompress.zstd_frame_decode(f, len(sub)) == sub


def _dev(ctx, a):
    p = ctx.alloc(max(a.nbytes, 1) + 64)
    if a.nbytes:
        ctx.htod(p, a)
    return p


@pytest.mark.parametrize("fmt", [0, 1])
@pytest.mark.parametrize("chunk", [65536, 4096, 999])
def test_encode_block_dev_reference_abi(gctx, fmt, chunk):
    """gc_zstd_match_find_dev (with literals) then gc_zstd_encode_block_dev: the reference's pair of kernels,
    both record layouts, sub-chunk sizes other than 64 KiB."""
    print(f"[zenc abi] fmt={fmt} chunk={chunk}", flush=True)
    data = datagen.mix(2 * 65536 + 4321, 12)
    n = data.size
    nsub = -(-n // chunk)
    max_seq = chunk // 3
    rec = 16 if fmt == 0 else 12
    slot = chunk + 16
    L = gcompress.lib()
    d_in = _dev(gctx, data)
    d_s, d_l = gctx.alloc(nsub * max_seq * rec), gctx.alloc(nsub * chunk)
    d_sc, d_lc = gctx.alloc(nsub * 4), gctx.alloc(nsub * 4)
    d_o, d_os = gctx.alloc(nsub * slot), gctx.alloc(nsub * 4)
    rc = L.gc_zstd_match_find_dev(gctx.h, d_in, d_s, fmt, d_l, d_sc, d_lc, chunk, nsub, n, 16, max_seq)
    assert rc == 0, L.gc_last_error()
    rc = L.gc_zstd_encode_block_dev(gctx.h, d_in, d_s, fmt, d_l, d_sc, d_lc, d_o, d_os, chunk, nsub, n, max_seq, slot)
    assert rc == 0, L.gc_last_error()
    frames = np.zeros(nsub * slot, np.uint8)
    sizes = np.zeros(nsub, np.uint32)
    gctx.dtoh(frames, d_o)
    gctx.dtoh(sizes, d_os)
    for p in (d_in, d_s, d_l, d_sc, d_lc, d_o, d_os):
        gctx.free(p)
    seq, sc, _, _ = oracle.zstd_match_find(data, chunk, 16, max_seq)
    rows = oracle.zstd_as_zstd_sequences(seq, sc)
    for s in range(nsub):
        sub = data[s * chunk:(s + 1) * chunk]
        f = frames[s * slot: s * slot + sizes[s]].tobytes()
        assert f == oracle.zstd_frame(sub, rows[s])
        assert gcompress.zstd_frame_decode(f, sub.size) == sub.tobytes()
    # a frame slot smaller than the raw frame is refused
    assert L.gc_zstd_encode_block_dev(gctx.h, 8, 8, fmt, 8, 8, 8, 8, 8, chunk, nsub, n, max_seq, chunk) != 0


@pytest.mark.parametrize("n", [0, 7, SUB, 5 * SUB + 9, (1 << 30) + 3 * SUB + 5])  # the last: two pieces
def test_nvzs_gpu_container(gctx, n):
    data = gcompress.corpus(n, 34, (1, 1, 1, 1), 4096) if n else np.zeros(0, np.uint8)
    print(f"[zenc nvzs] n={n}", flush=True)
    blob = gcompress.nvzs_compress_gpu(data, 1)
    algo, orig, ch, nch = gcompress.container_info(blob)
    assert (algo, orig, ch, nch) == (gcompress.ALGO_ZSTD, n, SUB, -(-n // SUB))
    assert gcompress.decompress(blob) == data.tobytes()


# the oracle coder's own corpus, on which it makes every choice it can (tests/test_zstd_frame_oracle.py)
from test_zstd_frame_oracle import CORPUS as CHOICES  # noqa: E402


@pytest.mark.parametrize("name,data", CHOICES, ids=[c[0] for c in CHOICES])
def test_every_choice_equals_oracle(gctx, name, data):
    print(f"[zenc choices] {name} n={data.size}", flush=True)
    frames, _ = gcompress.zstd_frames_gpu(data, 1)
    want = oracle_frames(data, 1)
    for s, (f, w) in enumerate(zip(frames, want)):
        assert f == w, f"sub-block {s}: frame differs from the oracle's ({len(f)} vs {len(w)} bytes)"


def test_ratio_on_the_zstd_mix(gctx):
    from gcompress import workloads as W

    data, inputs = W.config2()
    o, _, _, _ = inputs[0]
    x = data[o:o + 64 * SUB]
    print(f"[zenc ratio] {x.size} bytes", flush=True)
    frames, sizes = gcompress.zstd_frames_gpu(x, 1)
    assert frames == oracle_frames(x, 1)
    ratio = x.size / sum(sizes)
    print(f"[zenc ratio] {ratio:.4f}", flush=True)
    assert ratio >= 2.25


def test_repeat_offsets_with_equal_offsets_in_a_row(gctx):
    """Sequences a match finder does not make but a caller may hand over: every match of 6 bytes or more
    split in two, the second half with no literals and the same offset (the offset history then holds
    one offset twice, which the coder walks one sequence at a time).  Frames equal the oracle coder's
    and decode."""
    data = datagen.mix(3 * SUB + 1234, 21)
    n = data.size
    nsub = -(-n // SUB)
    max_seq = SUB // 3
    seq, sc, _, _ = oracle.zstd_match_find(data, SUB, 16, max_seq)
    rows = oracle.zstd_as_zstd_sequences(seq, sc)
    seqs = np.zeros((nsub, max_seq, 4), np.uint32)
    lits = np.zeros(nsub * SUB, np.uint8)
    counts = np.zeros(nsub, np.uint32)
    lcounts = np.zeros(nsub, np.uint32)
    split_rows = []
    for s in range(nsub):
        sub = data[s * SUB:(s + 1) * SUB]
        out, pos, lp = [], 0, 0
        for off, ll, ml, _ in rows[s]:
            lits[s * SUB + lp: s * SUB + lp + ll] = sub[pos:pos + ll]
            lp += ll
            if ml >= 6:
                out += [(off, ll, 3, 0), (off, 0, ml - 3, 0)]
            else:
                out.append((off, ll, ml, 0))
            pos += ll + ml
        lits[s * SUB + lp: s * SUB + lp + sub.size - pos] = sub[pos:]
        lcounts[s] = lp + sub.size - pos
        assert len(out) <= max_seq
        counts[s] = len(out)
        if out:
            seqs[s, :len(out)] = np.array(out, np.uint32)
        split_rows.append(np.array(out, np.uint32).reshape(-1, 4))
    assert sum(int((r[1:, 1] == 0).sum() if len(r) > 1 else 0) for r in split_rows) > 100
    print(f"[zenc split] {int(counts.sum())} sequences", flush=True)
    slot = SUB + 16
    L = gcompress.lib()
    d_in, d_s, d_l = _dev(gctx, data), _dev(gctx, seqs), _dev(gctx, lits)
    d_sc, d_lc = _dev(gctx, counts), _dev(gctx, lcounts)
    d_o, d_os = gctx.alloc(nsub * slot), gctx.alloc(nsub * 4)
    rc = L.gc_zstd_encode_block_dev(gctx.h, d_in, d_s, 0, d_l, d_sc, d_lc, d_o, d_os, SUB, nsub, n, max_seq, slot)
    assert rc == 0, L.gc_last_error()
    frames = np.zeros(nsub * slot, np.uint8)
    sizes = np.zeros(nsub, np.uint32)
    gctx.dtoh(frames, d_o)
    gctx.dtoh(sizes, d_os)
    for p in (d_in, d_s, d_l, d_sc, d_lc, d_o, d_os):
        gctx.free(p)
    for s in range(nsub):
        sub = data[s * SUB:(s + 1) * SUB]
        f = frames[s * slot: s * slot + sizes[s]].tobytes()
        assert f == oracle.zstd_frame(sub, split_rows[s]), f"sub-block {s}"
        assert gcompress.zstd_frame_decode(f, sub.size) == sub.tobytes()
