"""GPU LZMA2 coding (gc_lzma2_encode_dev, gc_lzma2_streams_gpu, gc_nvlz_compress_gpu; the coder of
compress_lzma2_custom.rs:344-465 / :601-664 moved to the device): every stream equals the host coder's
(gc_lzma2_encode_matches / gc_cpu_encode_from_matches, whose streams liblzma and Python's lzma decode) for
the same matches, byte for byte, and decodes back to its sub-block."""
import numpy as np
import pytest

import datagen
import gcompress
import oracle

pytestmark = pytest.mark.gpu

SUB = 65536
MiB = 1 << 20


def _cases():
    rng = np.random.default_rng(91)
    out = [(f"mix{n}", datagen.mix(n, n % 89)) for n in (1, 2, 3, 5, 100, 4096, 65535, 65536, 65537,
                                                          3 * 65536 + 777)]
    out += [("text", datagen.text(2 * SUB + 5, 1)), ("backrefs", datagen.backrefs(2 * SUB, 2)),
            ("structured", datagen.structured(2 * SUB, 3)), ("random", datagen.random_bytes(SUB, 4)),
            ("zero", np.zeros(SUB + 3, np.uint8)), ("period5", datagen.periodic(SUB, 5, 5)),
            ("period300", datagen.periodic(SUB, 300, 6))]
    # runs of one byte broken at random places: repeat lengths that cross windows and the 273 cap
    runs = np.repeat(rng.integers(0, 3, 2000, dtype=np.uint8), rng.integers(1, 400, 2000))[:2 * SUB]
    out.append(("runs", np.ascontiguousarray(runs)))
    out.append(("kinds", np.concatenate([datagen.random_bytes(SUB, 8), np.zeros(SUB, np.uint8),
                                         datagen.text(SUB, 9), rng.integers(0, 4, SUB, dtype=np.uint8)])))
    # the disk-image-like corpus of the 640 MiB workload (its kinds, 1 MiB segments, a zero gap)
    out.append(("disk", gcompress.corpus(6 * MiB, 3, (15, 25, 20, 40), MiB, True, 4 * MiB, MiB)))
    return out


CASES = _cases()


@pytest.mark.parametrize("name,data", CASES, ids=[c[0] for c in CASES])
def test_streams_equal_host_coder_and_decode(gctx, name, data):
    print(f"[lzenc] {name} n={data.size}", flush=True)
    streams, sizes = gcompress.lzma2_streams_gpu(data)
    packed = gctx.lzma2_packed(data)
    want, wsizes = gcompress.cpu_encode_from_matches(data, packed, 4)
    assert sizes == wsizes and len(streams) == len(want)
    for s, (g, w) in enumerate(zip(streams, want)):
        assert g == w, f"sub-block {s}: stream differs from the host coder's ({len(g)} vs {len(w)} bytes)"
        sub = data[s * SUB:(s + 1) * SUB].tobytes()
        assert gcompress.lzma2_raw_decode(g, len(sub)) == sub


def _dev(ctx, a):
    p = ctx.alloc(max(a.nbytes, 1) + 64)
    if a.nbytes:
        ctx.htod(p, a)
    return p


def _encode_dev(ctx, data, packed):
    n = data.size
    nsub = -(-n // SUB)
    L = gcompress.lib()
    d_in, d_pk = _dev(ctx, data), _dev(ctx, packed)
    d_out, d_sz = ctx.alloc(nsub * gcompress.LZMA2_GPU_SLOT), ctx.alloc(nsub * 4)
    try:
        assert L.gc_lzma2_encode_dev(ctx.h, d_in, d_pk, d_out, d_sz, n) == 0, L.gc_last_error()
        ctx.sync()
        slots = np.empty(nsub * gcompress.LZMA2_GPU_SLOT, np.uint8)
        sizes = np.empty(nsub, np.uint32)
        ctx.dtoh(slots, d_out)
        ctx.dtoh(sizes, d_sz)
    finally:
        for p in (d_in, d_pk, d_out, d_sz):
            ctx.free(p)
    o = gcompress.LZMA2_GPU_HEAD
    return [slots[s * gcompress.LZMA2_GPU_SLOT + o:s * gcompress.LZMA2_GPU_SLOT + o + int(sizes[s])].tobytes()
            for s in range(nsub)]


@pytest.mark.parametrize("source", ["oracle", "synthetic"])
@pytest.mark.parametrize("name", ["mix3", "text", "zero", "runs", "kinds"])
def test_encode_dev_any_matches(gctx, name, source):
    """The device ABI on matches it did not find: the oracle's, and valid-format words that claim matches
    the bytes do not have (the coder codes what it is given; the host coder with verify = 0 does the same)."""
    data = dict(CASES)[name]
    print(f"[lzenc dev] {name} {source} n={data.size}", flush=True)
    if source == "oracle":
        packed = oracle.lzma2_packed(data).astype(np.uint32)
    else:
        rng = np.random.default_rng(data.size)
        dist = rng.integers(0, 1 << 16, data.size, dtype=np.uint32)
        ln = rng.integers(0, 1 << 16, data.size, dtype=np.uint32)  # lengths past 273 are cut to 273
        ln[rng.random(data.size) < 0.6] = 0
        packed = (dist << 16) | ln
    got = _encode_dev(gctx, data, packed)
    for s, g in enumerate(got):
        sub = data[s * SUB:(s + 1) * SUB]
        w = gcompress.lzma2_encode_matches(sub, packed[s * SUB:s * SUB + sub.size], verify=False)
        assert g == w, f"sub-block {s}: stream differs from the host coder's ({len(g)} vs {len(w)} bytes)"
        if source == "oracle":
            assert gcompress.lzma2_raw_decode(g, sub.size) == sub.tobytes()


@pytest.mark.parametrize("dict_mb", [1, 128])
def test_nvlz_gpu_container(gctx, dict_mb):
    data = dict(CASES)["disk"]
    print(f"[lzenc nvlz] dict_mb={dict_mb} n={data.size}", flush=True)
    blob = gcompress.nvlz_compress_gpu(data, dict_mb)
    assert blob == gcompress.nvlz_compress_matches(data, dict_mb, threads=8)
    assert gcompress.decompress(blob) == data.tobytes()
