"""Configurations 2-4 at their full stated sizes on the GPU, against the oracle: every sub-block's
sequences (Zstd) or per-position match (LZMA2) of the plain per-chunk drop-in, the static batched form
and the persistent batched form (device work queue, hipGraph replay) equals the oracle's, compared
through SHA-256 digests of one canonical layout; the containers round-trip byte-exactly.

  2  1 GiB mix in 16 MiB chunks at Zstd L1 and L2 (compress_zstd_custom.rs:91-116), NVZS round trip
  3  640 MiB disk-image-like mix in 128 MiB chunks, LZMA2 (compress_lzma2_custom.rs:518-549): packed
     matches, the GpuMatchData layout on one whole chunk, triage, NVLZ round trips
  4  4096 inputs, 14.71 GiB, a third each Zstd L1 / L2 / LZMA2: batched (both forms) == oracle ==
     plain per input
"""
import json
import os

import numpy as np
import pytest

import fullsize as F
import gcompress
import oracle
from gcompress import workloads as W

pytestmark = pytest.mark.gpu

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "config_sha256.json")))
MODES = (("static", gcompress.BATCH_STATIC), ("persistent", gcompress.BATCH_PERSISTENT))


def _say(msg):
    print(f"[configs] {msg}", flush=True)


def _batched(gctx, data, inputs):
    """Digests of every input's results from both batched forms (the arena on the device)."""
    d = gctx.alloc(data.size + 256)
    out = {}
    try:
        gctx.htod(d, data)
        b = gcompress.Batch(gctx, d, data.size + 256, inputs)
        try:
            for name, mode in MODES:
                _say(f"batched {name}: {len(inputs)} inputs")
                b.set_mode(mode)
                b.run()
                gctx.sync()
                b.check()
                out[name] = [F.batch_digest(b, i, algo) for i, (_, _, algo, _) in enumerate(inputs)]
        finally:
            b.close()
    finally:
        gctx.free(d)
    return out


def test_config2_dropin_zstd_1GiB_L1_L2(gctx):
    _say("config 2: generating 1 GiB")
    data, inputs = W.config2()
    assert W.inputs_sha256(data, inputs) == GOLDEN["config2"]["sha256"]
    want = [F.oracle_zstd(data[o:o + s], lv) for (o, s, _, lv) in inputs]
    _say("config 2: plain per chunk")
    for k, (o, s, _, lv) in enumerate(inputs):
        assert F.plain_zstd(gctx, data[o:o + s], lv) == want[k], f"plain chunk at {o}, level {lv}"
    got = _batched(gctx, data, inputs)
    for name, _ in MODES:
        bad = [k for k in range(len(inputs)) if got[name][k] != want[k]]
        assert not bad, f"{name}: inputs {bad[:5]} differ from the oracle"
    _say("config 2: NVZS round trip")
    blob = gcompress.nvzs_compress(data, 1, threads=F.threads())
    assert gcompress.container_info(blob)[1:] == (data.size, 65536, data.size // 65536)
    assert gcompress.decompress(blob) == data.tobytes()
    ratio = data.size / len(blob)
    _say(f"config 2: NVZS ratio {ratio:.4f}")
    assert ratio > 2.0


def test_config3_dropin_lzma2_640MiB(gctx):
    _say("config 3: generating 640 MiB")
    data, inputs = W.config3()
    assert W.inputs_sha256(data, inputs) == GOLDEN["config3"]["sha256"]
    want = []
    for k, (o, s, _, _) in enumerate(inputs):
        x = data[o:o + s]
        _say(f"config 3: chunk {k} oracle + plain")
        ref = oracle.lzma2_packed(x, threads=F.threads())
        want.append(F.digest(ref))
        got = gctx.lzma2_packed(x)
        assert np.array_equal(got, ref), f"chunk {k}: packed matches differ"
        tri = gctx.lzma2_triage(x)
        nsub = -(-s // 65536)
        want_tri = np.array([ref[j * 65536:(j + 1) * 65536].any() for j in range(nsub)], np.uint8)
        assert np.array_equal(tri, want_tri), f"chunk {k}: triage differs"
        if k == 0:
            # the reference's GpuMatchData layout on a whole 128 MiB chunk (68 B per position)
            md = gcompress.gpu_find_matches(x)
            dist, lens = md.match_distances.reshape(-1, 8), md.match_lengths.reshape(-1, 8)
            assert np.array_equal(md.match_counts[:s], (ref != 0).astype(np.uint32))
            assert not md.match_counts[s:].any()
            assert np.array_equal(dist[:s, 0], ref >> 16) and np.array_equal(lens[:s, 0], ref & 0xFFFF)
            assert not dist[:, 1:].any() and not lens[:, 1:].any() and not dist[s:].any()
            del md, dist, lens
        del ref, got
    got = _batched(gctx, data, inputs)
    for name, _ in MODES:
        assert got[name] == want, f"{name}: batched LZMA2 matches differ from the oracle"
    _say("config 3: NVLZ from the GPU's matches (range-coder pool), round trip")
    blob = gcompress.nvlz_compress_matches(data, 128, threads=min(F.threads(), 8))
    assert gcompress.decompress(blob) == data.tobytes()
    _say(f"config 3: NVLZ ratio {data.size / len(blob):.4f}")
    # the reference's consumer (triage + liblzma preset 9 per sub-block), on the first 8 MiB: it runs
    # at a few MB/s (a 64 MiB-dictionary encoder set up per 64 KiB call)
    head = data[: 8 << 20]
    blob9 = gcompress.nvlz_compress(head, 1, 128, threads=min(F.threads(), 8))
    assert gcompress.decompress(blob9) == head.tobytes()


def test_config4_batched_4096_inputs(gctx):
    _say("config 4: generating 4096 inputs")
    data, inputs = W.config4(threads=F.threads())
    assert W.inputs_sha256(data, inputs) == GOLDEN["config4"]["sha256"]
    got = _batched(gctx, data, inputs)
    _say("config 4: oracle and plain per input")
    bad = []
    for i, (o, s, algo, lv) in enumerate(inputs):
        x = data[o:o + s]
        if algo == gcompress.ALGO_LZMA2:
            want, plain = F.oracle_lzma2(x), F.plain_lzma2(gctx, x)
        else:
            want, plain = F.oracle_zstd(x, lv), F.plain_zstd(gctx, x, lv)
        if not (got["static"][i] == got["persistent"][i] == plain == want):
            bad.append(i)
    assert not bad, f"inputs {bad[:8]} differ (of {len(bad)})"
