"""Configuration 1 at its full size on the CPU (the plain method: oracle sequences of the reference
kernel on the 64 MiB seed-1 mix, entropy-coded by libzstd ZSTD_compressSequences, one frame per 64 KiB
sub-chunk, round trip through an NVZS container), and the pinned digests of the other configurations'
inputs (configuration 4's is checked where its 14.7 GiB are generated, in the GPU test)."""
import hashlib
import json
import os
import struct

import numpy as np

import fullsize
import gcompress
import oracle
from gcompress import workloads as W

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "config_sha256.json")))


def test_config1_cpu_plain_method_round_trip():
    data = W.config1()
    assert hashlib.sha256(data.tobytes()).hexdigest() == GOLDEN["config1"]["sha256"]
    t = fullsize.threads()
    seq, sc, _, _ = oracle.zstd_match_find(data, depth=16, threads=t)
    rows = np.concatenate(oracle.zstd_as_zstd_sequences(seq, sc))
    frames, sizes = gcompress.zstd_encode_sequences(data, rows, sc, threads=t)
    assert len(frames) == 1024
    # every frame decodes to its sub-chunk (libzstd, the reader's decoder)
    for s in (0, 1, 511, 1023):
        assert gcompress.zstd_frame_decode(frames[s], 65536) == data[s * 65536:(s + 1) * 65536].tobytes()
    blob = (b"NVZS" + struct.pack("<QQQ", data.size, 65536, len(frames)) +
            b"".join(struct.pack("<Q", z) for z in sizes) + b"".join(frames))
    assert gcompress.decompress(blob) == data.tobytes()
    ratio = data.size / sum(sizes)
    assert 2.0 < ratio < 3.0, ratio
    # one encode thread (the reference's shape) gives the same bytes
    f1, _ = gcompress.zstd_encode_sequences(data[:4 << 20], np.concatenate(
        oracle.zstd_as_zstd_sequences(seq[:64], sc[:64])), sc[:64], threads=1)
    assert f1 == frames[:64]


def test_config_input_digests():
    for name, make in (("config2", W.config2), ("config3", W.config3)):
        data, inputs = make()
        assert W.inputs_sha256(data, inputs) == GOLDEN[name]["sha256"], name
        assert data.size == GOLDEN[name]["bytes"]
    pool = W.config5_pool()
    assert hashlib.sha256(pool.tobytes()).hexdigest() == GOLDEN["config5_pool"]["sha256"]
    sizes, _, _, kinds = W.config4_layout()
    assert int(sizes.sum()) == GOLDEN["config4"]["bytes"] and len(kinds) == GOLDEN["config4"]["inputs"]
