"""Streaming container reader (gc_decompress_file, decompress.rs:9-143 pipelined) on the CPU: containers
are assembled here from CPU-made frames (oracle sequences through libzstd, oracle matches through the
LZMA2 coder, liblzma level 0 on dictionary chunks), written to files, and read back through the reader
with several thread counts; the whole-buffer reader (gc_decompress) must agree; damaged files fail."""
import os
import struct

import numpy as np
import pytest

import datagen
import gcompress
import oracle

SUB = 65536


def nvzs_blob(data):
    seq, sc, _, _ = oracle.zstd_match_find(data, depth=16)
    rows = oracle.zstd_as_zstd_sequences(seq, sc)
    packed = np.concatenate(rows) if rows else np.zeros((0, 4), np.uint32)
    frames, sizes = gcompress.zstd_encode_sequences(data, packed, sc, 2)
    nch = len(frames) if data.size else 0
    head = b"NVZS" + struct.pack("<QQQ", data.size, SUB, nch) + b"".join(struct.pack("<Q", s) for s in sizes[:nch])
    return head + b"".join(frames[:nch])


def nvlz_blob(data):
    streams, sizes = gcompress.cpu_encode_from_matches(data, oracle.lzma2_packed(data), threads=2)
    nch = len(streams)
    head = b"NVLZ" + struct.pack("<QQQ", data.size, SUB, nch) + b"".join(struct.pack("<Q", s) for s in sizes)
    return head + b"".join(streams)


def _data(name):
    return {"mix": datagen.mix(5 * SUB + 321, 41), "text": datagen.text(3 * SUB, 42),
            "zero": np.zeros(2 * SUB, np.uint8), "random": datagen.random_bytes(SUB + 9, 43),
            "one": np.array([5], np.uint8), "empty": np.zeros(0, np.uint8)}[name]


@pytest.mark.parametrize("threads", [1, 2, 5])
@pytest.mark.parametrize("kind", ["nvzs", "nvlz", "nvlz_l0"])
@pytest.mark.parametrize("name", ["mix", "text", "zero", "random", "one", "empty"])
def test_stream_round_trip(tmp_path, kind, name, threads):
    data = _data(name)
    if kind == "nvzs":
        blob = nvzs_blob(data)
    elif kind == "nvlz":
        blob = nvlz_blob(data)
    else:
        blob = gcompress.nvlz_compress_l0(data, dict_mb=1, preset=0, threads=2)
    src, dst = tmp_path / "in.nv", tmp_path / "out.bin"
    src.write_bytes(blob)
    n = gcompress.decompress_file(src, dst, threads)
    assert n == data.size
    assert dst.read_bytes() == data.tobytes()
    assert gcompress.decompress(blob) == data.tobytes()


def test_many_frames_in_order(tmp_path):
    # more frames than the window holds, decoded out of order by 7 threads, written in order
    data = gcompress.corpus(40 * SUB + 17, 44, (1, 2, 3, 4), 4096)
    src, dst = tmp_path / "in.nv", tmp_path / "out.bin"
    src.write_bytes(nvlz_blob(data))
    assert gcompress.decompress_file(src, dst, 7) == data.size
    assert dst.read_bytes() == data.tobytes()


def test_damaged_files(tmp_path):
    data = datagen.text(3 * SUB, 45)
    good = nvzs_blob(data)
    out = tmp_path / "out.bin"
    cases = {
        "magic": b"NVXX" + good[4:],
        "header": good[:20],
        "count": good[:20] + struct.pack("<Q", 7) + good[28:],
        "table": good[:28 + 8],
        "payload": good[:-10],
        "frame": good[:-40] + bytes(40),
        "size": good[:28] + struct.pack("<Q", 1 << 40) + good[36:],
        # a huge chunk of one entry: refused before any allocation of its size
        "huge_chunk": b"NVZS" + struct.pack("<QQQ", 1 << 40, 1 << 40, 1) + struct.pack("<Q", 100) + bytes(100),
        # an entry count the file cannot hold a size table for
        "huge_count": b"NVZS" + struct.pack("<QQQ", (1 << 40) * SUB, SUB, 1 << 40) + bytes(64),
    }
    for name, blob in cases.items():
        p = tmp_path / f"{name}.nv"
        p.write_bytes(blob)
        with pytest.raises(gcompress.GcError):
            gcompress.decompress_file(p, out, 3)
        if name in ("huge_chunk", "huge_count"):
            with pytest.raises(gcompress.GcError, match="out of range|truncated"):
                gcompress.decompress_file(p, out, 3)
    with pytest.raises(gcompress.GcError, match="cannot open"):
        gcompress.decompress_file(tmp_path / "missing.nv", out, 1)
    p = tmp_path / "good.nv"
    p.write_bytes(good)
    with pytest.raises(gcompress.GcError, match="cannot create"):
        gcompress.decompress_file(p, tmp_path / "no" / "such" / "dir", 1)
    assert gcompress.decompress_file(p, out, 3) == data.size and out.read_bytes() == data.tobytes()
    assert os.path.getsize(out) == data.size
