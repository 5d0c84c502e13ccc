"""End to end on the GPU: match finding -> libzstd / liblzma -> NVZS / NVLZ -> byte-exact round trip."""
import numpy as np
import pytest

import datagen
import gcompress
import oracle

pytestmark = pytest.mark.gpu
SUB = 65536


@pytest.mark.parametrize("level", [1, 2])
def test_compress_chunk_zstd_custom_round_trip(level):
    data = datagen.mix(6 * SUB + 999, 21)
    print(f"[e2e zstd] level={level} n={data.size}", flush=True)
    frames, sizes = gcompress.compress_chunk_zstd_custom(data, 0, level)
    assert len(frames) == 7 and sizes == [len(f) for f in frames]
    for s, f in enumerate(frames):
        want = data[s * SUB:(s + 1) * SUB].tobytes()
        assert gcompress.zstd_frame_decode(f, len(want)) == want
    # the frames are exactly those of the oracle's sequences through the same consumer
    seq, sc, _, _ = oracle.zstd_match_find(data, depth={1: 16, 2: 64}[level])
    rows = np.concatenate(oracle.zstd_as_zstd_sequences(seq, sc))
    ofr, _ = gcompress.zstd_encode_sequences(data, rows, sc, 1)
    assert ofr == frames


def test_packed_sequences_match_full_layout(gctx):
    data = datagen.mix(4 * SUB + 17, 22)
    seqs, counts = gctx.zstd_sequences(data, 1)
    nsub = counts.size
    packed = np.zeros((nsub * gcompress.MAX_SEQ, 4), np.uint32)
    cnt = np.zeros(nsub, np.uint32)
    import ctypes as C
    total = C.c_size_t()
    rc = gcompress.lib().gc_zstd_find_sequences_packed(gctx.h, data.ctypes.data, data.size, 1, packed.ctypes.data,
                                                        packed.shape[0], cnt.ctypes.data, C.byref(total))
    assert rc == 0
    np.testing.assert_array_equal(cnt, counts)
    np.testing.assert_array_equal(packed[: total.value], np.concatenate([seqs[s, : counts[s]] for s in range(nsub)]))


@pytest.mark.parametrize("n", [0, 1, 100, SUB, 3 * SUB + 5])
def test_nvzs_round_trip(n):
    data = datagen.mix(n, 30) if n else np.zeros(0, np.uint8)
    print(f"[e2e nvzs] n={n}", flush=True)
    blob = gcompress.nvzs_compress(data, 1)
    algo, orig, ch, nch = gcompress.container_info(blob)
    assert (algo, orig, ch, nch) == (gcompress.ALGO_ZSTD, n, SUB, -(-n // SUB))
    assert gcompress.decompress(blob) == data.tobytes()


@pytest.mark.parametrize("n", [0, 5, SUB, 5 * SUB + 3])
def test_nvlz_round_trip_and_triage(n):
    data = gcompress.corpus(n, 31, (15, 25, 20, 40), 4096) if n else np.zeros(0, np.uint8)
    print(f"[e2e nvlz] n={n}", flush=True)
    blob = gcompress.nvlz_compress(data, 1, dict_mb=1, threads=4)
    algo, orig, ch, nch = gcompress.container_info(blob)
    assert (algo, orig, ch, nch) == (gcompress.ALGO_LZMA2, n, SUB, -(-n // SUB))
    assert gcompress.decompress(blob) == data.tobytes()
    # each entry is exactly what the reference consumer makes from the oracle's match counts
    if n:
        _, c = oracle.lzma2_match_find(data, SUB, 8, 32)
        sizes = np.frombuffer(blob[28:28 + 8 * nch], "<u8")
        pos = 28 + 8 * nch
        for s in range(nch):
            entry = blob[pos:pos + sizes[s]]
            sub = data[s * SUB:(s + 1) * SUB]
            assert entry == gcompress.encode_single_sub_block(sub, c[s * SUB:s * SUB + sub.size])
            pos += sizes[s]


@pytest.mark.parametrize("n", [0, 5, SUB, 3 * SUB + 99])
def test_nvlz_from_gpu_matches_round_trip(n):
    """NVLZ whose sub-blocks are coded from the GPU's matches: the container reads back, and each entry
    is exactly the stream the CPU encoder makes from the oracle's matches (the GPU's are identical)."""
    data = gcompress.corpus(n, 32, (15, 25, 20, 40), 4096) if n else np.zeros(0, np.uint8)
    print(f"[e2e nvlz matches] n={n}", flush=True)
    blob = gcompress.nvlz_compress_matches(data, dict_mb=1, threads=3)
    algo, orig, ch, nch = gcompress.container_info(blob)
    assert (algo, orig, ch, nch) == (gcompress.ALGO_LZMA2, n, SUB, -(-n // SUB))
    assert gcompress.decompress(blob) == data.tobytes()
    if n:
        streams, _ = gcompress.cpu_encode_from_matches(data, oracle.lzma2_packed(data), threads=1)
        assert blob[28 + 8 * nch:] == b"".join(streams)


@pytest.mark.parametrize("algo", ["nvzs", "nvlz", "nvlz_matches"])
def test_containers_through_the_streaming_reader(tmp_path, algo):
    """GPU-made containers written to a file and read back by the pipelined reader (decompress.rs:9-143)."""
    data = gcompress.corpus(6 * SUB + 1000, 33, (1, 1, 1, 1), 4096)
    print(f"[e2e stream] {algo}", flush=True)
    blob = {"nvzs": lambda: gcompress.nvzs_compress(data, 1, threads=2),
            "nvlz": lambda: gcompress.nvlz_compress(data, 1, dict_mb=1, threads=2),
            "nvlz_matches": lambda: gcompress.nvlz_compress_matches(data, dict_mb=1, threads=2)}[algo]()
    src, dst = tmp_path / "c.nv", tmp_path / "d.bin"
    src.write_bytes(blob)
    assert gcompress.decompress_file(src, dst, 4) == data.size
    assert dst.read_bytes() == data.tobytes()


@pytest.mark.parametrize("n", [1, 3 * 65536 + 17])
def test_gpu_containers_into_caller_buffer(n):
    """The GPU containers written into an output buffer the caller keeps equal the returned bytes, on two
    calls in a row into the same buffer (the second over a shorter input)."""
    x = datagen.mix(n, seed=n)
    buf_z = np.full(gcompress.lib().gc_nvzs_gpu_bound(n), 0xAB, np.uint8)
    buf_l = np.full(gcompress.lib().gc_container_bound(n, gcompress.ALGO_LZMA2), 0xAB, np.uint8)
    for y in (x, x[: max(1, n // 2)]):
        m = gcompress.nvzs_compress_gpu(y, 1, out=buf_z)
        assert buf_z[:m].tobytes() == gcompress.nvzs_compress_gpu(y, 1)
        m = gcompress.nvlz_compress_gpu(y, 128, out=buf_l)
        blob = gcompress.nvlz_compress_gpu(y, 128)
        assert buf_l[:m].tobytes() == blob
        assert gcompress.decompress(blob) == y.tobytes()
