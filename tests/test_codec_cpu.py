"""Consumers and containers on the CPU: oracle sequences through libzstd, sub-block triage through
liblzma, NVZS/NVLZ framing, and the error paths a reader must catch."""
import numpy as np
import pytest

import datagen
import gcompress
import oracle

SUB = 65536


def oracle_packed_seqs(data, depth=16):
    seq, sc, _, _ = oracle.zstd_match_find(data, depth=depth)
    rows = oracle.zstd_as_zstd_sequences(seq, sc)
    packed = np.concatenate(rows) if rows else np.zeros((0, 4), np.uint32)
    return packed, sc


@pytest.mark.parametrize("threads", [1, 4])
@pytest.mark.parametrize("name", ["mix", "text", "zero", "random", "tail"])
def test_zstd_frames_round_trip(name, threads):
    data = {"mix": datagen.mix(3 * SUB + 1234, 1), "text": datagen.text(2 * SUB, 2),
            "zero": np.zeros(SUB + 5, np.uint8), "random": datagen.random_bytes(SUB, 3),
            "tail": datagen.mix(7, 4)}[name]
    seqs, counts = oracle_packed_seqs(data)
    frames, sizes = gcompress.zstd_encode_sequences(data, seqs, counts, threads)
    assert len(frames) == -(-data.size // SUB)
    for s, f in enumerate(frames):
        want = data[s * SUB:(s + 1) * SUB].tobytes()
        assert gcompress.zstd_frame_decode(f, len(want)) == want
        assert sizes[s] == len(f)


def test_zstd_thread_count_does_not_change_bytes():
    data = datagen.mix(5 * SUB, 9)
    seqs, counts = oracle_packed_seqs(data)
    assert gcompress.zstd_encode_sequences(data, seqs, counts, 1) == gcompress.zstd_encode_sequences(data, seqs,
                                                                                                   counts, 3)


def test_zstd_sequences_compress_better_than_raw():
    data = datagen.text(4 * SUB, 5)
    seqs, counts = oracle_packed_seqs(data)
    frames, _ = gcompress.zstd_encode_sequences(data, seqs, counts)
    assert sum(map(len, frames)) < data.size * 0.8


def test_zstd_invalid_sequences_fall_back_to_level1():
    """compress_zstd_custom.rs:173-190: libzstd rejects the sequences -> ZSTD_compress(1) frame.

    libzstd 1.4.8 does not validate offsets by default (an out-of-range offset gives a corrupt
    frame, not an error), so the rejected case here is a literal run longer than the sub-chunk."""
    data = datagen.text(SUB, 6)
    bad = np.array([[5, 70000, 10, 0]], np.uint32)
    frames, _ = gcompress.zstd_encode_sequences(data, bad, np.array([1], np.uint32))
    assert gcompress.zstd_frame_decode(frames[0], SUB) == data.tobytes()


def test_lzma2_triage_raw_and_compressed():
    rnd = datagen.random_bytes(SUB, 1)
    blob = gcompress.encode_single_sub_block(rnd, np.zeros(SUB, np.uint32))
    # raw chunk: control 0x01, size-1 big endian, data, end marker
    assert blob[0] == 0x01 and blob[1:3] == bytes([0xFF, 0xFF]) and blob[-1] == 0 and len(blob) == SUB + 4
    assert gcompress.lzma2_raw_decode(blob, SUB) == rnd.tobytes()
    txt = datagen.text(SUB, 2)
    counts = np.zeros(SUB, np.uint32)
    counts[100] = 1
    blob = gcompress.encode_single_sub_block(txt, counts)
    assert len(blob) < SUB // 2
    assert gcompress.lzma2_raw_decode(blob, SUB) == txt.tobytes()


def test_lzma2_decode_rejects_long_or_trailing_entries():
    # an entry that decodes to more than the expected size, or has bytes after its end marker, is
    # damaged: it is not truncated to fit
    txt = datagen.text(SUB, 6)
    counts = np.ones(SUB, np.uint32)
    blob = gcompress.encode_single_sub_block(txt, counts)
    assert gcompress.lzma2_raw_decode(blob, SUB) == txt.tobytes()
    with pytest.raises(gcompress.GcError, match="past its size"):
        gcompress.lzma2_raw_decode(blob, SUB - 1)
    with pytest.raises(gcompress.GcError, match="after its end marker"):
        gcompress.lzma2_raw_decode(blob + b"\x00\x00", SUB)


def test_lzma2_triage_matches_oracle_counts():
    for data in (datagen.mix(SUB, 3), datagen.random_bytes(SUB, 4), datagen.structured(SUB, 5)):
        _, c = oracle.lzma2_match_find(data, SUB, 8, 32)
        blob = gcompress.encode_single_sub_block(data, c[:data.size])
        assert gcompress.lzma2_raw_decode(blob, data.size) == data.tobytes()


def test_nvlz_l0_container_round_trip():
    data = datagen.mix(3 * SUB + 77, 8)
    blob = gcompress.nvlz_compress_l0(data, dict_mb=1, preset=0, threads=2)
    algo, orig, ch, nch = gcompress.container_info(blob)
    assert (algo, orig, ch, nch) == (gcompress.ALGO_LZMA2, data.size, 1 << 20, 1)
    assert gcompress.decompress(blob) == data.tobytes()


def test_container_errors():
    data = datagen.mix(2 * SUB, 8)
    blob = gcompress.nvlz_compress_l0(data, dict_mb=1, preset=0, threads=1)
    with pytest.raises(gcompress.GcError, match="expected NVZS or NVLZ"):
        gcompress.decompress(b"XXXX" + blob[4:])
    with pytest.raises(gcompress.GcError, match="truncated"):
        gcompress.decompress(blob[:-10])
    with pytest.raises(gcompress.GcError):
        gcompress.decompress(blob[:20])
    bad = bytearray(blob)
    bad[20] = 7  # entry count no longer matches the sizes
    with pytest.raises(gcompress.GcError, match="entry count"):
        gcompress.decompress(bytes(bad))


def test_empty_container():
    blob = gcompress.nvlz_compress_l0(b"", dict_mb=1)
    assert gcompress.container_info(blob)[1:] == (0, 1 << 20, 0)
    assert gcompress.decompress(blob) == b""
