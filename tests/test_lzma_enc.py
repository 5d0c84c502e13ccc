"""LZMA2 coded from match lists (gc_lzma2_encode_matches, gc_cpu_encode_from_matches): every stream is
decoded by Python's lzma module (a decoder this library does not contain) and must give back the input.
The matches are the oracle's (the GPU's are bit-identical to them, tests/test_gpu_parity.py).  Encoded
bytes have no reference vectors (the reference ships none): parity here is the round trip, plus the
coding decisions that the format fixes (raw chunks for incompressible data, one chunk per sub-block)."""
import lzma

import numpy as np
import pytest

import datagen
import gcompress
import oracle

SUB = 65536


def decode(stream: bytes) -> bytes:
    return lzma.decompress(stream, format=lzma.FORMAT_RAW,
                           filters=[{"id": lzma.FILTER_LZMA2, "dict_size": 1 << 20}])


def _cases():
    rng = np.random.default_rng(5)
    # repeat-distance traffic: the same few distances over and over, with fresh bytes between
    base = rng.integers(0, 256, 4096, dtype=np.uint8)
    reps = [base]
    for i in range(120):
        d = int([8, 24, 24, 100, 8][i % 5])
        reps.append(np.concatenate(reps)[-d:][: int(rng.integers(2, 40))])
        reps.append(rng.integers(0, 256, int(rng.integers(1, 4)), dtype=np.uint8))
    return [
        ("text", datagen.text(SUB, 1)),
        ("backrefs", datagen.backrefs(SUB, 2)),
        ("structured", datagen.structured(SUB, 3)),
        ("random", datagen.random_bytes(SUB, 4)),
        ("zero", np.zeros(SUB, np.uint8)),
        ("period1", datagen.periodic(SUB, 1, 1)),
        ("period3", datagen.periodic(SUB, 3, 3)),
        ("period200", datagen.periodic(SUB, 200, 5)),
        ("reps", np.concatenate(reps)[:SUB]),
        ("len1", datagen.text(1, 6)),
        ("len2", np.array([7, 7], np.uint8)),
        ("len5", np.zeros(5, np.uint8)),
        ("odd", datagen.text(SUB - 1, 7)),
        ("short", datagen.mix(3000, 8)),
    ]


CASES = _cases()


@pytest.mark.parametrize("verify", [False, True])
@pytest.mark.parametrize("name,data", CASES, ids=[c[0] for c in CASES])
def test_sub_block_round_trip(name, data, verify):
    enc = gcompress.lzma2_encode_matches(data, oracle.lzma2_packed(data), verify=verify)
    assert decode(enc) == data.tobytes()
    assert enc[-1] == 0x00  # end marker
    assert len(enc) <= data.size + 3 * -(-data.size // SUB) + 1  # never above raw chunks


def test_coding_decisions():
    # incompressible: raw chunks (control 0x01); compressible: one LZMA chunk with dictionary reset,
    # new properties (lc 3, lp 0, pb 2) and the chunk's sizes in its header
    rnd = datagen.random_bytes(SUB, 11)
    assert gcompress.lzma2_encode_matches(rnd, oracle.lzma2_packed(rnd))[0] == 0x01
    txt = datagen.text(SUB, 12)
    enc = gcompress.lzma2_encode_matches(txt, oracle.lzma2_packed(txt))
    assert enc[0] == 0xE0 | ((SUB - 1) >> 16)
    assert ((enc[0] & 0x1F) << 16 | enc[1] << 8 | enc[2]) + 1 == SUB
    assert (enc[3] << 8 | enc[4]) + 1 == len(enc) - 7
    assert enc[5] == (2 * 5 + 0) * 9 + 3
    # a zero run is a handful of bytes: one literal then repeat matches at distance 1
    z = np.zeros(SUB, np.uint8)
    assert len(gcompress.lzma2_encode_matches(z, oracle.lzma2_packed(z))) < 128


def test_matches_drive_the_ratio():
    # the same data coded with its matches and with no matches at all: the matches must pay
    txt = datagen.text(SUB, 13)
    with_m = gcompress.lzma2_encode_matches(txt, oracle.lzma2_packed(txt))
    without = gcompress.lzma2_encode_matches(txt, np.zeros(SUB, np.uint32))
    assert decode(without) == txt.tobytes()
    assert len(with_m) < 0.6 * len(without)


def test_wrong_matches_fall_back_when_verified():
    # match words that point at bytes that differ: the stream decodes to something else, the
    # verification catches it and raw chunks replace it
    data = datagen.random_bytes(SUB, 14)
    rng = np.random.default_rng(3)
    pos = np.arange(SUB, dtype=np.uint64)
    dist = np.minimum(rng.integers(1, 1000, SUB).astype(np.uint64), np.maximum(pos, 1))
    bogus = ((dist << 16) | 20).astype(np.uint32)
    bogus[:4] = 0
    enc = gcompress.lzma2_encode_matches(data, bogus, verify=True)
    assert decode(enc) == data.tobytes()
    assert enc[0] == 0x01


@pytest.mark.parametrize("threads", [1, 3])
def test_chunk_streams(threads):
    data = np.concatenate([datagen.text(SUB, 15), datagen.random_bytes(SUB, 16), np.zeros(SUB, np.uint8),
                           datagen.structured(SUB + 777, 17)])
    streams, sizes = gcompress.cpu_encode_from_matches(data, oracle.lzma2_packed(data), threads=threads)
    assert len(streams) == -(-data.size // SUB)
    for s, st in enumerate(streams):
        assert len(st) == sizes[s]
        assert decode(st) == data[s * SUB:(s + 1) * SUB].tobytes()
    # thread count does not change the bytes
    assert streams == gcompress.cpu_encode_from_matches(data, oracle.lzma2_packed(data), threads=1)[0]


def test_against_preset9_on_the_mix():
    # a single candidate per position and a greedy parse against liblzma preset 9's full search: within
    # a bounded factor of its size on the disk-image-like mix (measured 3.41 vs 3.74 on 16 MiB)
    data = gcompress.corpus(8 * SUB, 3, (15, 25, 20, 40), 1 << 20, True, 0, 0)
    pk = oracle.lzma2_packed(data)
    ours = sum(gcompress.cpu_encode_from_matches(data, pk)[1])
    p9 = sum(len(gcompress.encode_single_sub_block(data[s * SUB:(s + 1) * SUB],
                                                    (pk[s * SUB:(s + 1) * SUB] != 0).astype(np.uint32)))
             for s in range(8))
    assert p9 <= ours <= 1.35 * p9
