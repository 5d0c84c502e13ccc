"""The oracle's Zstd frame coder (oracle/gc_oracle.c orc_zstd_frame, the checker of the GPU frame coder)
writes frames that stock libzstd decodes back to the input, on every data kind, sub-block sizes from 1
byte to 64 KiB and sequence sets from none to the cap; every choice it can make is made somewhere in
the set (predefined, RLE and FSE_Compressed tables in each dimension; raw, RLE and Huffman literals;
repeat-offset codes with and without literals before them), read back from the frames' own headers;
and the ratio on the Zstd mix is at least libzstd's own with the same sequences less 2%."""
import numpy as np
import pytest

import datagen
import gcompress
import oracle
from gcompress import workloads as W

SUB = 65536


def frames_for(data, depth=16, max_seq=None):
    seq, sc, _, _ = oracle.zstd_match_find(data, depth=depth, max_seq=max_seq)
    rows = oracle.zstd_as_zstd_sequences(seq, sc)
    return [(data[s * SUB:(s + 1) * SUB], rows[s], oracle.zstd_frame(data[s * SUB:(s + 1) * SUB], rows[s]))
            for s in range(len(rows))]


def parse(frame):
    """(block type, literals type, (LL, OF, ML) modes or None): the frame's own headers."""
    b = frame[4]
    o = 5 + (2 if b == 0x60 else 1)
    bh = frame[o] | frame[o + 1] << 8 | frame[o + 2] << 16
    btype = (bh >> 1) & 3
    if btype != 2:
        return btype, None, None
    o += 3
    h0 = frame[o]
    ltype, sf = h0 & 3, (h0 >> 2) & 3
    if ltype in (0, 1):
        hs = 1 if sf in (0, 2) else (2 if sf == 1 else 3)
        v = int.from_bytes(frame[o:o + hs], "little")
        size = v >> 3 if hs == 1 else v >> 4
        o += hs + (size if ltype == 0 else 1)
    else:
        hs = {0: 3, 1: 3, 2: 4, 3: 5}[sf]
        v = int.from_bytes(frame[o:o + hs], "little")
        bits = {3: 10, 4: 14, 5: 18}[hs]
        csize = (v >> (4 + bits)) & ((1 << bits) - 1)
        o += hs + csize
    nseq = frame[o]
    if nseq == 0:
        return btype, ltype, None
    o += 1 if nseq < 128 else 2
    m = frame[o]
    return btype, ltype, ((m >> 6) & 3, (m >> 4) & 3, (m >> 2) & 3)


def _corpus():
    out = [("text", datagen.text(4 * SUB, 1)), ("mix", datagen.mix(4 * SUB + 99, 4)),
           ("random", datagen.random_bytes(SUB + 5, 3)), ("zero", np.zeros(SUB, np.uint8)),
           ("structured", datagen.structured(2 * SUB, 3)), ("backrefs", datagen.backrefs(2 * SUB, 2)),
           ("period5", datagen.periodic(SUB, 5, 5)), ("period37", datagen.periodic(SUB, 37, 6)),
           ("highbytes", (datagen.text(SUB, 7) | 0x80).astype(np.uint8)),
           ("corpus", gcompress.corpus(4 * SUB, 8, (1, 1, 1, 1), 4096, False))]
    for n in (1, 2, 3, 4, 5, 31, 32, 63, 64, 65, 100, 255, 256, 257, 1023, 1024, 4095, 4096, 16383, 16384, 40000):
        out.append((f"text{n}", datagen.text(n, n)))
    # literals of one byte value between matches: an RLE literals block
    r = datagen.random_bytes(4096, 9)
    out.append(("rlelits", np.concatenate([np.concatenate([r, np.full(7, 65, np.uint8)]) for _ in range(12)])))
    return out


CORPUS = _corpus()


@pytest.mark.parametrize("name,data", CORPUS, ids=[c[0] for c in CORPUS])
@pytest.mark.parametrize("depth", [16, 1])
def test_frames_decode(name, data, depth):
    for sub, rows, f in frames_for(data, depth):
        assert gcompress.zstd_frame_decode(f, sub.size) == sub.tobytes()


def test_sequence_cap_and_no_sequences():
    data = datagen.text(2 * SUB, 11)
    for max_seq in (0, 1, 2, 100):
        for sub, rows, f in frames_for(data, 16, max_seq):
            assert rows.shape[0] <= max_seq
            assert gcompress.zstd_frame_decode(f, sub.size) == sub.tobytes()


def test_every_choice_is_made():
    seen_lit, seen_mode, reps = set(), set(), {"ll>0": 0, "ll=0": 0}
    for name, data in CORPUS:
        for sub, rows, f in frames_for(data):
            btype, ltype, modes = parse(f)
            if btype == 2:
                seen_lit.add(ltype)
            if modes:
                seen_mode.update((d, m) for d, m in zip("LOM", modes))
            # repeat offsets the format's history offers for these rows (RFC 8878 3.1.2.5)
            r = [1, 4, 8]
            for off, ll, _, _ in rows:
                cands = (r[0], r[1], r[2]) if ll > 0 else (r[1], r[2], r[0] - 1)
                if off in cands:
                    reps["ll>0" if ll > 0 else "ll=0"] += 1
                if ll > 0 and off == r[0]:
                    continue
                k = 1 if off == r[1] else (2 if off == r[2] else 0)
                r = [off, r[0], r[2]] if k == 1 else [off, r[0], r[1]]
    assert seen_lit == {0, 1, 2}, seen_lit
    assert {(d, m) for d in "LOM" for m in (0, 1, 2)} <= seen_mode, seen_mode
    assert reps["ll>0"] > 100 and reps["ll=0"] > 0, reps


def test_ratio_on_the_zstd_mix():
    data, inputs = W.config2()
    o, s, _, _ = inputs[0]
    x = data[o:o + 64 * SUB]
    total = sum(len(f) for _, _, f in frames_for(x))
    seq, sc, _, _ = oracle.zstd_match_find(x)
    rows = oracle.zstd_as_zstd_sequences(seq, sc)
    frames, _ = gcompress.zstd_encode_sequences(x, np.concatenate(rows), sc, threads=4)
    ref = sum(len(f) for f in frames)
    assert x.size / total >= 0.98 * x.size / ref, (x.size / total, x.size / ref)
    assert x.size / total >= 2.25
