"""The segmented parse (tests/segment_model.py, the protocol of gc_zstd_parse_seg_kernel) gives the
serial parse's sequences exactly: against the oracle on the four kinds of data, ragged sizes, the
kernel's 16 segments and other counts, both depths and the sequence cap; and the fix-up rounds end
where the protocol says they do."""
import numpy as np
import pytest

import datagen
import gcompress
import oracle
import segment_model as M


def serial(data, depth, max_seq=None):
    seq, sc, _, _ = oracle.zstd_match_find(data, depth=depth, max_seq=max_seq)
    out, pos = [], 0
    for ll, ml3, off in seq[0, : sc[0]]:
        pos += int(ll)
        out.append((pos, int(ml3) + 3, int(off)))
        pos += int(ml3) + 3
    return out


CASES = [("mix", lambda: gcompress.corpus(65536, 7, (1, 1, 1, 1), 4096, False)),
         ("text", lambda: datagen.text(65536, 8)),
         ("structured", lambda: gcompress.corpus(65536, 9, (0, 0, 0, 1), 65536, True)),
         ("random", lambda: datagen.random_bytes(65536, 10)),
         ("backrefs", lambda: gcompress.corpus(65536, 11, (0, 0, 1, 0), 65536, True)),
         ("ragged", lambda: gcompress.corpus(40001, 12, (1, 2, 1, 3), 2048, False)),
         ("zeros", lambda: np.zeros(30000, np.uint8)),
         ("tiny", lambda: datagen.text(70, 15))]


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("segments", [16, 7, 64])
def test_segmented_equals_serial(name, make, segments):
    data = make()
    for depth in (16, 1):
        got, rounds = M.segmented(data, segments, depth)
        assert got == serial(data, depth), (name, segments, depth, rounds)
        assert rounds <= segments - 1


def test_bounds():
    assert M.bounds(65536, 16) == ([4096 * j for j in range(16)], [4096 * (j + 1) for j in range(16)])
    starts, ends = M.bounds(1000, 16)  # 64-byte segments, the last ones empty
    assert starts[:3] == [0, 64, 128] and ends[15] == 1000 and starts[15] == 960
    starts, ends = M.bounds(70, 16)
    assert ends[0] == 64 and ends[1] == 70 and starts[2:] == [70] * 14


def test_sequence_cap():
    data = datagen.text(65536, 13)
    want = serial(data, 16, max_seq=500)
    assert len(want) == 500
    got, _ = M.segmented(data, 16, 16, max_seq=500)
    assert got == want


def test_guess_that_is_wrong_is_fixed():
    # a periodic block whose period is not among the initial offsets: every speculative start guesses
    # wrong offsets, and the fix-ups still give the serial parse, each meeting its earlier walk
    base = datagen.random_bytes(37, 14)
    data = np.tile(base, 65536 // 37 + 1)[:65536]
    stats = []
    got, rounds = M.segmented(data, 16, 16, stats=stats)
    assert got == serial(data, 16)
    assert rounds == 1 and all(met for (_, _, _, met) in stats)


def test_long_match_spans_segments():
    # one match covers most of the block: the segments it spans are empty, their end states pass on
    data = np.concatenate([datagen.random_bytes(5000, 16)] * 13)[:65536]
    got, _ = M.segmented(data, 16, 16)
    assert got == serial(data, 16)


def test_offsets_carried_through_transparent_segments():
    # random bytes with a rare planted repeat: the offsets it makes current travel through every later
    # segment (no sequence there to meet an earlier walk at); they are carried through the segments
    # without sequences at once, so one fix-up round redoes every later segment together
    data = datagen.random_bytes(65536, 17).copy()
    data[3000:3010] = data[1000:1010]
    stats = []
    got, rounds = M.segmented(data, 16, 16, stats=stats)
    assert got == serial(data, 16)
    assert rounds == 1 and sorted(j for (_, j, _, _) in stats) == list(range(1, 16))


def test_transparent_guess_that_fails():
    # 3-byte repeats at an offset that only the carried offsets hold (their 4th bytes differ, so no hash
    # candidate finds them): segments that found nothing with the guessed offsets find sequences with
    # the carried ones, the speculation past each is redone in a further round, and the result is exact
    data = datagen.random_bytes(65536, 18).copy()
    data[3000:3010] = data[1000:1010]  # makes offset 2000 current
    for at in (30000, 45000, 52000):
        data[at:at + 3] = data[at - 2000:at - 1997]
        data[at + 3] = data[at - 1997] ^ 0x55
    want = serial(data, 16)
    assert [h for (h, ln, off) in want if off == 2000] == [3000, 30000, 45000, 52000]
    got, rounds = M.segmented(data, 16, 16)
    assert got == want
    assert 2 <= rounds <= 15
