"""The kernels' lane-level algorithms (tests/kernel_model.py) equal the oracle on the CPU.

This checks the window / strip logic of gc_zstd_info_kernel and gc_zstd_parse_kernel without a GPU
(long matches that slide the window past its ahead registers, repeat offsets carried across slides,
lazy steps at the window edge)."""
import numpy as np
import pytest

import datagen
import gcompress
import kernel_model as km
import oracle


def inputs():
    rng = np.random.default_rng(11)
    out = [("mix", gcompress.corpus(12000, 3, rng.dirichlet([1, 1, 1, 1])))]
    out.append(("backrefs", datagen.backrefs(12000, 4)))
    out.append(("structured", datagen.structured(12000, 5)))
    out.append(("text", datagen.text(12000, 6)))
    out.append(("zeros+tail", np.concatenate([np.zeros(3000, np.uint8), datagen.random_bytes(500, 7)])))
    long_rep = np.concatenate([datagen.random_bytes(200, 8)] * 30)  # period 200: matches of ~5800 bytes
    out.append(("period200", long_rep))
    out.append(("period70", np.resize(datagen.random_bytes(70, 9), 9000)))
    out.append(("short", datagen.mix(100, 10)))
    # long literal stretches (the look-ahead): random bytes with a few planted repeats, at repeat
    # offsets 1/4/8 and at new offsets, some straddling a 4-position lane group or a 256-position step
    r = datagen.random_bytes(20000, 12).copy()
    for at, off, ln in [(700, 4, 5), (3001, 8, 4), (5000, 1, 6), (7777, 300, 9), (9300, 9300 - 7777, 7),
                        (12543, 1, 3), (15000, 4000, 12), (19990, 5, 6)]:
        r[at:at + ln] = r[at - off:at - off + ln] if off > ln else np.resize(r[at - off:at], ln)
    out.append(("random+planted", r))
    out.append(("random", datagen.random_bytes(9000, 13)))
    out.append(("text+random", np.concatenate([datagen.text(3000, 14), datagen.random_bytes(6000, 15),
                                               datagen.text(2000, 16)])))
    return out


CASES = inputs()


@pytest.mark.parametrize("name,data", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("strip", [4, 8])
def test_info_strips_equal_direct(name, data, strip):
    d = data.tolist()
    n = len(d)
    cand = km.cands(d)
    info = km.zstd_info(d, strip)
    for p in range(n - 3):
        c = cand[p] if cand[p] is not None else 0
        ln = km.common(d, c, p, 3, min(n - p, km.SHORT)) if p > 0 and km.v3(d, c) == km.v3(d, p) else 0
        assert info[p] == (ln << 16 | ((p - c) & 0xFFFF)), p


@pytest.mark.parametrize("name,data", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("depth", [1, 16])
@pytest.mark.parametrize("form", ["sliding", "aligned", "aligned+scan"])
def test_parse_window_equals_oracle(name, data, depth, form):
    d = data.tolist()
    if form == "sliding":
        got = km.zstd_parse(d, km.zstd_info(d), depth)
    else:
        got = km.zstd_parse_fixed(d, km.zstd_info(d), depth, scan=form == "aligned+scan")
    seq, sc, _, _ = oracle.zstd_match_find(data, depth=depth)
    assert got == [tuple(x) for x in seq[0, : sc[0]].tolist()]
