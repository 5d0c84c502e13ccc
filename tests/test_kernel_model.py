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
        assert info[p] == (ln << 16 | c), p


@pytest.mark.parametrize("name,data", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("depth", [1, 16])
@pytest.mark.parametrize("form", ["sliding", "aligned"])
def test_parse_window_equals_oracle(name, data, depth, form):
    d = data.tolist()
    parse = km.zstd_parse if form == "sliding" else km.zstd_parse_fixed
    got = parse(d, km.zstd_info(d), depth)
    seq, sc, _, _ = oracle.zstd_match_find(data, depth=depth)
    assert got == [tuple(x) for x in seq[0, : sc[0]].tolist()]
