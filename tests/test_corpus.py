"""The synthetic corpus generator is deterministic and has the four kinds it names."""
import hashlib

import numpy as np

import gcompress


def test_deterministic_and_seeded():
    a = gcompress.corpus(1 << 20, 1)
    b = gcompress.corpus(1 << 20, 1)
    c = gcompress.corpus(1 << 20, 2)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, c)


def test_kinds():
    rnd = gcompress.corpus(1 << 20, 3, (1, 0, 0, 0))
    counts = np.bincount(rnd, minlength=256)
    assert counts.min() > 3000  # uniform bytes
    txt = gcompress.corpus(1 << 20, 3, (0, 1, 0, 0))
    alpha = set(b"abcdefghijklmnopqrstuvwxyzABCDEFGHIJKLMNOPQRSTUVWXYZ0123456789-_ ")
    assert set(np.unique(txt).tolist()) <= alpha
    st = gcompress.corpus(1 << 20, 3, (0, 0, 0, 1))
    assert (st == 0).mean() > 0.2  # zero runs and zero fields


def test_gaps():
    MiB = 1 << 20
    d = gcompress.corpus(40 * MiB, 3, (15, 25, 20, 40), MiB, True, 12 * MiB, 2 * MiB)
    assert not d[12 * MiB:14 * MiB].any()
    assert not d[26 * MiB:28 * MiB].any()
    assert d[14 * MiB:15 * MiB].any()


def test_golden_sha256():
    """Pinned digests: a change of the generator changes every benchmark input."""
    import json
    import os

    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "corpus_sha256.json")))
    for name, spec in g.items():
        d = gcompress.corpus(spec["n"], spec["seed"], spec["weights"], spec["segment"], spec["exact"],
                             spec["gap_every"], spec["gap_len"])
        assert hashlib.sha256(d.tobytes()).hexdigest() == spec["sha256"], name
