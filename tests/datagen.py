"""Small seeded inputs for the tests (numpy only; the benchmark corpora come
from the library's own generator)."""
import numpy as np


def random_bytes(n, seed=0):
    return np.random.default_rng(seed).integers(0, 256, n, dtype=np.uint8)


def text(n, seed=0, vocab=300):
    rng = np.random.default_rng(seed)
    alpha = np.frombuffer(b"abcdefghijklmnopqrstuvwxyzABCDEFGHIJKLMNOPQRSTUVWXYZ0123456789-_", dtype=np.uint8)
    words = [bytes(alpha[rng.integers(0, 64, rng.integers(3, 13))]) for _ in range(vocab)]
    ranks = np.arange(1, vocab + 1, dtype=np.float64) ** -1.1
    ranks /= ranks.sum()
    out = bytearray()
    while len(out) < n:
        idx = rng.choice(vocab, size=256, p=ranks)
        out += b" ".join(words[i] for i in idx) + b" "
    return np.frombuffer(bytes(out[:n]), dtype=np.uint8).copy()


def backrefs(n, seed=0):
    rng = np.random.default_rng(seed)
    out = bytearray(rng.integers(0, 256, 64, dtype=np.uint8).tobytes())
    while len(out) < n:
        if rng.random() < 0.3:
            out += rng.integers(0, 256, rng.integers(1, 16), dtype=np.uint8).tobytes()
            continue
        dist = int(min(len(out), np.exp(rng.uniform(0, np.log(len(out) + 1)))))
        dist = max(dist, 1)
        ln = int(rng.geometric(1 / 32))
        start = len(out) - dist
        for i in range(ln):
            b = out[start + i]
            if rng.random() < 0.02:
                b = int(rng.integers(0, 256))
            out.append(b)
    return np.frombuffer(bytes(out[:n]), dtype=np.uint8).copy()


def structured(n, seed=0):
    rng = np.random.default_rng(seed)
    out = bytearray()
    counter, drift = 0, 1000
    while len(out) < n:
        if rng.random() < 1 / 64:
            out += bytes(int(rng.geometric(1 / 4096)))
            continue
        counter += 1
        if rng.random() < 0.05:
            drift += int(rng.integers(-2, 3))
        rec = np.array([counter, drift & 0xFFFFFFFF, int(rng.integers(0, 4)), 0xC0FFEE], dtype="<u4")
        out += rec.tobytes()
    return np.frombuffer(bytes(out[:n]), dtype=np.uint8).copy()


KINDS = [random_bytes, text, backrefs, structured]


def mix(n, seed=0, seg=65536):
    """Segments of the four kinds in turn."""
    parts = []
    i = 0
    total = 0
    while total < n:
        k = KINDS[(i + seed) % 4]
        m = min(seg, n - total)
        parts.append(k(m, seed * 1000 + i))
        total += m
        i += 1
    return np.concatenate(parts)


def periodic(n, period, seed=0):
    rng = np.random.default_rng(seed)
    pat = rng.permutation(256)[:period].astype(np.uint8)
    return np.resize(pat, n).astype(np.uint8)
