"""Seeded synthetic inputs of the five benchmark configurations (gc_corpus_generate), shared by the
tests and the benchmark so that both run on the same bytes.  Their SHA-256 digests are pinned in
tests/golden/config_sha256.json.

  1  64 MiB four-kind mix (64 KiB segments, equal shares), seed 1: CPU oracle + libzstd.
  2  1 GiB four-kind mix, seed 2, in 16 MiB chunks, each chunk at Zstd L1 and L2.
  3  640 MiB disk-image-like mix (15% random, 25% text, 20% back-references, 40% structured, 1 MiB
     segments, an 8 MiB zero gap after every 96 MiB of content), seed 3, in 128 MiB LZMA2 chunks.
  4  4096 inputs, sizes log-uniform 4 KiB-32 MiB, per-input Dirichlet(1,1,1,1) kind weights, a third
     each Zstd L1 / Zstd L2 / LZMA2 L1 (14.71 GiB realised: the log-uniform mean is ~3.7 MiB).
  5  serving pool: 1 GiB four-kind mix, seed 5, from which requests take their bytes.
"""
from __future__ import annotations

import hashlib
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import ALGO_LZMA2, ALGO_ZSTD, corpus

KiB, MiB, GiB = 1 << 10, 1 << 20, 1 << 30


def config1() -> np.ndarray:
    return corpus(64 * MiB, 1, (1, 1, 1, 1), 65536, True)


def config2(size: int = GiB, chunk: int = 16 * MiB):
    """(data, inputs [(offset, size, algo, level)]): every chunk at L1, then every chunk at L2."""
    data = corpus(size, 2, (1, 1, 1, 1), 65536, True)
    chunks = [(o, min(chunk, size - o)) for o in range(0, size, chunk)]
    return data, [(o, s, ALGO_ZSTD, lv) for lv in (1, 2) for (o, s) in chunks]


def config3(size: int = 640 * MiB, chunk: int = 128 * MiB):
    data = corpus(size, 3, (15, 25, 20, 40), MiB, True, 96 * MiB, 8 * MiB)
    return data, [(o, min(chunk, size - o), ALGO_LZMA2, 1) for o in range(0, size, chunk)]


def config4_layout(n_inputs: int = 4096):
    """Sizes, kind weights, arena offsets (256-byte aligned) and (algo, level) of each input."""
    rng = np.random.default_rng(4)
    sizes = np.exp(rng.uniform(np.log(4 * KiB), np.log(32 * MiB), n_inputs)).astype(np.int64)
    weights = rng.dirichlet([1, 1, 1, 1], n_inputs)
    kinds = [(ALGO_ZSTD, 1), (ALGO_ZSTD, 2), (ALGO_LZMA2, 1)]
    offs = np.concatenate([[0], np.cumsum((sizes + 255) // 256 * 256)])
    return sizes, weights, offs, [kinds[i % 3] for i in range(n_inputs)]


def config4(n_inputs: int = 4096, threads: int = 16):
    """(arena, inputs): input i is corpus(size_i, 1000 + i, weights_i, 64 KiB, drawn kinds) at its
    offset; generated by a thread pool (each input has its own seed, so the bytes do not depend on
    the thread count)."""
    sizes, weights, offs, kinds = config4_layout(n_inputs)
    data = np.zeros(int(offs[-1]), np.uint8)

    def gen(i):
        o, s = int(offs[i]), int(sizes[i])
        data[o: o + s] = corpus(s, 1000 + i, weights[i], 65536, False)

    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(gen, range(n_inputs)))
    inputs = [(int(offs[i]), int(sizes[i]), *kinds[i]) for i in range(n_inputs)]
    return data, inputs


def config5_pool(size: int = GiB) -> np.ndarray:
    return corpus(size, 5, (1, 1, 1, 1), 65536, True)


def inputs_sha256(data: np.ndarray, inputs) -> str:
    """Digest of the inputs' bytes, back to back in list order (distinct inputs only once)."""
    h = hashlib.sha256()
    seen = set()
    for (o, s, _, _) in inputs:
        if (o, s) in seen:
            continue
        seen.add((o, s))
        h.update(memoryview(data[o: o + s]))
    return h.hexdigest()
