"""Helpers of the full-size configuration tests: results of the oracle, the plain per-chunk GPU form
and the batched form in one canonical layout (per-sub-block counts, then the used records as
(offset, litLength, matchLength) rows; LZMA2: one packed u32 per position), compared by SHA-256 so
that a result needs to be held only until it is hashed."""
import ctypes as C
import hashlib
import os

import numpy as np

import gcompress
import oracle

MAX_SEQ = 65536 // 3


def threads() -> int:
    """Host threads for the oracle and the CPU encoders (the host's CPU share, at most 16)."""
    return max(1, min(len(os.sched_getaffinity(0)), 16))


def digest(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def used_rows(slots: np.ndarray, counts: np.ndarray) -> np.ndarray:
    """The used records of per-sub-block slot arrays [n_sub, max_seq, k], in order."""
    mask = np.arange(slots.shape[1])[None, :] < counts[:, None]
    return slots[mask]


def oracle_zstd(x: np.ndarray, level: int) -> str:
    seq, sc, _, _ = oracle.zstd_match_find(x, depth={1: 16, 2: 64}.get(level, 1), threads=threads())
    r = used_rows(seq, sc)  # literal_length, match_length - 3, offset
    rows = np.stack([r[:, 2], r[:, 0], r[:, 1] + 3], axis=1).astype(np.uint32)
    return digest(sc.astype(np.uint32), rows)


def oracle_lzma2(x: np.ndarray) -> str:
    return digest(oracle.lzma2_packed(x, threads=threads()))


def slots_zstd(seqs: np.ndarray, counts: np.ndarray) -> str:
    """ZSTD_Sequence slots [n_sub, max_seq, 4] (rep must be 0) -> digest."""
    r = used_rows(seqs, counts)
    assert not r[:, 3].any(), "rep field set"
    return digest(counts.astype(np.uint32), np.ascontiguousarray(r[:, :3]))


def plain_zstd(ctx, x: np.ndarray, level: int) -> str:
    """The plain drop-in (gc_zstd_find_sequences_packed: one chunk, host in, host out)."""
    L = gcompress.lib()
    nsub = -(-x.size // 65536)
    cap = nsub * MAX_SEQ
    seqs = np.empty((cap, 4), np.uint32)
    counts = np.empty(nsub, np.uint32)
    total = C.c_size_t()
    rc = L.gc_zstd_find_sequences_packed(ctx.h, x.ctypes.data, x.size, level, seqs.ctypes.data, cap,
                                         counts.ctypes.data, C.byref(total))
    assert rc == 0, L.gc_last_error()
    r = seqs[: total.value]
    assert not r[:, 3].any(), "rep field set"
    return digest(counts, np.ascontiguousarray(r[:, :3]))


def plain_lzma2(ctx, x: np.ndarray) -> str:
    return digest(ctx.lzma2_packed(x))


def batch_digest(b, i, algo) -> str:
    if algo == gcompress.ALGO_LZMA2:
        return digest(b.lzma2(i))
    seqs, counts = b.zstd(i)
    return slots_zstd(seqs, counts)
