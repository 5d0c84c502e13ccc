"""ctypes bindings of the CPU oracle (oracle/gc_oracle.c -> oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product library never
uses it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
SUB_BLOCK = 65536
MAX_SEQ = SUB_BLOCK // 3

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `python gpu-compressor_amd/build.py`")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.orc_hash4.restype = C.c_uint32
        L.orc_hash4.argtypes = [vp]
        L.orc_zstd_match_find.restype = C.c_int
        L.orc_zstd_match_find.argtypes = [vp, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, vp, vp, vp, vp, C.c_int]
        L.orc_zstd_frame.restype = C.c_size_t
        L.orc_zstd_frame.argtypes = [vp, C.c_uint32, vp, C.c_uint32, vp, C.c_size_t]
        L.orc_lzma2_match_find.restype = C.c_int
        L.orc_lzma2_match_find.argtypes = [vp, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, vp, vp, C.c_int]
        _lib = L
    return _lib


def _u8(data) -> np.ndarray:
    if isinstance(data, np.ndarray):
        return np.ascontiguousarray(data.view(np.uint8).reshape(-1))
    return np.frombuffer(bytes(data), dtype=np.uint8)


def hash4(four: bytes) -> int:
    b = np.frombuffer(bytes(four[:4]), dtype=np.uint8).copy()
    return lib().orc_hash4(b.ctypes.data_as(C.c_void_p))


def zstd_match_find(data, chunk: int = SUB_BLOCK, depth: int = 16, max_seq: int | None = None,
                    literals: bool = False, threads: int = 1):
    """Reference-layout Zstd result: (seq[n_sub, max_seq, 3], seq_counts, lit_counts, literals|None)."""
    src = _u8(data)
    if max_seq is None:
        max_seq = chunk // 3
    nsub = -(-src.size // chunk)
    seq = np.zeros((nsub, max_seq, 3), dtype=np.uint32)
    sc = np.zeros(nsub, dtype=np.uint32)
    lc = np.zeros(nsub, dtype=np.uint32)
    lit = np.zeros(nsub * chunk, dtype=np.uint8) if literals else None
    vp = C.c_void_p
    rc = lib().orc_zstd_match_find(src.ctypes.data_as(vp), src.size, chunk, depth, max_seq,
                                   seq.ctypes.data_as(vp), lit.ctypes.data_as(vp) if lit is not None else None,
                                   sc.ctypes.data_as(vp), lc.ctypes.data_as(vp), threads)
    if rc != 0:
        raise ValueError("orc_zstd_match_find rejected its arguments")
    return seq, sc, lc, lit


def lzma2_match_find(data, sub: int = SUB_BLOCK, max_matches: int = 8, depth: int = 32, threads: int = 1):
    """Reference-layout LZMA2 result: (matches[n_sub*sub, max_matches, 2], counts[n_sub*sub])."""
    src = _u8(data)
    nsub = -(-src.size // sub)
    m = np.zeros((nsub * sub, max_matches, 2), dtype=np.uint32)
    c = np.zeros(nsub * sub, dtype=np.uint32)
    vp = C.c_void_p
    rc = lib().orc_lzma2_match_find(src.ctypes.data_as(vp), src.size, sub, max_matches, depth,
                                    m.ctypes.data_as(vp), c.ctypes.data_as(vp), threads)
    if rc != 0:
        raise ValueError("orc_lzma2_match_find rejected its arguments")
    return m, c


def lzma2_packed(data, sub: int = SUB_BLOCK, depth: int = 32, threads: int = 1) -> np.ndarray:
    """Oracle result in the compact form: distance << 16 | length per position (0 = none)."""
    src = _u8(data)
    m, c = lzma2_match_find(src, sub, 1, depth, threads)
    packed = (m[:, 0, 0].astype(np.uint32) << 16) | m[:, 0, 1]
    packed[c == 0] = 0
    return packed[: src.size].copy()


def zstd_as_zstd_sequences(seq: np.ndarray, counts: np.ndarray):
    """Reference records -> list (per sub-block) of ZSTD_Sequence rows [offset, litLength, matchLength, rep]."""
    out = []
    for s in range(seq.shape[0]):
        r = seq[s, : counts[s]]
        z = np.zeros((r.shape[0], 4), dtype=np.uint32)
        z[:, 0] = r[:, 2]
        z[:, 1] = r[:, 0]
        z[:, 2] = r[:, 1] + 3
        out.append(z)
    return out


def zstd_frame(sub, rows) -> bytes:
    """One Zstd frame for a sub-block (<= 65536 bytes) from its ZSTD_Sequence rows {offset, litLength,
    matchLength, rep}: per dimension the cheapest of predefined, RLE and FSE_Compressed tables, repeat
    offset codes, Huffman or RLE literals when smaller than raw -- the checker of the GPU frame coder."""
    src = _u8(sub)
    r = np.ascontiguousarray(rows, dtype=np.uint32).reshape(-1, 4)
    cap = src.size + 64
    out = np.empty(cap, np.uint8)
    n = lib().orc_zstd_frame(src.ctypes.data, src.size, r.ctypes.data if r.size else None, r.shape[0],
                             out.ctypes.data, cap)
    if n == 0:
        raise RuntimeError("frame did not fit")
    return out[:n].tobytes()
