"""Python mirror of the reference's match-finding host interface, over libgcompress.so.

The functions keep the reference's names and argument meaning:

    gpu_find_matches(chunk, device_id) -> GpuMatchData        compress_lzma2_custom.rs:483-559
    compress_chunk_zstd_custom(chunk, device_id, level)        compress_zstd_custom.rs:47-203
        -> (list[bytes], list[int])

All compute goes through the HIP library (libgcompress.so, built in-tree by
gpu-compressor_amd/build.py); there is no CPU fallback: if the library is
missing this module raises on import of the library.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG, "libgcompress.so")

SUB_BLOCK = 65536            # constants.rs:5 / :11
ZSTD_DEPTH_LAZY = 16         # constants.rs:6
LZMA2_MAX_MATCHES = 8        # constants.rs:12
LZMA2_DEPTH = 32             # constants.rs:13
MAX_SEQ = SUB_BLOCK // 3     # compress_zstd_custom.rs:55

SEQ_FORMAT_ZSTD = 0
SEQ_FORMAT_REF = 1

_u8p = C.POINTER(C.c_uint8)
_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)
_vp = C.c_void_p
_sz = C.c_size_t

# (name, restype, argtypes) for every symbol of include/gcompress.h
SIGNATURES = [
    ("gc_open", _vp, [C.c_int]),
    ("gc_close", None, [_vp]),
    ("gc_last_error", C.c_char_p, []),
    ("gc_device_count", C.c_int, []),
    ("gc_dev_alloc", _vp, [_vp, _sz]),
    ("gc_dev_free", C.c_int, [_vp, _vp]),
    ("gc_htod", C.c_int, [_vp, _vp, _vp, _sz]),
    ("gc_dtoh", C.c_int, [_vp, _vp, _vp, _sz]),
    ("gc_dev_zero", C.c_int, [_vp, _vp, _sz]),
    ("gc_sync", C.c_int, [_vp]),
    ("gc_stream", _vp, [_vp]),
    ("gc_zstd_match_find_dev", C.c_int,
     [_vp, _vp, _vp, C.c_uint32, _vp, _vp, _vp, C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint32, C.c_uint32]),
    ("gc_lzma2_match_find_dev", C.c_int,
     [_vp, _vp, _vp, _vp, C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint32, C.c_uint32]),
    ("gc_lzma2_match_find_packed_dev", C.c_int, [_vp, _vp, _vp, C.c_uint32, C.c_uint64, C.c_uint32]),
    ("gc_gpu_find_matches", C.c_int, [_vp, _vp, _sz, _vp, _vp, _vp]),
    ("gc_lzma2_find_matches_packed", C.c_int, [_vp, _vp, _sz, _vp]),
    ("gc_zstd_find_sequences", C.c_int, [_vp, _vp, _sz, C.c_uint32, _vp, _vp]),
]


class GcError(RuntimeError):
    pass


_lib = None


def lib():
    """Load libgcompress.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GcError(f"{LIB_PATH} is missing: run `python gpu-compressor_amd/build.py` first")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc: int, what: str):
    if rc != 0:
        msg = lib().gc_last_error()
        raise GcError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_vp)


def _as_u8(data) -> np.ndarray:
    if isinstance(data, np.ndarray):
        return np.ascontiguousarray(data.view(np.uint8).reshape(-1))
    return np.frombuffer(bytes(data), dtype=np.uint8)


class Context:
    """One device, one stream, the code object loaded once (gc_open)."""

    def __init__(self, device: int = 0):
        self.device = device
        h = lib().gc_open(device)
        if not h:
            raise GcError(f"gc_open({device}) failed: {lib().gc_last_error().decode()}")
        self.h = h

    def close(self):
        if self.h:
            lib().gc_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- device memory ---------------------------------------------------
    def alloc(self, nbytes: int) -> int:
        p = lib().gc_dev_alloc(self.h, nbytes)
        if not p:
            raise GcError(f"gc_dev_alloc({nbytes}): {lib().gc_last_error().decode()}")
        return p

    def free(self, p: int):
        _check(lib().gc_dev_free(self.h, p), "gc_dev_free")

    def htod(self, dst: int, a: np.ndarray):
        _check(lib().gc_htod(self.h, dst, _ptr(a), a.nbytes), "gc_htod")

    def dtoh(self, a: np.ndarray, src: int):
        _check(lib().gc_dtoh(self.h, _ptr(a), src, a.nbytes), "gc_dtoh")

    def zero(self, p: int, nbytes: int):
        _check(lib().gc_dev_zero(self.h, p, nbytes), "gc_dev_zero")

    def sync(self):
        _check(lib().gc_sync(self.h), "gc_sync")

    # -- host-level finders ----------------------------------------------
    def lzma2_packed(self, chunk) -> np.ndarray:
        """Packed LZMA2 matches, one u32 per position: distance << 16 | length."""
        src = _as_u8(chunk)
        out = np.zeros(src.size, dtype=np.uint32)
        _check(lib().gc_lzma2_find_matches_packed(self.h, _ptr(src), src.size, _ptr(out)),
               "gc_lzma2_find_matches_packed")
        return out

    def zstd_sequences(self, chunk, level: int):
        """(seqs[n_sub, MAX_SEQ, 4] u32 as ZSTD_Sequence, counts[n_sub])."""
        src = _as_u8(chunk)
        nsub = -(-src.size // SUB_BLOCK)
        seqs = np.zeros((nsub, MAX_SEQ, 4), dtype=np.uint32)
        counts = np.zeros(nsub, dtype=np.uint32)
        _check(lib().gc_zstd_find_sequences(self.h, _ptr(src), src.size, level, _ptr(seqs), _ptr(counts)),
               "gc_zstd_find_sequences")
        return seqs, counts


_contexts: dict[int, Context] = {}


def context(device_id: int = 0) -> Context:
    """A process-wide context per device (the reference re-creates one per call)."""
    ctx = _contexts.get(device_id)
    if ctx is None:
        ctx = Context(device_id)
        _contexts[device_id] = ctx
    return ctx


@dataclass
class GpuMatchData:
    """compress_lzma2_custom.rs:472-479."""
    match_distances: np.ndarray
    match_lengths: np.ndarray
    match_counts: np.ndarray
    num_sub_blocks: int
    sub_block_size: int
    max_matches: int


def gpu_find_matches(chunk_data, device_id: int = 0) -> GpuMatchData:
    """compress_lzma2_custom.rs:483-559: LZMA2 match finding on the GPU."""
    src = _as_u8(chunk_data)
    nsub = -(-src.size // SUB_BLOCK)
    slots = nsub * SUB_BLOCK
    d = np.empty(slots * LZMA2_MAX_MATCHES, dtype=np.uint32)
    l = np.empty(slots * LZMA2_MAX_MATCHES, dtype=np.uint32)
    c = np.empty(slots, dtype=np.uint32)
    ctx = context(device_id)
    _check(lib().gc_gpu_find_matches(ctx.h, _ptr(src), src.size, _ptr(d), _ptr(l), _ptr(c)), "gc_gpu_find_matches")
    return GpuMatchData(d, l, c, nsub, SUB_BLOCK, LZMA2_MAX_MATCHES)
