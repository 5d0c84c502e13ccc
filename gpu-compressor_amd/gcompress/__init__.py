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
# GCOMPRESS_LIB: load another build of the library (kernel variants side by side)
LIB_PATH = os.environ.get("GCOMPRESS_LIB") or os.path.join(_PKG, "libgcompress.so")

SUB_BLOCK = 65536            # constants.rs:5 / :11
LZMA2_GPU_SLOT = 65552       # gc_lzma2_encode_dev: bytes per sub-block's output slot
LZMA2_GPU_HEAD = 2           # ... and where in the slot the stream starts
ZSTD_DEPTH_LAZY = 16         # constants.rs:6
LZMA2_MAX_MATCHES = 8        # constants.rs:12
LZMA2_DEPTH = 32             # constants.rs:13
MAX_SEQ = SUB_BLOCK // 3     # compress_zstd_custom.rs:55
SERVE_ZSTD_SLOT = 24576      # records per sub-block slot of a served device-result Zstd request

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
    ("gc_set_segmented_parse", C.c_int, [_vp, C.c_uint32]),
    ("gc_zstd_match_find_dev", C.c_int,
     [_vp, _vp, _vp, C.c_uint32, _vp, _vp, _vp, C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint32, C.c_uint32]),
    ("gc_zstd_match_find_dev_noalloc", C.c_int,
     [_vp, _vp, _vp, C.c_uint32, _vp, _vp, _vp, C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint32, C.c_uint32]),
    ("gc_lzma2_match_find_dev", C.c_int,
     [_vp, _vp, _vp, _vp, C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint32, C.c_uint32]),
    ("gc_lzma2_match_find_packed_dev", C.c_int, [_vp, _vp, _vp, C.c_uint32, C.c_uint64, C.c_uint32]),
    ("gc_gpu_find_matches", C.c_int, [_vp, _vp, _sz, _vp, _vp, _vp]),
    ("gc_lzma2_find_matches_packed", C.c_int, [_vp, _vp, _sz, _vp]),
    ("gc_zstd_find_sequences", C.c_int, [_vp, _vp, _sz, C.c_uint32, _vp, _vp]),
    ("gc_zstd_find_sequences_packed", C.c_int, [_vp, _vp, _sz, C.c_uint32, _vp, _sz, _vp, _vp]),
    ("gc_zstd_custom_bound", _sz, [_sz]),
    ("gc_zstd_encode_sequences", C.c_int, [_vp, _sz, _vp, _vp, _vp, _sz, _vp, _vp, C.c_int]),
    ("gc_compress_chunk_zstd_custom", C.c_int, [_vp, _vp, _sz, C.c_uint32, _vp, _sz, _vp, _vp]),
    ("gc_zstd_frame_decode", C.c_int, [_vp, _sz, _vp, _sz, _vp]),
    ("gc_lzma2_raw_encode", C.c_int, [_vp, _sz, C.c_uint32, _vp, _sz, _vp]),
    ("gc_lzma2_raw_decode", C.c_int, [_vp, _sz, _vp, _sz, _vp]),
    ("gc_encode_single_sub_block", C.c_int, [_vp, _sz, _vp, _vp, _sz, _vp]),
    ("gc_lzma2_encode_bound", _sz, [_sz]),
    ("gc_lzma2_encode_matches", C.c_int, [_vp, _sz, _vp, C.c_int, _vp, _sz, _vp]),
    ("gc_cpu_encode_from_matches", C.c_int, [_vp, _sz, _vp, C.c_int, _vp, _sz, _vp, _vp]),
    ("gc_nvlz_compress_matches", C.c_int, [_vp, _vp, _sz, C.c_uint32, C.c_int, _vp, _sz, _vp]),
    ("gc_decompress_file", C.c_int, [C.c_char_p, C.c_char_p, C.c_int, _vp]),
    ("gc_zstd_encode_block_dev", C.c_int, [_vp, _vp, _vp, C.c_uint32, _vp, _vp, _vp, _vp, _vp, C.c_uint32,
                                           C.c_uint32, C.c_uint64, C.c_uint32, C.c_uint32]),
    ("gc_zstd_frames_gpu", C.c_int, [_vp, _vp, _sz, C.c_uint32, _vp, _sz, _vp, _vp]),
    ("gc_nvzs_gpu_bound", _sz, [_sz]),
    ("gc_nvzs_compress_gpu", C.c_int, [_vp, _vp, _sz, C.c_uint32, _vp, _sz, _vp]),
    ("gc_lzma2_encode_dev", C.c_int, [_vp, _vp, _vp, _vp, _vp, C.c_uint64]),
    ("gc_lzma2_streams_gpu", C.c_int, [_vp, _vp, _sz, _vp, _sz, _vp, _vp]),
    ("gc_nvlz_compress_gpu", C.c_int, [_vp, _vp, _sz, C.c_uint32, _vp, _sz, _vp]),
    ("gc_container_bound", _sz, [_sz, C.c_uint32]),
    ("gc_nvzs_compress", C.c_int, [_vp, _vp, _sz, C.c_uint32, C.c_int, _vp, _sz, _vp]),
    ("gc_nvlz_compress", C.c_int, [_vp, _vp, _sz, C.c_uint32, C.c_uint32, C.c_int, _vp, _sz, _vp]),
    ("gc_nvlz_l0_bound", _sz, [_sz, C.c_uint32]),
    ("gc_nvlz_compress_l0", C.c_int, [_vp, _sz, C.c_uint32, C.c_uint32, C.c_int, _vp, _sz, _vp]),
    ("gc_container_info", C.c_int, [_vp, _sz, _vp, _vp, _vp, _vp]),
    ("gc_decompress", C.c_int, [_vp, _sz, _vp, _sz, _vp]),
    ("gc_batch_create", _vp, [_vp, _vp, C.c_uint64, _vp, C.c_uint32]),
    ("gc_batch_destroy", C.c_int, [_vp]),
    ("gc_batch_run", C.c_int, [_vp]),
    ("gc_batch_set_mode", C.c_int, [_vp, C.c_uint32]),
    ("gc_batch_check", C.c_int, [_vp]),
    ("gc_batch_run_timed", C.c_int, [_vp, _vp]),
    ("gc_batch_set_tune", C.c_int, [_vp, _vp]),
    ("gc_batch_get_tune", C.c_int, [_vp, _vp]),
    ("gc_batch_autotune", C.c_int, [_vp, C.c_uint32, _vp, _vp]),
    ("gc_batch_info", C.c_int, [_vp, C.c_uint32, _vp, _vp, _vp]),
    ("gc_batch_device_outputs", C.c_int, [_vp, _vp, _vp, _vp]),
    ("gc_batch_copy_zstd", C.c_int, [_vp, C.c_uint32, _vp, _vp]),
    ("gc_batch_copy_lzma2", C.c_int, [_vp, C.c_uint32, _vp]),
    ("gc_server_start", _vp, [_vp, C.c_uint64, C.c_uint32]),
    ("gc_lzma2_triage", C.c_int, [_vp, _vp, C.c_size_t, _vp]),
    ("gc_lzma2_triage_dev", C.c_int, [_vp, _vp, C.c_uint64, _vp]),
    ("gc_server_submit", C.c_int64, [_vp, _vp, C.c_uint64, C.c_uint32, C.c_uint32]),
    ("gc_server_submit_into", C.c_int64, [_vp, _vp, C.c_uint64, C.c_uint32, C.c_uint32, _vp, C.c_uint64, _vp]),
    ("gc_server_submit_dev", C.c_int64, [_vp, _vp, C.c_uint64, C.c_uint32, C.c_uint32]),
    ("gc_server_submit_dev_many", C.c_int, [_vp, _vp, _vp, _vp, _vp, C.c_uint32, _vp]),
    ("gc_server_wait_release_many", C.c_int, [_vp, _vp, C.c_uint32, _vp, _vp]),
    ("gc_server_wait", C.c_int, [_vp, C.c_uint64, C.c_int, _vp]),
    ("gc_server_result_dev", C.c_int, [_vp, C.c_uint64, _vp, _vp]),
    ("gc_server_result_zstd", C.c_int, [_vp, C.c_uint64, _vp, C.c_uint64, _vp]),
    ("gc_server_result_lzma2", C.c_int, [_vp, C.c_uint64, _vp]),
    ("gc_server_release", C.c_int, [_vp, C.c_uint64]),
    ("gc_server_get_stats", C.c_int, [_vp, _vp]),
    ("gc_server_stop", C.c_int, [_vp]),
    ("gc_corpus_generate", C.c_int, [_vp, C.c_uint64, C.c_uint64, _vp, C.c_uint64, C.c_uint32, C.c_uint64,
                                     C.c_uint64]),
]

ALGO_ZSTD = 0
ALGO_LZMA2 = 1
BATCH_STATIC = 0
BATCH_QUEUE_GRAPH = 1
BATCH_PERSISTENT = 2
SEG_DEFAULT_MAX_SUB = 2048  # GC_SEG_DEFAULT_MAX_SUB


class BatchInput(C.Structure):
    _fields_ = [("offset", C.c_uint64), ("size", C.c_uint64), ("algo", C.c_uint32), ("level", C.c_uint32)]


class Tune(C.Structure):
    _fields_ = [("cand_wg_per_cu", C.c_uint32), ("parse_wg_per_cu", C.c_uint32),
                ("elem_wg_per_cu", C.c_uint32), ("info_strip", C.c_uint32)]


class RequestInfo(C.Structure):
    _fields_ = [("size", C.c_uint64), ("algo", C.c_uint32), ("n_sub", C.c_uint32), ("n_sequences", C.c_uint64),
                ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_start", C.c_double), ("t_done", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("batches", C.c_uint64), ("items", C.c_uint64), ("completed", C.c_uint64), ("bytes", C.c_uint64),
                ("hbm_in_use", C.c_uint64), ("pending", C.c_uint64), ("t_batches", C.c_double),
                ("t_fetch", C.c_double), ("t_host", C.c_double), ("t_kernel", C.c_double * 6),
                ("launches", C.c_uint64 * 6), ("zstd_bytes", C.c_uint64), ("zstd_sub_blocks", C.c_uint64),
                ("zstd_sequences", C.c_uint64), ("lzma2_bytes", C.c_uint64)]


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

    def lzma2_triage(self, data) -> np.ndarray:
        """One byte per 64 KiB sub-block: 1 if any position has an LZMA2 match (gc_lzma2_triage)."""
        d = _as_u8(data)
        out = np.zeros(max(-(-d.size // 65536), 1), np.uint8)
        _check(lib().gc_lzma2_triage(self.h, _ptr(d), d.size, _ptr(out)), "gc_lzma2_triage")
        return out[: -(-d.size // 65536)]

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

    def set_segmented_parse(self, max_sub_blocks: int):
        """Calls of at most max_sub_blocks sub-chunks use the segmented parse (0: never)."""
        _check(lib().gc_set_segmented_parse(self.h, max_sub_blocks), "gc_set_segmented_parse")

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


def corpus(n: int, seed: int, weights=(1, 1, 1, 1), segment: int = 65536, exact_shares: bool = True,
           gap_every: int = 0, gap_len: int = 0) -> np.ndarray:
    """Deterministic four-kind synthetic corpus (gc_corpus_generate)."""
    out = np.empty(n, dtype=np.uint8)
    w = np.asarray(weights, dtype=np.float64)
    _check(lib().gc_corpus_generate(_ptr(out), n, seed, _ptr(w), segment, int(exact_shares), gap_every, gap_len),
           "gc_corpus_generate")
    return out


class Batch:
    """Many inputs in one pass, one launch per kernel (gc_batch_*).

    `inputs` is a list of (offset, size, algo, level) over a device arena."""

    def __init__(self, ctx: Context, d_arena: int, arena_size: int, inputs):
        self.ctx = ctx
        arr = (BatchInput * len(inputs))(*[BatchInput(*t) for t in inputs])
        self.inputs = list(inputs)
        h = lib().gc_batch_create(ctx.h, d_arena, arena_size, C.cast(arr, _vp), len(inputs))
        if not h:
            raise GcError(f"gc_batch_create: {lib().gc_last_error().decode()}")
        self.h = h

    def close(self):
        if self.h:
            lib().gc_batch_destroy(self.h)
            self.h = None

    def run(self):
        _check(lib().gc_batch_run(self.h), "gc_batch_run")

    def set_mode(self, mode):
        _check(lib().gc_batch_set_mode(self.h, mode), "gc_batch_set_mode")

    def check(self):
        """Waits for the queued passes; raises if a persistent pass's work queue reported an error."""
        _check(lib().gc_batch_check(self.h), "gc_batch_check")

    def run_timed(self):
        """Per-kernel ms: [candidates, LZMA2 lengths, Zstd verify, Zstd parse]."""
        ms = (C.c_float * 4)()
        _check(lib().gc_batch_run_timed(self.h, C.cast(ms, _vp)), "gc_batch_run_timed")
        return list(ms)

    def set_tune(self, cand, parse, elem, info_strip=8):
        t = Tune(cand, parse, elem, info_strip)
        _check(lib().gc_batch_set_tune(self.h, C.byref(t)), "gc_batch_set_tune")

    def tune(self):
        t = Tune()
        _check(lib().gc_batch_get_tune(self.h, C.byref(t)), "gc_batch_get_tune")
        return (t.cand_wg_per_cu, t.parse_wg_per_cu, t.elem_wg_per_cu, t.info_strip)

    def autotune(self, reps=2):
        t = Tune()
        ms = C.c_float()
        _check(lib().gc_batch_autotune(self.h, reps, C.byref(t), C.byref(ms)), "gc_batch_autotune")
        return (t.cand_wg_per_cu, t.parse_wg_per_cu, t.elem_wg_per_cu, t.info_strip), ms.value

    def n_sub(self, i):
        ns = C.c_uint32()
        _check(lib().gc_batch_info(self.h, i, C.byref(ns), None, None), "gc_batch_info")
        return ns.value

    def zstd(self, i, with_seqs=True):
        ns = self.n_sub(i)
        counts = np.zeros(ns, np.uint32)
        seqs = np.zeros((ns, MAX_SEQ, 4), np.uint32) if with_seqs else None
        _check(lib().gc_batch_copy_zstd(self.h, i, _ptr(seqs) if with_seqs else None, _ptr(counts)),
               "gc_batch_copy_zstd")
        return seqs, counts

    def lzma2(self, i):
        out = np.zeros(self.inputs[i][1], np.uint32)
        _check(lib().gc_batch_copy_lzma2(self.h, i, _ptr(out)), "gc_batch_copy_lzma2")
        return out


# ---- consumers and containers ------------------------------------------------

def compress_chunk_zstd_custom(chunk_data, device_id: int = 0, level: int = 1):
    """compress_zstd_custom.rs:47-203 -> (list of frames, list of sizes)."""
    src = _as_u8(chunk_data)
    nsub = -(-src.size // SUB_BLOCK)
    cap = lib().gc_zstd_custom_bound(src.size)
    out = np.empty(max(cap, 1), np.uint8)
    sizes = np.zeros(max(nsub, 1), np.uint64)
    n = C.c_size_t()
    _check(lib().gc_compress_chunk_zstd_custom(context(device_id).h, _ptr(src), src.size, level, _ptr(out), cap,
                                               _ptr(sizes), C.byref(n)), "gc_compress_chunk_zstd_custom")
    frames, o = [], 0
    for s in sizes[:nsub]:
        frames.append(out[o:o + int(s)].tobytes())
        o += int(s)
    return frames, [int(s) for s in sizes[:nsub]]


def encode_single_sub_block(sub_data, match_counts) -> bytes:
    """compress_lzma2_custom.rs:570-597."""
    src = _as_u8(sub_data)
    mc = np.ascontiguousarray(match_counts, dtype=np.uint32)
    cap = src.size + src.size // 8 + 1024
    out = np.empty(cap, np.uint8)
    n = C.c_size_t()
    _check(lib().gc_encode_single_sub_block(_ptr(src), src.size, _ptr(mc), _ptr(out), cap, C.byref(n)),
           "gc_encode_single_sub_block")
    return out[: n.value].tobytes()


def lzma2_encode_matches(sub_data, packed, verify: bool = True) -> bytes:
    """One sub-block (<= 65536 bytes) as an LZMA2 stream coded from its packed matches
    (compress_lzma2_custom.rs:344-465, with the LZMA format's full coding)."""
    src = _as_u8(sub_data)
    pk = np.ascontiguousarray(packed, dtype=np.uint32)
    if pk.size < src.size:
        raise ValueError("one packed match word per byte")
    cap = lib().gc_lzma2_encode_bound(src.size)
    out = np.empty(cap, np.uint8)
    n = C.c_size_t()
    _check(lib().gc_lzma2_encode_matches(_ptr(src), src.size, _ptr(pk), int(verify), _ptr(out), cap, C.byref(n)),
           "gc_lzma2_encode_matches")
    return out[: n.value].tobytes()


def cpu_encode_from_matches(chunk_data, packed, threads: int = 0):
    """compress_lzma2_custom.rs:601-664: one verified LZMA2 stream per 64 KiB sub-block -> (streams, sizes)."""
    src = _as_u8(chunk_data)
    pk = np.ascontiguousarray(packed, dtype=np.uint32)
    if pk.size < src.size:
        raise ValueError("one packed match word per byte")
    nsub = -(-src.size // SUB_BLOCK)
    cap = max(nsub, 1) * lib().gc_lzma2_encode_bound(SUB_BLOCK)
    out = np.empty(cap, np.uint8)
    sizes = np.zeros(max(nsub, 1), np.uint64)
    n = C.c_size_t()
    _check(lib().gc_cpu_encode_from_matches(_ptr(src), src.size, _ptr(pk), threads, _ptr(out), cap, _ptr(sizes),
                                            C.byref(n)), "gc_cpu_encode_from_matches")
    streams, o = [], 0
    for z in sizes[:nsub]:
        streams.append(out[o:o + int(z)].tobytes())
        o += int(z)
    return streams, [int(z) for z in sizes[:nsub]]


def nvlz_compress_matches(data, dict_mb: int = 128, device_id: int = 0, threads: int = 0) -> bytes:
    """NVLZ container with sub-blocks coded from the GPU's matches (no second search)."""
    src = _as_u8(data)
    cap = lib().gc_container_bound(src.size, ALGO_LZMA2)
    out = np.empty(cap, np.uint8)
    n = C.c_size_t()
    _check(lib().gc_nvlz_compress_matches(context(device_id).h, _ptr(src), src.size, dict_mb, threads, _ptr(out),
                                          cap, C.byref(n)), "gc_nvlz_compress_matches")
    return out[: n.value].tobytes()


def zstd_frames_gpu(data, level: int = 1, device_id: int = 0):
    """Match finding and Zstd frame coding on the GPU -> (frames, sizes), one frame per 64 KiB."""
    src = _as_u8(data)
    nsub = -(-src.size // SUB_BLOCK)
    cap = src.size + 16 * max(nsub, 1)
    out = np.empty(max(cap, 1), np.uint8)
    sizes = np.zeros(max(nsub, 1), np.uint64)
    n = C.c_size_t()
    _check(lib().gc_zstd_frames_gpu(context(device_id).h, _ptr(src), src.size, level, _ptr(out), cap, _ptr(sizes),
                                    C.byref(n)), "gc_zstd_frames_gpu")
    frames, o = [], 0
    for z in sizes[:nsub]:
        frames.append(out[o:o + int(z)].tobytes())
        o += int(z)
    return frames, [int(z) for z in sizes[:nsub]]


def _out_buffer(out, cap, name):
    if out.dtype != np.uint8 or not out.flags.c_contiguous or out.size < cap:
        raise ValueError(f"{name}: out must be a contiguous uint8 array of at least {cap} bytes")
    return out


def nvzs_compress_gpu(data, level: int = 1, device_id: int = 0, out=None):
    """NVZS container whose frames are coded on the GPU -> bytes; with `out` (a uint8 array of at least
    gc_nvzs_gpu_bound bytes, kept by the caller across calls) the container is written there and its
    length returned, with no copy of it in Python."""
    src = _as_u8(data)
    cap = lib().gc_nvzs_gpu_bound(src.size)
    buf = np.empty(cap, np.uint8) if out is None else _out_buffer(out, cap, "nvzs_compress_gpu")
    n = C.c_size_t()
    _check(lib().gc_nvzs_compress_gpu(context(device_id).h, _ptr(src), src.size, level, _ptr(buf), buf.size,
                                      C.byref(n)), "gc_nvzs_compress_gpu")
    return n.value if out is not None else buf[: n.value].tobytes()


def lzma2_streams_gpu(data, device_id: int = 0):
    """Match finding and LZMA2 coding on the GPU -> (streams, sizes), one stream per 64 KiB."""
    src = _as_u8(data)
    nsub = -(-src.size // SUB_BLOCK)
    cap = src.size + 4 * max(nsub, 1)
    out = np.empty(max(cap, 1), np.uint8)
    sizes = np.zeros(max(nsub, 1), np.uint64)
    n = C.c_size_t()
    _check(lib().gc_lzma2_streams_gpu(context(device_id).h, _ptr(src), src.size, _ptr(out), cap, _ptr(sizes),
                                      C.byref(n)), "gc_lzma2_streams_gpu")
    streams, o = [], 0
    for z in sizes[:nsub]:
        streams.append(out[o:o + int(z)].tobytes())
        o += int(z)
    return streams, [int(z) for z in sizes[:nsub]]


def nvlz_compress_gpu(data, dict_mb: int = 128, device_id: int = 0, out=None):
    """NVLZ container whose sub-blocks are coded on the GPU from the GPU's matches -> bytes; with `out` (a
    uint8 array of at least gc_container_bound bytes, kept by the caller across calls) the container is
    written there and its length returned, with no copy of it in Python."""
    src = _as_u8(data)
    cap = lib().gc_container_bound(src.size, ALGO_LZMA2)
    buf = np.empty(cap, np.uint8) if out is None else _out_buffer(out, cap, "nvlz_compress_gpu")
    n = C.c_size_t()
    _check(lib().gc_nvlz_compress_gpu(context(device_id).h, _ptr(src), src.size, dict_mb, _ptr(buf), buf.size,
                                      C.byref(n)), "gc_nvlz_compress_gpu")
    return n.value if out is not None else buf[: n.value].tobytes()


def nvzs_compress(data, level: int = 1, device_id: int = 0, threads: int = 1) -> bytes:
    src = _as_u8(data)
    cap = lib().gc_container_bound(src.size, ALGO_ZSTD)
    out = np.empty(cap, np.uint8)
    n = C.c_size_t()
    _check(lib().gc_nvzs_compress(context(device_id).h, _ptr(src), src.size, level, threads, _ptr(out), cap,
                                  C.byref(n)), "gc_nvzs_compress")
    return out[: n.value].tobytes()


def nvlz_compress(data, level: int = 1, dict_mb: int = 128, device_id: int = 0, threads: int = 0) -> bytes:
    src = _as_u8(data)
    cap = lib().gc_container_bound(src.size, ALGO_LZMA2)
    out = np.empty(cap, np.uint8)
    n = C.c_size_t()
    _check(lib().gc_nvlz_compress(context(device_id).h, _ptr(src), src.size, level, dict_mb, threads, _ptr(out),
                                  cap, C.byref(n)), "gc_nvlz_compress")
    return out[: n.value].tobytes()


def nvlz_compress_l0(data, dict_mb: int = 128, preset: int = 0, threads: int = 0) -> bytes:
    src = _as_u8(data)
    cap = lib().gc_nvlz_l0_bound(src.size, dict_mb)
    out = np.empty(cap, np.uint8)
    n = C.c_size_t()
    _check(lib().gc_nvlz_compress_l0(_ptr(src), src.size, dict_mb, preset, threads, _ptr(out), cap, C.byref(n)),
           "gc_nvlz_compress_l0")
    return out[: n.value].tobytes()


def decompress_file(in_path, out_path, threads: int = 0) -> int:
    """decompress.rs:9-143, pipelined: container file -> original bytes in out_path; returns bytes written."""
    n = C.c_uint64()
    _check(lib().gc_decompress_file(os.fsencode(in_path), os.fsencode(out_path), threads, C.byref(n)),
           "gc_decompress_file")
    return n.value


def container_info(blob: bytes):
    b = _as_u8(blob)
    algo, orig, ch, nch = C.c_uint32(), C.c_uint64(), C.c_uint64(), C.c_uint64()
    _check(lib().gc_container_info(_ptr(b), b.size, C.byref(algo), C.byref(orig), C.byref(ch), C.byref(nch)),
           "gc_container_info")
    return algo.value, orig.value, ch.value, nch.value


def decompress(blob: bytes) -> bytes:
    """decompress.rs:9-379 (NVZS via libzstd, NVLZ via liblzma)."""
    b = _as_u8(blob)
    _, orig, _, _ = container_info(b)
    out = np.empty(max(orig, 1), np.uint8)
    n = C.c_size_t()
    _check(lib().gc_decompress(_ptr(b), b.size, _ptr(out), orig, C.byref(n)), "gc_decompress")
    return out[: n.value].tobytes()


def zstd_encode_sequences(src, seqs: np.ndarray, counts: np.ndarray, threads: int = 1):
    """CPU consumer alone: packed ZSTD_Sequence rows + per-sub-chunk counts -> (frames, sizes)."""
    s = _as_u8(src)
    seqs = np.ascontiguousarray(seqs, dtype=np.uint32).reshape(-1, 4)
    counts = np.ascontiguousarray(counts, dtype=np.uint32)
    cap = lib().gc_zstd_custom_bound(s.size)
    out = np.empty(max(cap, 1), np.uint8)
    nsub = -(-s.size // SUB_BLOCK)
    sizes = np.zeros(max(nsub, 1), np.uint64)
    n = C.c_size_t()
    _check(lib().gc_zstd_encode_sequences(_ptr(s), s.size, _ptr(seqs), _ptr(counts), _ptr(out), cap, _ptr(sizes),
                                          C.byref(n), threads), "gc_zstd_encode_sequences")
    frames, o = [], 0
    for z in sizes[:nsub]:
        frames.append(out[o:o + int(z)].tobytes())
        o += int(z)
    return frames, [int(z) for z in sizes[:nsub]]


def zstd_frame_decode(frame: bytes, expected: int) -> bytes:
    f = _as_u8(frame)
    out = np.empty(max(expected, 1), np.uint8)
    n = C.c_size_t()
    _check(lib().gc_zstd_frame_decode(_ptr(f), f.size, _ptr(out), expected, C.byref(n)), "gc_zstd_frame_decode")
    return out[: n.value].tobytes()


def lzma2_raw_decode(blob: bytes, expected: int) -> bytes:
    f = _as_u8(blob)
    out = np.empty(max(expected, 1), np.uint8)
    n = C.c_size_t()
    _check(lib().gc_lzma2_raw_decode(_ptr(f), f.size, _ptr(out), expected, C.byref(n)), "gc_lzma2_raw_decode")
    return out[: n.value].tobytes()


class Server:
    """Served form (gc_server_*): continuous batching of match-finding requests."""

    def __init__(self, ctx: Context, hbm_budget: int, max_batch_items: int = 0):
        self.ctx = ctx
        h = lib().gc_server_start(ctx.h, hbm_budget, max_batch_items)
        if not h:
            raise GcError(f"gc_server_start: {lib().gc_last_error().decode()}")
        self.h = h
        self._keep = {}

    def submit(self, data: np.ndarray, algo: int, level: int = 1) -> int:
        """Submit a request; `data` is kept referenced until release()."""
        rid = lib().gc_server_submit(self.h, _ptr(data), data.size, algo, level)
        if rid < 0:
            _check(int(rid), "gc_server_submit")
        self._keep[rid] = data
        return rid

    def submit_into(self, data: np.ndarray, algo: int, level: int, out: np.ndarray, counts=None) -> int:
        """Submit with the results written straight into `out` (LZMA2: u32 per byte; Zstd: (k, 4) u32
        sequence records) and, for Zstd, `counts` (u32 per sub-block)."""
        cp = _ptr(counts) if counts is not None else None
        rid = lib().gc_server_submit_into(self.h, _ptr(data), data.size, algo, level, _ptr(out), out.nbytes, cp)
        if rid < 0:
            _check(int(rid), "gc_server_submit_into")
        self._keep[rid] = (data, out, counts)
        return rid

    def submit_dev(self, d_input: int, n: int, algo: int, level: int = 1) -> int:
        """Submit a request whose input is already in device memory; its results stay there."""
        rid = lib().gc_server_submit_dev(self.h, d_input, n, algo, level)
        if rid < 0:
            _check(int(rid), "gc_server_submit_dev")
        return rid

    @staticmethod
    def dev_batch(reqs):
        """[(device pointer, n, algo, level)] as the arrays submit_dev_many takes."""
        return (np.array([r[0] for r in reqs], np.uint64), np.array([r[1] for r in reqs], np.uint64),
                np.array([r[2] for r in reqs], np.uint32), np.array([r[3] for r in reqs], np.uint32))

    def submit_dev_many(self, batch) -> np.ndarray:
        """Submit a dev_batch at once (one lock, in order); returns their ids."""
        ptrs, sizes, algos, levels = batch
        k = ptrs.size
        ids = np.zeros(max(k, 1), np.int64)
        _check(lib().gc_server_submit_dev_many(self.h, _ptr(ptrs), _ptr(sizes), _ptr(algos), _ptr(levels), k,
                                               _ptr(ids)), "gc_server_submit_dev_many")
        return ids[:k]

    def wait_release_many(self, ids):
        """Wait for each id in order and release it; returns (infos, status codes)."""
        arr = np.ascontiguousarray(ids, np.uint64)
        k = arr.size
        infos = (RequestInfo * max(k, 1))()
        status = np.zeros(max(k, 1), np.int32)
        lib().gc_server_wait_release_many(self.h, _ptr(arr), k, infos, _ptr(status))
        for rid in arr:
            self._keep.pop(int(rid), None)
        return infos, status[:k]

    def result_dev(self, rid: int, n_sub: int):
        """(device pointer of the results, Zstd sequence counts as an array or None)."""
        p = C.c_void_p()
        cp = C.POINTER(C.c_uint32)()
        _check(lib().gc_server_result_dev(self.h, rid, C.byref(p), C.byref(cp)), "gc_server_result_dev")
        counts = np.ctypeslib.as_array(cp, (n_sub,)).copy() if cp else None
        return p.value, counts

    def wait(self, rid: int, timeout_ms: int = -1):
        info = RequestInfo()
        rc = lib().gc_server_wait(self.h, rid, timeout_ms, C.byref(info))
        if rc == 1:
            return None
        _check(rc, "gc_server_wait")
        return info

    def result_zstd(self, rid: int, n_sub: int, n_seq: int):
        seqs = np.zeros((max(n_seq, 1), 4), np.uint32)
        counts = np.zeros(max(n_sub, 1), np.uint32)
        _check(lib().gc_server_result_zstd(self.h, rid, _ptr(seqs), seqs.shape[0], _ptr(counts)),
               "gc_server_result_zstd")
        return seqs[:n_seq], counts[:n_sub]

    def result_lzma2(self, rid: int, n: int):
        out = np.zeros(max(n, 1), np.uint32)
        _check(lib().gc_server_result_lzma2(self.h, rid, _ptr(out)), "gc_server_result_lzma2")
        return out[:n]

    def release(self, rid: int):
        _check(lib().gc_server_release(self.h, rid), "gc_server_release")
        self._keep.pop(rid, None)

    def stats(self):
        st = ServerStats()
        _check(lib().gc_server_get_stats(self.h, C.byref(st)), "gc_server_get_stats")
        out = {}
        for k, _ in ServerStats._fields_:
            v = getattr(st, k)
            out[k] = list(v) if isinstance(v, C.Array) else v
        return out

    def stop(self):
        if self.h:
            lib().gc_server_stop(self.h)
            self.h = None
