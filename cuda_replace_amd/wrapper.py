"""ctypes binding of libcuda_replace.so with the reference wrapper's API.

Same classes, methods, argument meaning and errors as the reference's
``cuda_replace_wrapper.py`` (``CudaReplaceLib`` py:35-141, ``Session`` py:145-297,
``replace_unified`` py:299-301, ``gpu_replace_streaming`` py:303-366): every
method raises ``RuntimeError("... rc=N")`` on a non-zero return code.

Differences, all towards ``bytes.replace`` or speed:

* inputs are passed to the library without an extra host copy;
* ``Session.length()`` asks the library for the length instead of copying the
  whole buffer back (``cr_session_length``);
* ``gpu_replace_streaming`` moves a chunk's cut to the end of a kept match that
  straddles the chunk end (the reference cuts at the chunk end and loses bytes,
  py:355-364) and handles the empty pattern like ``bytes.replace``;
* 64-bit offsets everywhere, so the 2^30 expansion guard of py:327-340 is not
  needed.

There is no CPU fallback: if the library is missing, ``CudaReplaceLib`` raises.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Iterable, List, Optional, Sequence, Tuple

_HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_LIB = os.path.join(_HERE, "libcuda_replace.so")

c_size_t = C.c_size_t
c_uint8_p = C.POINTER(C.c_uint8)
c_void_p = C.c_void_p
c_int_p = C.POINTER(C.c_int)
c_longlong_p = C.POINTER(C.c_longlong)
c_size_t_p = C.POINTER(c_size_t)
c_uint8_pp = C.POINTER(c_uint8_p)

# symbols the library must export (the reference ABI plus extensions)
REQUIRED_SYMBOLS = (
    "cuda_replace_unified", "cuda_free_host", "cuda_replace_open", "cuda_replace_reset",
    "cuda_replace_build_index", "cuda_replace_apply", "cuda_replace_apply_seeded",
    "cuda_replace_query_prefix", "cuda_replace_apply_batch", "cuda_replace_result",
    "cuda_replace_close", "cuda_replace_version",
    "cr_session_length", "cr_query_prefix64", "cr_replace_device", "cr_pattern_mode",
    "cr_batch_create", "cr_batch_run", "cr_batch_destroy", "cr_stream_replace",
    "cr_server_create", "cr_server_submit", "cr_server_wait", "cr_server_release", "cr_server_destroy",
    "cr_apply_batch_ex", "cr_batch_fusable", "cr_autotune_set", "cr_autotune_reset", "cr_autotune_report",
    "cr_batch_set_mode", "cr_batch_info", "cr_replace_device_noalloc",
)


_bytes_from = C.pythonapi.PyBytes_FromStringAndSize
_bytes_from.restype = C.py_object
_bytes_from.argtypes = [c_void_p, C.c_ssize_t]


def _bytes_at(ptr, n: int) -> bytes:
    """Copies n bytes at ptr into a new bytes object; any size (ctypes.string_at
    takes a C int length and cuts results of 2 GiB and more)."""
    if not n:
        return b""
    addr = ptr if isinstance(ptr, int) else C.cast(ptr, c_void_p).value
    return _bytes_from(addr, n)


def _u8(b: bytes):
    """Pointer to the bytes object's own buffer (no copy; NULs allowed)."""
    return C.cast(C.c_char_p(b), c_uint8_p)


class _Arr:
    """A ctypes array with numpy's ``.ctypes.data_as`` spelling, so both forms of
    _blob_arrays' result are used the same way."""

    def __init__(self, arr):
        self.arr = arr
        self.ctypes = self

    def data_as(self, ptr_type):
        return C.cast(self.arr, ptr_type)


def _blob_arrays(items):
    """(blob, pointer array, length array) for a list of bytes: the items joined
    into one buffer (kept alive by the caller) and the address of each item in it.
    numpy builds the arrays when it is installed; without it (the reference wrapper
    needs only the standard library, py:1-6) ctypes arrays are built in Python."""
    k = len(items)
    blob = b"".join(items) + b"\0"
    base = C.cast(C.c_char_p(blob), c_void_p).value
    try:
        import numpy as np
    except ImportError:
        np = None
    if np is None:
        from itertools import accumulate
        lens = list(map(len, items))
        offs = [0] + list(accumulate(lens[:-1])) if k else [0]
        ptrs = (C.c_uint64 * max(k, 1))(*(base + o for o in offs[:max(k, 1)]))
        return blob, _Arr(ptrs), _Arr((C.c_int * max(k, 1))(*lens))
    lens = np.fromiter(map(len, items), dtype=np.int32, count=k)
    offs = np.zeros(max(k, 1), dtype=np.uint64)
    if k > 1:
        np.cumsum(lens[:-1], dtype=np.uint64, out=offs[1:k])
    ptrs = offs + np.uint64(base)
    if len(lens) == 0:
        lens = np.zeros(1, dtype=np.int32)
    return blob, ptrs, lens


def _pair_arrays(pairs):
    """(pats, pat_lens, reps, rep_lens, keepalive) C arrays for a pair list: the
    patterns and the replacements each joined into one buffer, with arrays of their
    addresses and lengths (a ctypes object per pair costs ~3 us each)."""
    pblob, pptrs, plens = _blob_arrays([bytes(p) for p, _ in pairs])
    rblob, rptrs, rlens = _blob_arrays([bytes(r) for _, r in pairs])
    return (pptrs.ctypes.data_as(c_uint8_pp), plens.ctypes.data_as(c_int_p), rptrs.ctypes.data_as(c_uint8_pp),
            rlens.ctypes.data_as(c_int_p), (pblob, pptrs, plens, rblob, rptrs, rlens))


def _load(path: str) -> C.CDLL:
    path = os.fspath(path)
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"libcuda_replace.so not found at {path}; build it with "
            f"`python -m cuda_replace_amd.build`")
    return C.CDLL(path)


class CudaReplaceLib:
    """Loads libcuda_replace.so and binds its signatures (py:35-121)."""

    def __init__(self, dll_path: Optional[str] = None):
        self.path = os.fspath(dll_path) if dll_path else DEFAULT_LIB
        self.lib = _load(self.path)
        lib = self.lib
        lib.cuda_replace_unified.argtypes = [c_uint8_p, c_size_t, c_uint8_p, C.c_int, c_uint8_p,
                                             C.c_int, C.POINTER(c_uint8_p), c_size_t_p]
        lib.cuda_replace_unified.restype = C.c_int
        lib.cuda_free_host.argtypes = [c_void_p]
        lib.cuda_free_host.restype = None
        lib.cuda_replace_open.argtypes = [C.POINTER(c_void_p), c_uint8_p, c_size_t]
        lib.cuda_replace_open.restype = C.c_int
        lib.cuda_replace_build_index.argtypes = [c_void_p, c_size_t]
        lib.cuda_replace_build_index.restype = C.c_int
        lib.cuda_replace_apply.argtypes = [c_void_p, c_uint8_p, C.c_int, c_uint8_p, C.c_int, c_size_t_p]
        lib.cuda_replace_apply.restype = C.c_int
        lib.cuda_replace_apply_batch.argtypes = [c_void_p, c_uint8_pp, c_int_p, c_uint8_pp, c_int_p,
                                                 C.c_int, c_size_t_p]
        lib.cuda_replace_apply_batch.restype = C.c_int
        lib.cr_apply_batch_ex.argtypes = [c_void_p, c_uint8_pp, c_int_p, c_uint8_pp, c_int_p, C.c_int, c_size_t_p,
                                          C.c_int, c_int_p]
        lib.cr_apply_batch_ex.restype = C.c_int
        lib.cr_batch_fusable.argtypes = [c_uint8_pp, c_int_p, c_uint8_pp, c_int_p, C.c_int]
        lib.cr_batch_fusable.restype = C.c_int
        lib.cr_autotune_set.argtypes = [c_uint8_p, C.c_int, c_size_t, C.c_int]
        lib.cr_autotune_set.restype = C.c_int
        lib.cr_autotune_reset.argtypes = []
        lib.cr_autotune_reset.restype = None
        lib.cr_autotune_report.argtypes = [C.c_char_p, c_size_t]
        lib.cr_autotune_report.restype = C.c_int
        lib.cuda_replace_result.argtypes = [c_void_p, C.POINTER(c_uint8_p), c_size_t_p]
        lib.cuda_replace_result.restype = C.c_int
        lib.cuda_replace_close.argtypes = [c_void_p]
        lib.cuda_replace_close.restype = None
        lib.cuda_replace_reset.argtypes = [c_void_p, c_uint8_p, c_size_t]
        lib.cuda_replace_reset.restype = C.c_int
        lib.cuda_replace_apply_seeded.argtypes = [c_void_p, c_uint8_p, C.c_int, c_uint8_p, C.c_int,
                                                  C.c_longlong, c_longlong_p, c_size_t_p]
        lib.cuda_replace_apply_seeded.restype = C.c_int
        lib.cuda_replace_query_prefix.argtypes = [c_void_p, c_size_t, C.POINTER(C.c_int),
                                                  C.POINTER(C.c_int)]
        lib.cuda_replace_query_prefix.restype = C.c_int
        lib.cuda_replace_version.argtypes = []
        lib.cuda_replace_version.restype = C.c_char_p
        lib.cr_session_length.argtypes = [c_void_p, c_size_t_p]
        lib.cr_session_length.restype = C.c_int
        lib.cr_query_prefix64.argtypes = [c_void_p, c_size_t, c_longlong_p, c_longlong_p]
        lib.cr_query_prefix64.restype = C.c_int
        lib.cr_replace_device.argtypes = [c_void_p, c_void_p, c_size_t, c_void_p, c_size_t, c_uint8_p,
                                          C.c_int, c_uint8_p, C.c_int, C.c_longlong, c_size_t_p,
                                          c_longlong_p, C.POINTER(C.c_float)]
        lib.cr_replace_device.restype = C.c_int
        lib.cr_replace_device_noalloc.argtypes = [c_void_p, c_size_t, c_void_p, c_size_t, c_uint8_p, C.c_int,
                                                  c_uint8_p, C.c_int, C.c_longlong, c_void_p, c_size_t_p]
        lib.cr_replace_device_noalloc.restype = C.c_int
        lib.cr_pattern_mode.argtypes = [c_uint8_p, C.c_int]
        lib.cr_pattern_mode.restype = C.c_int
        lib.cr_batch_create.argtypes = [C.POINTER(c_void_p)]
        lib.cr_batch_create.restype = C.c_int
        c_char_pp = C.POINTER(C.c_char_p)
        lib.cr_batch_run.argtypes = [c_void_p, C.c_int, c_char_pp, c_size_t_p, c_int_p, c_char_pp, c_int_p,
                                     c_char_pp, c_int_p, C.POINTER(c_uint8_p), c_size_t_p, c_size_t_p,
                                     C.POINTER(C.c_double)]
        lib.cr_batch_run.restype = C.c_int
        lib.cr_batch_destroy.argtypes = [c_void_p]
        lib.cr_batch_destroy.restype = None
        lib.cr_batch_set_mode.argtypes = [c_void_p, C.c_int]
        lib.cr_batch_set_mode.restype = C.c_int
        lib.cr_batch_info.argtypes = [c_void_p, C.POINTER(C.c_double), C.c_int]
        lib.cr_batch_info.restype = C.c_int
        lib.cr_stream_replace.argtypes = [c_void_p, C.c_char_p, c_size_t, c_char_pp, c_int_p, c_char_pp, c_int_p,
                                          C.c_int, c_size_t, C.POINTER(c_uint8_p), c_size_t_p,
                                          C.POINTER(C.c_double)]
        lib.cr_stream_replace.restype = C.c_int
        lib.cr_server_create.argtypes = [C.POINTER(c_void_p), c_size_t, C.c_int, c_size_t]
        lib.cr_server_create.restype = C.c_int
        lib.cr_server_submit.argtypes = [c_void_p, c_void_p, c_size_t, C.c_int, c_char_pp, c_int_p, c_char_pp,
                                         c_int_p, C.POINTER(c_void_p)]
        lib.cr_server_submit.restype = C.c_int
        lib.cr_server_wait.argtypes = [c_void_p, C.POINTER(c_uint8_p), c_size_t_p, C.POINTER(C.c_double)]
        lib.cr_server_wait.restype = C.c_int
        lib.cr_server_release.argtypes = [c_void_p]
        lib.cr_server_release.restype = None
        lib.cr_server_destroy.argtypes = [c_void_p]
        lib.cr_server_destroy.restype = None

    def version(self) -> str:
        return self.lib.cuda_replace_version().decode()

    def pattern_mode(self, pat: bytes) -> int:
        return int(self.lib.cr_pattern_mode(_u8(pat), len(pat)))

    def autotune_set(self, pat: bytes, n: int, group_tiles: int) -> None:
        """Pins the tiles per group (1, 2 or 4; 0 unpins) for pat's pattern class and
        the size bucket of n on the current device.  For a border-free pattern 1/2/4
        pin the 32-byte-per-thread form with that group size, 8 the three-launch
        64-byte form, 16 the single launch and 32 the single launch that holds each
        group in LDS."""
        rc = int(self.lib.cr_autotune_set(_u8(pat), len(pat), int(n), int(group_tiles)))
        if rc != 0:
            raise RuntimeError(f"cr_autotune_set failed rc={rc}")

    def autotune_reset(self) -> None:
        self.lib.cr_autotune_reset()

    def autotune_report(self) -> list:
        """The autotuner's current choices with their measured candidate times."""
        import json
        need = int(self.lib.cr_autotune_report(None, 0))
        buf = C.create_string_buffer(need + 1)
        self.lib.cr_autotune_report(buf, need + 1)
        return json.loads(buf.value.decode())

    def batch_fusable(self, pairs: Sequence[Tuple[bytes, bytes]]) -> bool:
        """True if the library would run these pairs as one fused pass (host check
        only, no GPU call)."""
        if not pairs:
            return False
        arrs = _pair_arrays(pairs)
        rc = int(self.lib.cr_batch_fusable(*arrs[:4], len(pairs)))
        if rc < 0:
            raise RuntimeError(f"cr_batch_fusable failed rc={rc}")
        return rc == 1

    def unified(self, src: bytes, pat: bytes, rep: bytes) -> bytes:
        """One-shot replace (py:124-141)."""
        out_ptr = c_uint8_p()
        out_len = c_size_t()
        rc = self.lib.cuda_replace_unified(_u8(src), len(src), _u8(pat), len(pat), _u8(rep), len(rep),
                                           C.byref(out_ptr), C.byref(out_len))
        if rc != 0:
            raise RuntimeError(f"cuda_replace_unified failed rc={rc}")
        try:
            return _bytes_at(out_ptr, out_len.value)
        finally:
            self.lib.cuda_free_host(out_ptr)


class Session:
    """A device-resident buffer with replace passes (py:145-297).

    Distinct sessions may be used from different threads concurrently; calls on
    one session are serialized by the library."""

    def __init__(self, lib: CudaReplaceLib, src: bytes):
        self.lib = lib
        self.h = c_void_p()
        rc = lib.lib.cuda_replace_open(C.byref(self.h), _u8(src), len(src))
        if rc != 0:
            raise RuntimeError(f"cuda_replace_open failed rc={rc}")

    def __enter__(self) -> "Session":
        return self

    def __exit__(self, exc_type, exc, tb):
        self.close()

    def reset(self, src: bytes):
        rc = self.lib.lib.cuda_replace_reset(self.h, _u8(src), len(src))
        if rc != 0:
            raise RuntimeError(f"cuda_replace_reset failed rc={rc}")

    def build_index(self, chunk_size: Optional[int] = None):
        rc = self.lib.lib.cuda_replace_build_index(self.h, 0 if chunk_size is None else int(chunk_size))
        if rc != 0:
            raise RuntimeError(f"cuda_replace_build_index failed rc={rc}")

    def apply(self, pat: bytes, rep: bytes) -> int:
        new_len = c_size_t()
        rc = self.lib.lib.cuda_replace_apply(self.h, _u8(pat), len(pat), _u8(rep), len(rep),
                                             C.byref(new_len))
        if rc != 0:
            raise RuntimeError(f"cuda_replace_apply failed rc={rc}")
        return int(new_len.value)

    def apply_seeded(self, pat: bytes, rep: bytes, prev_last_rel: int) -> Tuple[int, int]:
        new_len = c_size_t()
        last_rel = C.c_longlong(-1)
        rc = self.lib.lib.cuda_replace_apply_seeded(self.h, _u8(pat), len(pat), _u8(rep), len(rep),
                                                    C.c_longlong(int(prev_last_rel)), C.byref(last_rel),
                                                    C.byref(new_len))
        if rc != 0:
            raise RuntimeError(f"cuda_replace_apply_seeded failed rc={rc}")
        return int(new_len.value), int(last_rel.value)

    def apply_batch(self, pairs: Sequence[Tuple[bytes, bytes]], fuse: bool = True) -> int:
        """Applies the pairs in order.  With fuse (the default, as in
        cuda_replace_apply_batch) pairs that cr_batch_fusable accepts run as one
        pass; self.last_batch_fused says whether they did."""
        count = len(pairs)
        self.last_batch_fused = False
        if count == 0:
            return self.length()
        # the pointer arrays of the last pair list are reused when equal pairs come
        # again (the library keeps that list's table too): the arrays hold copies,
        # and the kept key is a snapshot of pairs of bytes, which cannot change
        key = tuple(pairs)
        cached = getattr(self, "_batch_arrays", None)
        if cached is not None and cached[0] == key:
            arrs = cached[1]
        else:
            arrs = _pair_arrays(pairs)
            immutable = all(type(pr) is tuple and len(pr) == 2 and type(pr[0]) is bytes and type(pr[1]) is bytes
                            for pr in key)
            self._batch_arrays = (key, arrs) if immutable else None
        new_len = c_size_t()
        fused = C.c_int(0)
        rc = self.lib.lib.cr_apply_batch_ex(self.h, *arrs[:4], count, C.byref(new_len), 1 if fuse else 0,
                                            C.byref(fused))
        if rc != 0:
            raise RuntimeError(f"cuda_replace_apply_batch failed rc={rc}")
        self.last_batch_fused = bool(fused.value)
        return int(new_len.value)

    def query_prefix(self, limit: int) -> Tuple[int, int]:
        kept = C.c_longlong(0)
        last = C.c_longlong(-1)
        rc = self.lib.lib.cr_query_prefix64(self.h, int(limit), C.byref(kept), C.byref(last))
        if rc != 0:
            raise RuntimeError(f"cuda_replace_query_prefix failed rc={rc}")
        return int(kept.value), int(last.value)

    def result(self) -> bytes:
        out_ptr = c_uint8_p()
        out_len = c_size_t()
        rc = self.lib.lib.cuda_replace_result(self.h, C.byref(out_ptr), C.byref(out_len))
        if rc != 0:
            raise RuntimeError(f"cuda_replace_result failed rc={rc}")
        try:
            return _bytes_at(out_ptr, out_len.value)
        finally:
            self.lib.lib.cuda_free_host(out_ptr)

    def length(self) -> int:
        n = c_size_t()
        rc = self.lib.lib.cr_session_length(self.h, C.byref(n))
        if rc != 0:
            raise RuntimeError(f"cuda_replace_result(len) failed rc={rc}")
        return int(n.value)

    def close(self):
        if self.h:
            self.lib.lib.cuda_replace_close(self.h)
            self.h = c_void_p(None)


def replace_unified(lib: CudaReplaceLib, src: bytes, pat: bytes, rep: bytes) -> bytes:
    """One-shot convenience around cuda_replace_unified() (py:299-301)."""
    return lib.unified(src, pat, rep)


def gpu_replace_streaming(lib: CudaReplaceLib, src: bytes, pairs: Iterable[Tuple[bytes, bytes]],
                          chunk_bytes: int = 256 * 1024 * 1024, stats: Optional[dict] = None) -> bytes:
    """Memory-bounded chunked replace (py:305-366) through the library's pipelined
    streaming (``cr_stream_replace``): per pair, chunks of ``chunk_bytes`` with an
    L-1 byte overlap, the seed carried from chunk to chunk, each chunk's output cut
    after the kept match that straddles its end; the next chunk uploads while the
    current one computes.  The result equals ``bytes.replace`` of the whole input
    (empty patterns included).  ``stats``, when given, receives wall ms, bytes in
    and out summed over pairs, and the number of chunks."""
    if chunk_bytes <= 0:
        chunk_bytes = 1
    pairs = list(pairs)
    k = len(pairs)
    pats = [p for p, _ in pairs]
    reps = [r for _, r in pairs]
    out_ptr = c_uint8_p()
    out_len = c_size_t()
    st = (C.c_double * 4)()
    with Session(lib, b"") as sess:
        rc = lib.lib.cr_stream_replace(sess.h, src, len(src), (C.c_char_p * max(k, 1))(*pats),
                                       (C.c_int * max(k, 1))(*map(len, pats)), (C.c_char_p * max(k, 1))(*reps),
                                       (C.c_int * max(k, 1))(*map(len, reps)), k, chunk_bytes,
                                       C.byref(out_ptr), C.byref(out_len), st)
        if rc != 0:
            raise RuntimeError(f"cr_stream_replace failed rc={rc}")
        try:
            out = _bytes_at(out_ptr, out_len.value)
        finally:
            if out_ptr:
                lib.lib.cuda_free_host(out_ptr)
    if stats is not None:
        stats.update({"wall_ms": st[0], "bytes_in": st[1], "bytes_out": st[2], "chunks": int(st[3])})
    return out


class BatchEngine:
    """Independent documents, each with its own ordered list of (pattern,
    replacement) pairs, replaced together on the GPU (``cr_batch_run``).  Every
    result equals applying the document's pairs in order with ``bytes.replace``.
    The engine keeps its device arenas and pinned staging between runs."""

    def __init__(self, lib: CudaReplaceLib, host_rounds: bool = False):
        """host_rounds=True keeps every run on the host-driven rounds (one readback
        of kept counts per round) instead of the device-driven, graph-captured
        rounds used by default."""
        self.lib = lib
        self.h = c_void_p()
        rc = lib.lib.cr_batch_create(C.byref(self.h))
        if rc != 0:
            raise RuntimeError(f"cr_batch_create failed rc={rc}")
        if host_rounds and lib.lib.cr_batch_set_mode(self.h, 1) != 0:
            raise RuntimeError("cr_batch_set_mode failed")
        self.last_stats = None

    def run(self, docs: Sequence[bytes], pairs: Sequence[Sequence[Tuple[bytes, bytes]]],
            copy: bool = True) -> list:
        """Results in document order: ``bytes`` objects, or with ``copy=False``
        read-only memoryviews into the engine's pinned output buffer, valid until
        the next ``run`` or ``close``."""
        if len(docs) != len(pairs):
            raise ValueError("one pair list per document")
        n = len(docs)
        pats = [p for per in pairs for p, _ in per]
        reps = [r for per in pairs for _, r in per]
        m = len(pats)
        doc_ptrs = (C.c_char_p * max(n, 1))(*docs)
        doc_lens = (c_size_t * max(n, 1))(*map(len, docs))
        counts = (C.c_int * max(n, 1))(*map(len, pairs))
        # patterns and replacements: one joined buffer each and numpy-built pointer
        # arrays (per-element ctypes conversion of tens of thousands of pairs costs
        # tens of milliseconds)
        pat_blob, pat_ptr_arr, pat_len_arr = _blob_arrays(pats)
        rep_blob, rep_ptr_arr, rep_len_arr = _blob_arrays(reps)
        pat_ptrs = pat_ptr_arr.ctypes.data_as(C.POINTER(C.c_char_p))
        pat_lens = pat_len_arr.ctypes.data_as(c_int_p)
        rep_ptrs = rep_ptr_arr.ctypes.data_as(C.POINTER(C.c_char_p))
        rep_lens = rep_len_arr.ctypes.data_as(c_int_p)
        base = c_uint8_p()
        offs = (c_size_t * max(n, 1))()
        lens = (c_size_t * max(n, 1))()
        stats = (C.c_double * 4)()
        rc = self.lib.lib.cr_batch_run(self.h, n, doc_ptrs, doc_lens, counts, pat_ptrs, pat_lens, rep_ptrs,
                                       rep_lens, C.byref(base), offs, lens, stats)
        if rc != 0:
            raise RuntimeError(f"cr_batch_run failed rc={rc}")
        info = (C.c_double * 7)()
        self.lib.lib.cr_batch_info(self.h, info, 7)
        self.last_stats = {"device_ms": stats[0], "wall_ms": stats[1], "algo_bytes": stats[2],
                           "rounds": int(stats[3]), "path": "device" if info[0] == 2 else "host",
                           "group_tiles": int(info[1]), "graph_replayed": bool(info[2]),
                           "tuned": bool(info[3]), "graph": bool(info[5]), "resumes": int(info[6])}
        addr = C.cast(base, c_void_p).value or 0
        if not copy:
            total = max((offs[i] + lens[i] for i in range(n)), default=0)
            whole = memoryview((C.c_uint8 * total).from_address(addr)).cast("B") if total else memoryview(b"")
            return [whole[offs[i]:offs[i] + lens[i]].toreadonly() for i in range(n)]
        return [_bytes_at(addr + offs[i], lens[i]) for i in range(n)]

    def close(self) -> None:
        if self.h:
            self.lib.lib.cr_batch_destroy(self.h)
            self.h = c_void_p()

    def __enter__(self) -> "BatchEngine":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Ticket:
    """One request queued on a ``Server``; ``result()`` blocks until it is done."""

    def __init__(self, server: "Server", handle: c_void_p, keep: tuple):
        self._server = server
        self._h = handle
        self._keep = keep          # the request's bytes stay alive until the result is read
        self._out: Optional[bytes] = None
        self.times: Optional[dict] = None

    def wait(self) -> memoryview:
        """Blocks until done and returns a read-only view of the result, valid
        until ``release()``; ``times`` is filled in."""
        if self._out is not None:
            return memoryview(self._out)
        out = c_uint8_p()
        n = c_size_t()
        t = (C.c_double * 3)()
        rc = self._server.lib.lib.cr_server_wait(self._h, C.byref(out), C.byref(n), t)
        if rc != 0:
            self.release()
            raise RuntimeError(f"cr_server_wait failed rc={rc}")
        self.times = {"queue_ms": t[0], "latency_ms": t[1], "batch_docs": int(t[2])}
        addr = C.cast(out, c_void_p).value
        if not n.value:
            return memoryview(b"")
        return memoryview((C.c_uint8 * n.value).from_address(addr)).cast("B").toreadonly()

    def release(self) -> None:
        """Frees the request and its result (views from ``wait()`` become invalid)."""
        if self._h:
            self._server.lib.lib.cr_server_release(self._h)
            self._h = c_void_p()
            self._keep = ()

    def __del__(self):
        # the queued request reads the bytes this ticket keeps alive: a ticket
        # dropped unread is released only once its request is done
        # (cr_server_release waits for it)
        try:
            self.release()
        except Exception:
            pass

    def result(self) -> bytes:
        if self._out is None:
            out = c_uint8_p()
            n = c_size_t()
            t = (C.c_double * 3)()
            rc = self._server.lib.lib.cr_server_wait(self._h, C.byref(out), C.byref(n), t)
            try:
                if rc != 0:
                    raise RuntimeError(f"cr_server_wait failed rc={rc}")
                self._out = _bytes_at(out, n.value)
                self.times = {"queue_ms": t[0], "latency_ms": t[1], "batch_docs": int(t[2])}
            finally:
                self._server.lib.lib.cr_server_release(self._h)
                self._h = c_void_p()
                self._keep = ()
        return self._out


class Server:
    """Continuous-batching request server (``cr_server_*``): requests are admitted
    in arrival order under an HBM budget and run in batches by ``workers`` batch
    engines, each on its own stream.  Each result equals applying the request's
    pairs in order with ``bytes.replace``."""

    def __init__(self, lib: CudaReplaceLib, hbm_budget: int = 0, workers: int = 4, max_batch_bytes: int = 0):
        self.lib = lib
        self.h = c_void_p()
        rc = lib.lib.cr_server_create(C.byref(self.h), hbm_budget, workers, max_batch_bytes)
        if rc != 0:
            raise RuntimeError(f"cr_server_create failed rc={rc}")

    def submit(self, doc: bytes, pairs: Sequence[Tuple[bytes, bytes]]) -> Ticket:
        return self.submit_at(doc, 0, len(doc), pairs)

    def submit_at(self, buf: bytes, offset: int, length: int, pairs: Sequence[Tuple[bytes, bytes]]) -> Ticket:
        """Queues ``buf[offset:offset + length]`` without copying it."""
        if offset < 0 or length < 0 or offset + length > len(buf):
            raise ValueError("slice outside the buffer")
        k = len(pairs)
        pats = [p for p, _ in pairs]
        reps = [r for _, r in pairs]
        pat_arr = (C.c_char_p * max(k, 1))(*pats)
        rep_arr = (C.c_char_p * max(k, 1))(*reps)
        t = c_void_p()
        addr = C.cast(C.c_char_p(buf), c_void_p).value + offset
        rc = self.lib.lib.cr_server_submit(self.h, addr, length, k, pat_arr, (C.c_int * max(k, 1))(*map(len, pats)),
                                           rep_arr, (C.c_int * max(k, 1))(*map(len, reps)), C.byref(t))
        if rc != 0:
            raise RuntimeError(f"cr_server_submit failed rc={rc}")
        return Ticket(self, t, (buf, pats, reps))

    def close(self) -> None:
        if self.h:
            self.lib.lib.cr_server_destroy(self.h)
            self.h = c_void_p()

    def __enter__(self) -> "Server":
        return self

    def __exit__(self, *exc) -> None:
        self.close()
