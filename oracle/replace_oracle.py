"""CPU oracle for the byte-pattern replace path.

TEST INFRASTRUCTURE ONLY: this module is the checker for the HIP library in
``cuda_replace_amd``.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import it; the product never does.

The stated contract of the reference is Python's ``bytes.replace``
(cuda_replace.cu:2-3, README.md:3-7), so the single-pair, seedless oracle *is*
CPython's ``bytes.replace``.  What ``bytes.replace`` cannot express -- the seed
carry of ``apply_seeded`` and the kept-start queries -- is restated here from the
reference source:

* match definition with bounds check ......... cuda_replace.cu:119-129
* greedy with an initial ``last`` seed ........ cuda_replace.cu:938-968
* output layout (kept start -> replacement, covered bytes skipped, all other
  bytes copied) .............................. cuda_replace.cu:1391-1467, 1520-1559
* prefix query (count, last kept below limit) . cuda_replace.cu:1825-1897
* batch = sequential seedless applies ........ cuda_replace.cu:2465-2527
* streaming protocol ......................... cuda_replace_wrapper.py:305-366,
  with the straddling-match cut fixed (a kept match that starts before the
  chunk end and ends after it moves the cut to its end).

Parity status of each function (pinned by tests/test_oracle.py):

* ``replace``            -- CPython ``bytes.replace`` itself.
* ``kept_positions``     -- pinned: CPython ``bytes.find`` loop, ``bytes.count``
  for border-free patterns, closed forms on periodic inputs, brute force.
* ``apply_seeded``       -- pinned: ``S[:F] + S[F:].replace(P, R)`` with
  ``F = max(0, seed + L)`` by brute force over every input of up to 5 bytes on
  ``{a, b, NUL}`` (plus 600 seeded random inputs of 6-16 bytes) with every seed
  in ``[-L-2, n+1]``.
* ``query_prefix``       -- pinned through ``kept_positions``.
* ``apply_batch``        -- ``functools.reduce`` of ``bytes.replace``.
* ``streaming``          -- pinned: equals ``bytes.replace`` of the whole input
  for every chunk size (brute force over small inputs).
"""
from __future__ import annotations

import ctypes as C
import functools
import os
from typing import Iterable, List, Optional, Sequence, Tuple

NO_SEED = None


def replace(src: bytes, pat: bytes, rep: bytes) -> bytes:
    """The contract: CPython bytes.replace (cuda_replace.cu:2-3)."""
    return src.replace(pat, rep)


def _floor(seed: Optional[int], L: int) -> int:
    """First position a candidate may be kept at, given a seed "last kept"
    position (cuda_replace.cu:947-953: keep i iff i - last >= L)."""
    if seed is None:
        return 0
    return max(0, seed + L)


def is_match(src: bytes, pos: int, pat: bytes) -> bool:
    """cuda_replace.cu:119-129: bounds-checked byte compare."""
    L = len(pat)
    if L > len(src) or pos + L > len(src):
        return False
    return src[pos:pos + L] == pat


def kept_positions(src: bytes, pat: bytes, seed: Optional[int] = None) -> List[int]:
    """Greedy kept starts, restating the loop of cuda_replace.cu:938-968 over
    the candidates of cuda_replace.cu:119-129.  Pure Python; small inputs."""
    L = len(pat)
    if L == 0:
        return list(range(len(src) + 1))
    kept: List[int] = []
    last = seed
    i = src.find(pat)
    while i != -1:
        if last is None or i - last >= L:
            kept.append(i)
            last = i
        i = src.find(pat, i + 1)
    return kept


def apply_seeded(src: bytes, pat: bytes, rep: bytes,
                 seed: Optional[int] = None) -> Tuple[bytes, int]:
    """One apply with a seed; returns (output, last kept start or -1).

    Output layout restated from cuda_replace.cu:1391-1467 / 1520-1559: walk the
    source, emit ``rep`` at a kept start and skip its L bytes, copy every other
    byte.  An empty pattern follows bytes.replace (seed ignored)."""
    L = len(pat)
    if L == 0:
        return src.replace(b"", rep), len(src)
    kept = kept_positions(src, pat, seed)
    out = bytearray()
    prev = 0
    for k in kept:
        out += src[prev:k]
        out += rep
        prev = k + L
    out += src[prev:]
    return bytes(out), (kept[-1] if kept else -1)


def query_prefix(src: bytes, pat: bytes, seed: Optional[int], limit: int) -> Tuple[int, int]:
    """(#kept starts < limit, last kept start < limit or -1), in the source
    coordinates of the apply (cuda_replace.cu:1825-1897); limit clamped to n."""
    limit = min(limit, len(src))
    ks = [k for k in kept_positions(src, pat, seed) if k < limit]
    return len(ks), (ks[-1] if ks else -1)


def apply_batch(src: bytes, pairs: Sequence[Tuple[bytes, bytes]]) -> bytes:
    """Sequential seedless applies (cuda_replace.cu:2480-2519)."""
    return functools.reduce(lambda s, pr: s.replace(pr[0], pr[1]), pairs, src)


def streaming(src: bytes, pairs: Iterable[Tuple[bytes, bytes]], chunk_bytes: int) -> bytes:
    """The chunked protocol of cuda_replace_wrapper.py:305-366 computed on the
    CPU: chunks of ``chunk_bytes`` with an L-1 byte overlap, the seed carried
    from the previous chunk, the prefix query cutting the output.  The cut is
    moved to the end of a kept match that straddles the chunk end (the wrapper
    cuts at the chunk end and drops bytes there)."""
    chunk_bytes = max(1, chunk_bytes)
    out = src
    for pat, rep in pairs:
        L = len(pat)
        if L == 0:
            out = out.replace(b"", rep)
            continue
        diff = len(rep) - L
        pos, prev_last, n = 0, None, len(out)
        pieces: List[bytes] = []
        while pos < n:
            base_end = min(pos + chunk_bytes, n)
            take_end = min(base_end + L - 1, n)
            chunk = out[pos:take_end]
            seed = None if prev_last is None else prev_last - pos
            gpu_out, _ = apply_seeded(chunk, pat, rep, seed)
            kept_before, last_before = query_prefix(chunk, pat, seed, base_end - pos)
            cut = base_end - pos
            if last_before >= 0:
                cut = max(cut, last_before + L)
            pieces.append(gpu_out[:cut + kept_before * diff])
            if last_before >= 0:
                prev_last = pos + last_before
            pos += cut
        out = b"".join(pieces)
    return out


# --------------------------------------------------------------------------
# The C restatement (oracle/cr_oracle.c) for inputs too large for Python loops.
# --------------------------------------------------------------------------
_HERE = os.path.dirname(os.path.abspath(__file__))
C_LIB_PATH = os.path.join(_HERE, "_build", "libcr_oracle.so")
_NO_SEED_C = -(1 << 63)
_clib = None


def c_lib():
    """Load the compiled C oracle (built by ``oracle/Makefile``)."""
    global _clib
    if _clib is None:
        if not os.path.exists(C_LIB_PATH):
            raise FileNotFoundError(f"C oracle not built: {C_LIB_PATH} (run make -C oracle)")
        lib = C.CDLL(C_LIB_PATH)
        i64, p = C.c_int64, C.c_void_p
        lib.cro_kept.argtypes = [p, i64, p, i64, i64, p, i64]
        lib.cro_kept.restype = i64
        lib.cro_apply.argtypes = [p, i64, p, i64, p, i64, i64, p, i64, p, p]
        lib.cro_apply.restype = i64
        lib.cro_apply_len.argtypes = [p, i64, p, i64, i64, i64]
        lib.cro_apply_len.restype = i64
        lib.cro_query_prefix.argtypes = [p, i64, p, i64, i64, i64, p, p]
        lib.cro_query_prefix.restype = None
        _clib = lib
    return _clib


def _cbuf(b: bytes):
    return C.c_char_p(b) if b else C.c_char_p(b"\0")


def c_apply_seeded(src: bytes, pat: bytes, rep: bytes,
                   seed: Optional[int] = None) -> Tuple[bytes, int, int]:
    """C restatement of apply: returns (output, kept count, last kept)."""
    lib = c_lib()
    s = _NO_SEED_C if seed is None else int(seed)
    n_out = lib.cro_apply_len(_cbuf(src), len(src), _cbuf(pat), len(pat), len(rep), s)
    out = C.create_string_buffer(max(1, n_out))
    kept, last = C.c_int64(), C.c_int64()
    got = lib.cro_apply(_cbuf(src), len(src), _cbuf(pat), len(pat), _cbuf(rep), len(rep), s,
                        out, n_out, C.byref(kept), C.byref(last))
    if got != n_out:
        raise RuntimeError("C oracle length mismatch")
    return out.raw[:n_out], kept.value, last.value


def c_kept_count(src: bytes, pat: bytes, seed: Optional[int] = None) -> int:
    """Kept count by the reference-shaped loop (cuda_replace.cu:938-968)."""
    s = _NO_SEED_C if seed is None else int(seed)
    return c_lib().cro_kept(_cbuf(src), len(src), _cbuf(pat), len(pat), s, None, 0)


def c_query_prefix(src: bytes, pat: bytes, seed: Optional[int], limit: int) -> Tuple[int, int]:
    s = _NO_SEED_C if seed is None else int(seed)
    cnt, last = C.c_int64(), C.c_int64()
    c_lib().cro_query_prefix(_cbuf(src), len(src), _cbuf(pat), len(pat), s, int(limit),
                             C.byref(cnt), C.byref(last))
    return cnt.value, last.value
