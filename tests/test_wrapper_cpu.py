"""Host-side helpers of the Python wrapper (no GPU call): the pointer and length
arrays it hands to the C ABI must address exactly the caller's bytes."""
import ctypes as C
import random

from cuda_replace_amd.wrapper import _blob_arrays, _pair_arrays


def test_blob_arrays_address_every_item():
    rng = random.Random(5)
    items = [bytes(rng.randrange(256) for _ in range(rng.randint(0, 40))) for _ in range(2000)]
    items += [b"", b"\x00", b"a" * 300]
    blob, ptrs, lens = _blob_arrays(items)
    assert len(ptrs) == len(items) and len(lens) == len(items)
    for i, it in enumerate(items):
        assert int(lens[i]) == len(it)
        assert C.string_at(int(ptrs[i]), int(lens[i])) == it


def test_blob_arrays_empty_list():
    blob, ptrs, lens = _blob_arrays([])
    assert len(ptrs) == 1 and len(lens) == 1 and int(lens[0]) == 0


def test_pair_arrays_match_pairs():
    pairs = [(b"&amp;", b"&"), (b"", b"-"), (bytearray(b"ab"), b"XYZ")]
    pats, pat_lens, reps, rep_lens, keep = _pair_arrays(pairs)
    for i, (p, r) in enumerate(pairs):
        assert pat_lens[i] == len(p) and rep_lens[i] == len(r)
        assert C.string_at(C.cast(pats[i], C.c_void_p).value or 0, pat_lens[i]) == bytes(p)
        assert C.string_at(C.cast(reps[i], C.c_void_p).value or 0, rep_lens[i]) == bytes(r)


def test_pair_arrays_without_numpy(monkeypatch):
    """The reference wrapper needs only the standard library (py:1-6); with numpy
    absent the pair arrays are built with ctypes and address the same bytes."""
    import sys
    monkeypatch.setitem(sys.modules, "numpy", None)   # `import numpy` raises ImportError
    rng = random.Random(9)
    items = [bytes(rng.randrange(256) for _ in range(rng.randint(0, 20))) for _ in range(300)] + [b""]
    blob, ptrs, lens = _blob_arrays(items)
    lp = ptrs.ctypes.data_as(C.POINTER(C.c_uint64))
    ll = lens.ctypes.data_as(C.POINTER(C.c_int))
    for i, it in enumerate(items):
        assert ll[i] == len(it)
        assert C.string_at(lp[i], ll[i]) == it
    pairs = [(b"&amp;", b"&"), (b"", b"-"), (b"ab", b"XYZ")]
    pats, pat_lens, reps, rep_lens, keep = _pair_arrays(pairs)
    for i, (p, r) in enumerate(pairs):
        assert pat_lens[i] == len(p) and rep_lens[i] == len(r)
        assert C.string_at(C.cast(pats[i], C.c_void_p).value or 0, pat_lens[i]) == p
        assert C.string_at(C.cast(reps[i], C.c_void_p).value or 0, rep_lens[i]) == r
    blob, ptrs, lens = _blob_arrays([])
    assert lens.ctypes.data_as(C.POINTER(C.c_int))[0] == 0
