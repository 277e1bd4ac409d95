"""The library bound the way the reference's ctypes wrapper binds it (gpu).

The reference wrapper (cuda_replace_wrapper.py:48-121) declares its own argtypes
and passes ``create_string_buffer`` copies: ``uint8_t*`` for every byte buffer,
``uint8_t**`` arrays for apply_batch (py:222-258), the *int* prefix query
(py:262-272), ``build_index`` on a live handle (py:178-185) and the length read
through ``cuda_replace_result`` (py:284-291).  These tests restate those call
shapes here (the reference module itself is never imported) and check every
result against the oracle, so the drop-in claim covers the exact forms a user of
the reference wrapper sends, not only this package's own wrapper."""
import ctypes as C
import random

import pytest

from oracle import replace_oracle as ro

pytestmark = pytest.mark.gpu

u8p = C.POINTER(C.c_uint8)
u8pp = C.POINTER(u8p)
intp = C.POINTER(C.c_int)
sizep = C.POINTER(C.c_size_t)
llp = C.POINTER(C.c_longlong)


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def raw(lib):
    """A second CDLL handle of the same library with the reference's signatures."""
    d = C.CDLL(lib.path)
    d.cuda_replace_unified.argtypes = [u8p, C.c_size_t, u8p, C.c_int, u8p, C.c_int, C.POINTER(u8p), sizep]
    d.cuda_replace_unified.restype = C.c_int
    d.cuda_free_host.argtypes = [C.c_void_p]
    d.cuda_free_host.restype = None
    d.cuda_replace_open.argtypes = [C.POINTER(C.c_void_p), u8p, C.c_size_t]
    d.cuda_replace_open.restype = C.c_int
    d.cuda_replace_build_index.argtypes = [C.c_void_p, C.c_size_t]
    d.cuda_replace_build_index.restype = C.c_int
    d.cuda_replace_apply.argtypes = [C.c_void_p, u8p, C.c_int, u8p, C.c_int, sizep]
    d.cuda_replace_apply.restype = C.c_int
    d.cuda_replace_apply_batch.argtypes = [C.c_void_p, u8pp, intp, u8pp, intp, C.c_int, sizep]
    d.cuda_replace_apply_batch.restype = C.c_int
    d.cuda_replace_result.argtypes = [C.c_void_p, C.POINTER(u8p), sizep]
    d.cuda_replace_result.restype = C.c_int
    d.cuda_replace_close.argtypes = [C.c_void_p]
    d.cuda_replace_close.restype = None
    d.cuda_replace_reset.argtypes = [C.c_void_p, u8p, C.c_size_t]
    d.cuda_replace_reset.restype = C.c_int
    d.cuda_replace_apply_seeded.argtypes = [C.c_void_p, u8p, C.c_int, u8p, C.c_int, C.c_longlong, llp, sizep]
    d.cuda_replace_apply_seeded.restype = C.c_int
    d.cuda_replace_query_prefix.argtypes = [C.c_void_p, C.c_size_t, intp, intp]
    d.cuda_replace_query_prefix.restype = C.c_int
    return d


def buf(b):
    """A NUL-terminated ctypes copy, as the reference wrapper makes for every input."""
    sb = C.create_string_buffer(b)
    return sb, C.cast(sb, u8p)


def read_result(raw, h):
    """cuda_replace_result + cuda_free_host; returns (bytes, length)."""
    out = u8p()
    n = C.c_size_t()
    assert raw.cuda_replace_result(h, C.byref(out), C.byref(n)) == 0
    try:
        data = C.string_at(out, n.value) if n.value else b""
    finally:
        raw.cuda_free_host(out)
    return data, n.value


def open_session(raw, src):
    keep, p = buf(src)
    h = C.c_void_p()
    assert raw.cuda_replace_open(C.byref(h), p, len(src)) == 0
    return h, keep


def text(n, seed):
    rng = random.Random(seed)
    words = [b"alpha", b"beta", b"&amp;", b"&lt;", b"&gt;", b"hello", b"aaab", b"a", b"\x00", b"ab"]
    out = bytearray()
    while len(out) < n:
        out += rng.choice(words) + rng.choice([b" ", b"\n", b""])
    return bytes(out[:n])


def test_unified_reference_shape(raw):
    src = text(50000, 1)
    for pat, rep in [(b"hello", b"goodbye"), (b"aa", b"X"), (b"&amp;", b"&"), (b"", b"|"), (b"a", b"")]:
        sk, sp = buf(src)
        pk, pp = buf(pat)
        rk, rp = buf(rep)
        out = u8p()
        n = C.c_size_t()
        assert raw.cuda_replace_unified(sp, len(src), pp, len(pat), rp, len(rep), C.byref(out), C.byref(n)) == 0
        got = C.string_at(out, n.value) if n.value else b""
        raw.cuda_free_host(out)
        assert got == src.replace(pat, rep), (pat, rep)


def test_build_index_then_apply(raw):
    """build_index on a live handle, at the default and at small chunk sizes
    (many chunks), then applies whose results the index must not change."""
    src = text(300000, 2)
    h, keep = open_session(raw, src)
    try:
        for cs in (0, 4096, 1, 1 << 30):
            assert raw.cuda_replace_build_index(h, cs) == 0, cs
        pk, pp = buf(b"hello")
        rk, rp = buf(b"HI")
        n = C.c_size_t()
        assert raw.cuda_replace_apply(h, pp, 5, rp, 2, C.byref(n)) == 0
        want = src.replace(b"hello", b"HI")
        assert n.value == len(want)
        assert raw.cuda_replace_build_index(h, 777) == 0
        data, ln = read_result(raw, h)
        assert data == want and ln == len(want)
    finally:
        raw.cuda_replace_close(h)


def batch_call(raw, h, pairs):
    """cuda_replace_apply_batch with the reference wrapper's argument shapes:
    one create_string_buffer per pattern and replacement, pointer arrays cast to
    uint8_t**, length arrays cast to int* (py:222-258)."""
    count = len(pairs)
    keep = []
    pat_ptrs, rep_ptrs = [], []
    pat_lens = (C.c_int * count)()
    rep_lens = (C.c_int * count)()
    for i, (p, r) in enumerate(pairs):
        pk, pp = buf(p)
        rk, rp = buf(r)
        keep += [pk, rk]
        pat_ptrs.append(pp)
        rep_ptrs.append(rp)
        pat_lens[i] = len(p)
        rep_lens[i] = len(r)
    pat_arr = (u8p * count)(*pat_ptrs)
    rep_arr = (u8p * count)(*rep_ptrs)
    new_len = C.c_size_t()
    rc = raw.cuda_replace_apply_batch(h, C.cast(pat_arr, u8pp), C.cast(pat_lens, intp), C.cast(rep_arr, u8pp),
                                      C.cast(rep_lens, intp), count, C.byref(new_len))
    return rc, new_len.value


@pytest.mark.parametrize("kind", ["fusable", "sequential", "mixed_modes"])
def test_apply_batch_reference_shape(raw, kind):
    src = text(400000, 3)
    if kind == "fusable":
        pairs = [(b"&amp;", b"1"), (b"&lt;", b"2"), (b"&gt;", b"3")]
    elif kind == "sequential":
        pairs = [(b"hello", b"ab"), (b"ab", b"X"), (b"X", b"")]
    else:
        pairs = [(b"aa", b"a"), (b"a" * 17, b"B"), (b"", b"-"), (b"beta", b"")] + [(b"\x00", b"NUL")]
    say("apply_batch", kind)
    h, keep = open_session(raw, src)
    try:
        rc, n = batch_call(raw, h, pairs)
        assert rc == 0
        want = ro.apply_batch(src, pairs)
        assert n == len(want)
        data, ln = read_result(raw, h)
        assert data == want and ln == len(want)
        # count == 0 returns the current length (the wrapper short-circuits it too)
        new_len = C.c_size_t()
        assert raw.cuda_replace_apply_batch(h, None, None, None, None, 0, C.byref(new_len)) == 0
        assert new_len.value == len(want)
    finally:
        raw.cuda_replace_close(h)


def test_seeded_and_int_query_reference_shape(raw):
    src = text(200000, 4)
    for pat, rep, seed in [(b"aa", b"Z", 1000), (b"hello", b"", -5), (b"&lt;", b"<<", 150000), (b"ab", b"", -1)]:
        h, keep = open_session(raw, src)
        try:
            pk, pp = buf(pat)
            rk, rp = buf(rep)
            last = C.c_longlong(-1)
            n = C.c_size_t()
            assert raw.cuda_replace_apply_seeded(h, pp, len(pat), rp, len(rep), seed, C.byref(last),
                                                 C.byref(n)) == 0
            want, want_last = ro.apply_seeded(src, pat, rep, seed)
            assert n.value == len(want) and last.value == want_last
            for limit in (0, 1, 777, 100000, len(src), len(src) + 10):
                kb = C.c_int(0)
                lb = C.c_int(-1)
                assert raw.cuda_replace_query_prefix(h, limit, C.byref(kb), C.byref(lb)) == 0
                assert (kb.value, lb.value) == ro.query_prefix(src, pat, seed, limit), (pat, limit)
            data, ln = read_result(raw, h)
            assert data == want
        finally:
            raw.cuda_replace_close(h)


def test_reset_and_length_reference_shape(raw):
    """reset with a create_string_buffer copy, then the wrapper's length(): a
    cuda_replace_result whose buffer is freed at once."""
    h, keep = open_session(raw, b"abc")
    try:
        for src in (text(9000, 5), b"", b"aaaa"):
            sk, sp = buf(src)
            assert raw.cuda_replace_reset(h, sp, len(src)) == 0
            pk, pp = buf(b"a")
            rk, rp = buf(b"xyz")
            n = C.c_size_t()
            assert raw.cuda_replace_apply(h, pp, 1, rp, 3, C.byref(n)) == 0
            data, ln = read_result(raw, h)
            assert ln == n.value == len(src.replace(b"a", b"xyz")) and data == src.replace(b"a", b"xyz")
    finally:
        raw.cuda_replace_close(h)


def test_device_pass_drops_the_query_record(lib):
    """cr_replace_device runs on the session's scratch: a prefix query after it
    answers 0 / -1 (no stale pattern or freed buffer is read) and the session's
    own result is untouched; the next apply restores the query."""
    from cuda_replace_amd import Session
    from tests.hipmem import DevBuf
    src = text(100000, 6)
    with Session(lib, src) as s:
        s.apply(b"hello", b"HI")
        want = src.replace(b"hello", b"HI")
        assert s.query_prefix(50000) == ro.query_prefix(src, b"hello", None, 50000)
        big = text(3 << 20, 7)
        with DevBuf.of(big) as d_src, DevBuf(len(big) * 3 + 16) as d_dst:
            out_len = C.c_size_t()
            rc = lib.lib.cr_replace_device(s.h, d_src.ptr, len(big), d_dst.ptr, d_dst.n,
                                           C.cast(C.c_char_p(b"a"), u8p), 1, C.cast(C.c_char_p(b"<A>"), u8p), 3,
                                           -1, C.byref(out_len), None, None)
            assert rc == 0
            assert d_dst.download(out_len.value) == big.replace(b"a", b"<A>")
        assert s.query_prefix(50000) == (0, -1)
        assert s.result() == want
        s.apply(b"beta", b"")
        w2 = want.replace(b"beta", b"")
        assert s.query_prefix(40000) == ro.query_prefix(want, b"beta", None, 40000)
        assert s.result() == w2
