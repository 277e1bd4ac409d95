"""cr_replace_device_noalloc against cr_replace_device (gpu).

The allocation-free entry point runs the same kernels with its scratch in a static
device array; its output must equal cr_replace_device's byte for byte (and
bytes.replace), on two calls in a row with the second input smaller, for every
pattern class it accepts, on the default stream and on a stream of the caller's.  Requests
beyond its limits return -5 before touching the GPU."""
import ctypes as C

import pytest

from tests.hipmem import DevBuf, Stream, synchronize

pytestmark = pytest.mark.gpu

u8p = C.POINTER(C.c_uint8)


def say(*a):
    print(*a, flush=True)


def text(n, seed):
    import numpy as np
    rng = np.random.default_rng(seed)
    words = [b"alpha ", b"beta\n", b"qarn ", b"hello ", b"aaab", b"aa", b"x", b"\x00", b"ab", b"abcab "]
    idx = rng.integers(0, len(words), n // 3 + 8)
    return b"".join(words[i] for i in idx)[:n]


def dev_pass(lib, fn, src_buf, n, pat, rep, seed, cap, stream=None, h=None):
    """One pass into a fresh destination filled with 0xEE; returns (rc, out_len,
    destination buffer)."""
    dst = DevBuf(cap, fill=0xEE)
    synchronize()
    out_len = C.c_size_t()
    P = C.cast(C.c_char_p(pat), u8p)
    R = C.cast(C.c_char_p(rep), u8p)
    if fn == "noalloc":
        rc = lib.lib.cr_replace_device_noalloc(src_buf.ptr, n, dst.ptr, cap, P, len(pat), R, len(rep), seed,
                                               stream, C.byref(out_len))
    else:
        rc = lib.lib.cr_replace_device(h, src_buf.ptr, n, dst.ptr, cap, P, len(pat), R, len(rep), seed,
                                       C.byref(out_len), None, None)
    synchronize()
    return rc, out_len.value, dst


CASES = [
    (b"qarn", b"0123456789AB", -4),      # border-free expand
    (b"hello", b"#", -5),                # border-free shrink
    (b"aa", b"X", -2),                   # carry tables, 2 states
    (b"abcab", b"", -5),                 # carry tables, 8 states
    (b"aaab", b"Z" * 40, 300000),        # border-free, seeded past the start
    (b"", b"|", 0),                      # empty pattern
    (b"x", b"", -1),                     # one-byte pattern, deletion
]


@pytest.mark.parametrize("pat,rep,seed", CASES, ids=[f"{p!r}->{len(r)}" for p, r, _ in CASES])
def test_noalloc_equals_session_pass(lib, pat, rep, seed):
    h = C.c_void_p()
    assert lib.lib.cuda_replace_open(C.byref(h), None, 0) == 0
    s2 = Stream()
    try:
        # two calls in a row, the second on a smaller input, and a third on a
        # stream of the caller's
        for k, (n, stream) in enumerate([(40 << 20, None), ((1 << 20) + 37, None), (300001, s2.handle)]):
            src = text(n, 100 + k)
            say("noalloc", pat, len(rep), seed, n)
            cap = len(src) * 14 + 64 if pat else len(src) * (len(rep) + 1) + len(rep) + 64
            with DevBuf.of(src) as sb:
                rc1, n1, d1 = dev_pass(lib, "noalloc", sb, n, pat, rep, seed, cap, stream=stream)
                rc2, n2, d2 = dev_pass(lib, "session", sb, n, pat, rep, seed, cap, h=h)
            assert rc1 == 0 and rc2 == 0
            assert n1 == n2
            a = d1.download(n1)
            assert a == d2.download(n2)
            f = 0 if not pat else max(0, seed + len(pat))
            assert a == src[:f] + src[f:].replace(pat, rep)
            # nothing is written past the output
            assert d1.download(64, n1) == b"\xee" * 64
            d1.free()
            d2.free()
    finally:
        s2.destroy()
        lib.lib.cuda_replace_close(h)


def test_noalloc_destination_too_small(lib):
    src = text(1 << 20, 7)
    with DevBuf.of(src) as sb:
        want = len(src.replace(b"qarn", b"0123456789AB"))
        rc, n, _ = dev_pass(lib, "noalloc", sb, len(src), b"qarn", b"0123456789AB", -4, len(src) + 16)
        assert rc == 1 and n == want          # hipErrorInvalidValue and the needed length
        rc, n, d = dev_pass(lib, "noalloc", sb, len(src), b"qarn", b"0123456789AB", -4, want + 16)
        assert rc == 0 and n == want
        assert d.download(n) == src.replace(b"qarn", b"0123456789AB")


def test_noalloc_limits(lib):
    out_len = C.c_size_t()
    P = C.cast(C.c_char_p(b"ab"), u8p)
    with DevBuf(64, fill=0) as t:

        def call(n, pat, rep):
            return lib.lib.cr_replace_device_noalloc(t.ptr, n, t.ptr + 32, 32, C.cast(C.c_char_p(pat), u8p),
                                                     len(pat), C.cast(C.c_char_p(rep), u8p), len(rep), -1, None,
                                                     C.byref(out_len))

        assert call((1 << 32) + 1, b"ab", b"X") == -5          # input over 4 GiB
        assert call(16, b"a" * 257, b"X") == -5               # pattern over 256 bytes
        assert call(16, b"ab", b"X" * 257) == -5              # replacement over 256 bytes
        assert call(16, b"a" * 17, b"X") == -5                # self-overlapping, longer than 16
        assert lib.lib.cr_replace_device_noalloc(None, 16, None, 0, P, 2, P, 2, -1, None, C.byref(out_len)) == -1
        assert lib.lib.cr_replace_device_noalloc(t.ptr, 16, t.ptr + 32, 32, P, 2, P, 2, -1, None, None) == -1
