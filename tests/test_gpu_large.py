"""Parity past 32-bit offsets (gpu): inputs of 2^31 bytes and more.

The reference keeps positions and prefixes in ``int`` (cuda_replace.cu:159-161,
609, 620-622, 684, 695, 1544), so every offset at or past 2^31 is where a port
would break.  These tests run each pass class over buffers that cross 2^31 (and
one that crosses 2^32) with matches planted across the boundaries, and check the
results with ``bytes.replace`` / ``count`` / ``rfind`` and closed forms:

* a border-free pattern, expand and shrink, plain and seeded with a seed past
  2^31, and prefix queries past INT_MAX (64-bit and the int ABI's -4 return);
* the carry-table pass (``aa`` -> ``X`` on all-'a') at 2^31 + 3 bytes, and at
  2^32 + 6 bytes, where the kept count itself exceeds INT_MAX;
* a fused multi-pattern apply_batch over 2^31 + 1 MiB, against the chained
  oracle and against the same pairs run pass by pass.
"""
import ctypes as C

import numpy as np
import pytest

from cuda_replace_amd import Session

pytestmark = pytest.mark.gpu

B31 = 1 << 31
INT_MAX = (1 << 31) - 1


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def abcd():
    """2^31 + 1 MiB + 123 bytes over {a,b,c,d} (seeded) with b"abcd" planted
    across 2^31 - 1 and at 2^31 + 10."""
    n = B31 + (1 << 20) + 123
    rng = np.random.default_rng(2031)
    arr = np.frombuffer(rng.bytes(n), dtype=np.uint8) & np.uint8(3)
    arr = arr + np.uint8(ord("a"))
    for p in (B31 - 2, B31 + 10, n - 4, 0):
        arr[p:p + 4] = np.frombuffer(b"abcd", dtype=np.uint8)
    src = arr.tobytes()
    del arr
    say(f"abcd input: {n} bytes, {src.count(b'abcd')} occurrences")
    return src


def int_query(lib, h, limit):
    k = C.c_int(0)
    last = C.c_int(-1)
    rc = lib.lib.cuda_replace_query_prefix(h, limit, C.byref(k), C.byref(last))
    return rc, k.value, last.value


def expect_query(src, pat, floor, limit):
    """Border-free pattern: the kept starts are the occurrences at or after floor."""
    lim = min(limit, len(src))
    if lim <= floor:
        return 0, -1
    end = lim + len(pat) - 1
    return src.count(pat, floor, end), src.rfind(pat, floor, end)


@pytest.mark.parametrize("rep", [b"0123456789AB", b"Z"], ids=["expand", "shrink"])
def test_border_free_past_2gib(lib, abcd, rep):
    src, pat = abcd, b"abcd"
    say("border-free", len(rep))
    with Session(lib, src) as s:
        n_out = s.apply(pat, rep)
        want = src.replace(pat, rep)
        assert n_out == len(want)
        assert s.result() == want
        del want
        for limit in (B31 - 100, B31 + 1000, len(src) + 5):
            assert s.query_prefix(limit) == expect_query(src, pat, 0, limit), limit
        rc, k, last = int_query(lib, s.h, B31 - 100)
        assert rc == 0 and (k, last) == expect_query(src, pat, 0, B31 - 100)
        # a kept start at 2^31 + 10 does not fit the int ABI
        assert int_query(lib, s.h, B31 + 1000)[0] == -4


def test_seeded_past_2gib(lib, abcd):
    src, pat, rep = abcd, b"abcd", b"0123456789AB"
    seed = B31 + 5
    floor = seed + len(pat)
    say("seeded past 2^31")
    with Session(lib, src) as s:
        n_out, last = s.apply_seeded(pat, rep, seed)
        want = src[:floor] + src[floor:].replace(pat, rep)
        assert n_out == len(want) and last == src.rfind(pat, floor)
        assert s.result() == want
        del want
        for limit in (1 << 30, B31 + 9, B31 + 11, B31 + 5000, len(src)):
            assert s.query_prefix(limit) == expect_query(src, pat, floor, limit), limit
        rc, k, lst = int_query(lib, s.h, 1 << 30)
        assert rc == 0 and (k, lst) == (0, -1)
        assert int_query(lib, s.h, B31 + 5000)[0] == -4


def test_carry_tables_past_2gib(lib):
    n = B31 + 3
    say("all-a", n)
    with Session(lib, b"a" * n) as s:
        kept = n // 2
        assert s.apply(b"aa", b"X") == n - kept
        assert s.result() == b"X" * kept + b"a" * (n % 2)
        assert s.query_prefix(n) == (kept, 2 * (kept - 1))
        assert s.query_prefix(B31 - 1) == (B31 // 2, B31 - 2)
        assert int_query(lib, s.h, n)[0] == -4            # last kept start 2^31
        rc, k, last = int_query(lib, s.h, 1000)
        assert rc == 0 and (k, last) == (500, 998)


def test_kept_count_past_int_max(lib):
    """2^32 + 6 bytes of 'a': 2^31 + 3 kept starts, so even the count overflows int."""
    n = (1 << 32) + 6
    say("all-a", n)
    with Session(lib, b"a" * n) as s:
        kept = n // 2
        assert kept > INT_MAX
        assert s.apply(b"aa", b"X") == kept
        assert s.query_prefix(n) == (kept, n - 2)
        assert s.query_prefix(B31 + 1) == (B31 // 2 + 1, B31)
        assert int_query(lib, s.h, n)[0] == -4
        out = s.result()
        assert len(out) == kept and out == b"X" * kept


ENTITIES = [(b"&ab;", b"1"), (b"&bcd;", b"22"), (b"&cab;", b"333"), (b"&dd;", b"4444"),
            (b"&abc;", b"5"), (b"&bad;", b"66"), (b"&cc;", b"777"), (b"&dab;", b"8888")]


def test_fused_apply_batch_past_2gib(lib):
    n = B31 + (1 << 20)
    rng = np.random.default_rng(4031)
    lut = np.frombuffer(b"abcd ", dtype=np.uint8)
    arr = lut[np.frombuffer(rng.bytes(n), dtype=np.uint8) % np.uint8(5)]
    # one entity every 200 bytes (0.5% of positions), one straddling 2^31
    pos = np.arange(7, n - 8, 200, dtype=np.int64)
    pos = np.append(pos, B31 - 3)
    which = rng.integers(0, len(ENTITIES), len(pos))
    for e, (pat, _) in enumerate(ENTITIES):
        at = pos[which == e]
        for j, byte in enumerate(pat):
            arr[at + j] = byte
    src = arr.tobytes()
    del arr
    want = src
    for pat, rep in ENTITIES:
        want = want.replace(pat, rep)
    say(f"fused batch input {n} bytes, output {len(want)}")
    assert lib.batch_fusable(ENTITIES)
    with Session(lib, src) as s:
        assert s.apply_batch(ENTITIES) == len(want)
        assert s.last_batch_fused
        assert s.result() == want
        s.reset(src)
        assert s.apply_batch(ENTITIES, fuse=False) == len(want)
        assert not s.last_batch_fused
        assert s.result() == want
