"""Failed device allocations (gpu): the session keeps its data and stays usable.

When growing a buffer fails, the reference keeps the old buffers alive and the
session's data as it was (cuda_replace.cu:1030-1066, 1142-1220).  The test hook
CR_FAIL_GROW_OVER=bytes, read at every buffer growth, makes each growth above that
size fail as an out-of-memory hipMalloc would.  Every case here forces such a
failure in a call that needs a larger buffer (an expanding apply, a fused
apply_batch, a reset or open with a larger input, a one-shot call), checks that the
call reports the error and that the session's bytes and length are unchanged, then
lifts the limit and checks that the same session works and matches the oracle.
"""
import os

import numpy as np
import pytest

from oracle import replace_oracle as ro
from cuda_replace_amd import Session

pytestmark = pytest.mark.gpu

KEY = "CR_FAIL_GROW_OVER"


@pytest.fixture
def limit():
    saved = os.environ.get(KEY)

    def set_limit(v):
        if v is None:
            os.environ.pop(KEY, None)
        else:
            os.environ[KEY] = str(int(v))

    yield set_limit
    if saved is None:
        os.environ.pop(KEY, None)
    else:
        os.environ[KEY] = saved


def text(n, seed):
    rng = np.random.default_rng(seed)
    return np.frombuffer(b"abcd #", dtype=np.uint8)[rng.integers(0, 6, size=n)].tobytes()


def test_expanding_apply(lib, limit):
    src = text(1 << 20, 1)
    with Session(lib, src) as s:
        s.apply(b"ab", b"X")
        before = s.result()
        limit(2 << 20)
        with pytest.raises(RuntimeError):
            s.apply(b"a", b"0123456789" * 3)   # its output buffer would need about 30 MiB
        assert s.length() == len(before)
        assert s.result() == before
        limit(None)
        n2 = s.apply(b"a", b"0123")
        want = before.replace(b"a", b"0123")
        assert n2 == len(want) and s.result() == want
        assert s.query_prefix(1000) == ro.query_prefix(before, b"a", None, 1000)


def test_expanding_fused_apply_batch(lib, limit):
    src = text(1 << 20, 2)
    pairs = [(b"#ab", b"0" * 20), (b"#cd", b"1" * 20)]
    assert lib.batch_fusable(pairs)
    with Session(lib, src) as s:
        s.apply(b"dd", b"")
        before = s.result()
        limit(2 << 20)
        with pytest.raises(RuntimeError):
            s.apply_batch(pairs)
        assert s.length() == len(before)
        assert s.result() == before
        limit(None)
        s.apply_batch(pairs)
        assert s.last_batch_fused
        assert s.result() == ro.apply_batch(before, pairs)


def test_reset_and_open_with_larger_input(lib, limit):
    small, big = text(64 << 10, 3), text(8 << 20, 4)
    with Session(lib, small) as s:
        limit(1 << 20)
        with pytest.raises(RuntimeError):
            s.reset(big)
        assert s.length() == len(small)
        assert s.result() == small
        s.apply(b"ab", b"Y")                       # still usable under the limit
        assert s.result() == small.replace(b"ab", b"Y")
        with pytest.raises(RuntimeError):
            Session(lib, big)
        limit(None)
        s.reset(big)
        s.apply(b"ab", b"Y")
        assert s.result() == big.replace(b"ab", b"Y")


def test_one_shot_call(lib, limit):
    src = text(4 << 20, 5)
    limit(1 << 20)
    with pytest.raises(RuntimeError):
        lib.unified(src, b"ab", b"XYZ")
    limit(None)
    assert lib.unified(src, b"ab", b"XYZ") == src.replace(b"ab", b"XYZ")
