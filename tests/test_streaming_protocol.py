"""The host-side chunking protocol of gpu_replace_streaming, checked on the CPU.

The wrapper's Session is swapped for a stand-in that answers each call with the
oracle, so the chunk/overlap/seed/cut logic of the wrapper is exercised without a
GPU and compared with bytes.replace of the whole input."""
import random

import pytest

import cuda_replace_amd.wrapper as W
from oracle import replace_oracle as ro
from tests.cases import STREAMING_REGRESSIONS


class OracleSession:
    def __init__(self, lib, src):
        self.buf = src
        self.last_apply = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def reset(self, src):
        self.buf = src
        self.last_apply = None

    def apply(self, pat, rep):
        self.last_apply = (self.buf, pat, None)
        self.buf = ro.apply_seeded(self.buf, pat, rep)[0]
        return len(self.buf)

    def apply_seeded(self, pat, rep, seed):
        self.last_apply = (self.buf, pat, seed)
        self.buf, last = ro.apply_seeded(self.buf, pat, rep, seed)
        return len(self.buf), last

    def query_prefix(self, limit):
        src, pat, seed = self.last_apply
        return ro.query_prefix(src, pat, seed, limit)

    def result(self):
        return self.buf

    def close(self):
        pass


@pytest.fixture
def oracle_session(monkeypatch):
    monkeypatch.setattr(W, "Session", OracleSession)


@pytest.mark.parametrize("case", STREAMING_REGRESSIONS, ids=[c[0] for c in STREAMING_REGRESSIONS])
def test_straddle_regressions(oracle_session, case):
    name, src, pat, rep, chunk = case
    assert W.gpu_replace_streaming(None, src, [(pat, rep)], chunk) == src.replace(pat, rep)


def test_protocol_random(oracle_session):
    rng = random.Random(5)
    alpha = b"ab\x00"
    for _ in range(300):
        n = rng.randint(0, 30)
        src = bytes(rng.choice(alpha) for _ in range(n))
        pairs = [(bytes(rng.choice(alpha) for _ in range(rng.randint(0, 4))),
                  bytes(rng.choice(b"XYZ") for _ in range(rng.randint(0, 4))))
                 for _ in range(rng.randint(1, 3))]
        want = ro.apply_batch(src, pairs)
        for chunk in (1, 2, 3, 4, 7, 100):
            assert W.gpu_replace_streaming(None, src, pairs, chunk) == want, (src, pairs, chunk)


def test_protocol_text(oracle_session):
    text = (b"hello world &amp; hello &lt;x&gt; aaaa " * 200)
    pairs = [(b"&amp;", b"&"), (b"hello", b"hi"), (b"aa", b"b"), (b"&lt;", b"<")]
    want = ro.apply_batch(text, pairs)
    for chunk in (1, 5, 64, 1000, 1 << 20):
        assert W.gpu_replace_streaming(None, text, pairs, chunk) == want
