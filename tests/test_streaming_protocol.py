"""The chunk protocol of the library's streaming pipeline (cr_stream_replace),
restated step for step on the CPU and checked against bytes.replace of the whole
input.

Per pair, chunk k covers source positions [kC, (k+1)C) plus an (L-1)-byte overlap
at fixed offsets, so uploads never wait for results.  Its floor (the seed) is the
last kept start so far plus L.  Its output starts after the bytes covered by a
match from an earlier chunk (dropped) and ends at cut = max((k+1)C, last kept + L),
with the kept count below C taken from the prefix query.  The last chunk keeps
its whole output.  An empty pattern inserts once at each chunk seam.  The GPU tests
(tests/test_gpu_stream.py) run the same protocol in the library.
"""
import random

import pytest

from oracle import replace_oracle as ro
from tests.cases import STREAMING_REGRESSIONS


def stream_pair_model(x: bytes, pat: bytes, rep: bytes, C: int) -> bytes:
    n, L, r = len(x), len(pat), len(rep)
    if n == 0:
        return rep if L == 0 else b""
    H = L - 1 if L else 0
    K = (n + C - 1) // C
    out, last_global, cut_prev = [], None, 0
    for k in range(K):
        start = k * C
        chunk = x[start:min(n, start + C + H)]
        last_chunk = k == K - 1
        floor = 0 if last_global is None else max(0, last_global + L - start)
        seed = floor - L                       # apply_seeded's carry for that floor
        res, _ = ro.apply_seeded(chunk, pat, rep, seed if floor > 0 else None)
        drop = max(0, cut_prev - start)
        if L == 0:
            end = len(res) if last_chunk else len(chunk) * (r + 1)
            cut = start + len(chunk)
        elif last_chunk:
            end, cut = len(res), n
        else:
            kept_before, last_rel = ro.query_prefix(chunk, pat, seed if floor > 0 else None, C)
            if last_rel >= 0:
                last_global = start + last_rel
            cut = start + C
            if last_global is not None:
                cut = max(cut, last_global + L)
            end = (cut - start) + kept_before * (r - L)
        out.append(res[drop:end] if end > drop else b"")
        cut_prev = cut
    return b"".join(out)


def stream_model(src: bytes, pairs, C: int) -> bytes:
    for pat, rep in pairs:
        src = stream_pair_model(src, pat, rep, C)
    return src


@pytest.mark.parametrize("case", STREAMING_REGRESSIONS, ids=[c[0] for c in STREAMING_REGRESSIONS])
def test_straddle_regressions(case):
    name, src, pat, rep, chunk = case
    assert stream_model(src, [(pat, rep)], chunk) == src.replace(pat, rep)


def test_protocol_random():
    rng = random.Random(5)
    alpha = b"ab\x00"
    for _ in range(300):
        n = rng.randint(0, 30)
        src = bytes(rng.choice(alpha) for _ in range(n))
        pairs = [(bytes(rng.choice(alpha) for _ in range(rng.randint(0, 5))),
                  bytes(rng.choice(b"XYZ") for _ in range(rng.randint(0, 4))))
                 for _ in range(rng.randint(1, 3))]
        want = ro.apply_batch(src, pairs)
        for chunk in (1, 2, 3, 4, 7, 100):
            assert stream_model(src, pairs, chunk) == want, (src, pairs, chunk)


def test_protocol_pattern_longer_than_chunk():
    # one kept match covers several chunk ends: the cut must come from an earlier chunk
    for n in range(0, 40):
        src = b"a" * n
        for L in (3, 5, 9):
            for chunk in (1, 2, 4):
                assert stream_model(src, [(b"a" * L, b"Z")], chunk) == src.replace(b"a" * L, b"Z")


def test_protocol_text():
    text = (b"hello world &amp; hello &lt;x&gt; aaaa " * 200)
    pairs = [(b"&amp;", b"&"), (b"hello", b"hi"), (b"aa", b"b"), (b"&lt;", b"<"), (b"", b"|")]
    want = ro.apply_batch(text, pairs)
    for chunk in (1, 5, 64, 1000, 1 << 20):
        assert stream_model(text, pairs, chunk) == want


def test_planted_cut_mistake_is_caught():
    # cutting at the chunk end (the reference wrapper's rule) must fail the checks
    def bad(x, pat, rep, C):
        out, pos = [], 0
        while pos < len(x):
            out.append(x[pos:pos + C].replace(pat, rep))
            pos += C
        return b"".join(out)
    src = b"aaaa"
    assert bad(src, b"aa", b"X", 3) != src.replace(b"aa", b"X")
