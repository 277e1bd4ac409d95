"""The bitmap-mode passes (gpu): patterns the carry tables and the LDS halo cannot take.

A self-overlapping pattern of 17-256 bytes, or any pattern over 256 bytes, runs in
bitmap mode: candidates are marked, then the leftmost non-overlapping rule
(cuda_replace.cu:905-968, the serial greedy) picks the kept starts.  The library
selects them in parallel: per-tile tables of (carry out, kept) for every entry
state, composed 64 at a time level by level, walked from the top and expanded back
to the tiles, then every tile resolved; border-free long patterns keep every
candidate.  Sizes give 1, 2, 64, 65 and 4097 tiles (one, two and three table
levels); inputs are all-'a', periodic and biased random ones where chains of
overlapping occurrences cross tile and level boundaries.  Every output, seeded
apply and prefix query is checked against the oracle and against the one-wave
serial greedy (CR_BITMAP_SERIAL=1) on the same input.
"""
import os

import numpy as np
import pytest

from oracle import replace_oracle as ro
from cuda_replace_amd import Session

pytestmark = pytest.mark.gpu

TILE = 8192


def say(*a):
    print(*a, flush=True)


@pytest.fixture
def serial_env():
    saved = os.environ.get("CR_BITMAP_SERIAL")
    yield os.environ
    if saved is None:
        os.environ.pop("CR_BITMAP_SERIAL", None)
    else:
        os.environ["CR_BITMAP_SERIAL"] = saved


def biased(n, seed, p_a=0.97):
    rng = np.random.default_rng(seed)
    return (np.where(rng.random(n) < p_a, 97, 98).astype(np.uint8)).tobytes()


SELF = [(b"a" * 17, b"X"), (b"ab" * 9 + b"a", b""), (b"a" * 40, b"0123456789" * 5), (b"abc" * 20 + b"ab", b"#"),
        (b"a" * 256, b"Z"), (b"a" * 100 + b"b" + b"a" * 100, b"mid")]


@pytest.mark.parametrize("tiles", [1, 2, 64, 65, 4097])
def test_self_overlapping_long_patterns(lib, tiles):
    for pat, rep in SELF:
        assert lib.pattern_mode(pat) == 3, pat
        n = tiles * TILE - (5 if tiles > 1 else 100)
        inputs = [b"a" * n, biased(n, tiles + len(pat))]
        if pat.startswith(b"ab"):
            inputs.append((b"ab" * (n // 2 + 1))[:n])
        if pat.startswith(b"abc"):
            inputs.append((b"abc" * (n // 3 + 1))[:n])
        for src in inputs:
            say("bitmap self", len(pat), tiles, src[:3])
            assert lib.unified(src, pat, rep) == src.replace(pat, rep), (len(pat), tiles)


def test_border_free_long_patterns(lib):
    rng = np.random.default_rng(11)
    words = [bytes(rng.integers(97, 123, size=int(rng.integers(2, 9))).astype(np.uint8)) for _ in range(500)]
    text = b" ".join(words[int(i)] for i in rng.integers(0, 500, size=400000))
    for L in (257, 300, 1000, 9000, 20000):
        pat = text[123456:123456 + L]
        if not pat or lib.pattern_mode(pat) != 3:
            continue
        # plant the pattern around tile boundaries, back to back, and long enough
        # to span several tiles
        b = bytearray(text)
        for pos in (0, TILE - 7, 5 * TILE - L // 2, 40000, 40000 + L, len(b) - L):
            if 0 <= pos <= len(b) - L:
                b[pos:pos + L] = pat
        src = bytes(b)
        for rep in (b"", b"<r>", b"R" * 700):
            say("bitmap border-free", L, len(rep))
            assert lib.unified(src, pat, rep) == src.replace(pat, rep), (L, len(rep))


def test_bitmap_seeded_and_queries(lib):
    src = biased(65 * TILE + 3, 3, 0.99)
    n = len(src)
    for pat, rep in [(b"a" * 17, b"Q"), (b"a" * 130, b""), (src[70000:70400], b"!")]:
        for seed in (-20, 0, 17, TILE - 1, TILE, 300001, n - 17):
            with Session(lib, src) as s:
                n2, last = s.apply_seeded(pat, rep, seed)
                want, want_last = ro.apply_seeded(src, pat, rep, seed)
                assert s.result() == want and n2 == len(want) and last == want_last, (len(pat), seed)
                for lim in (0, TILE, TILE + 1, 300000, n):
                    assert s.query_prefix(lim) == ro.query_prefix(src, pat, seed, lim), (len(pat), seed, lim)


def test_parallel_equals_serial_greedy(lib, serial_env):
    """The same passes with the one-wave serial greedy give the same bytes."""
    src = biased(70 * TILE + 11, 9, 0.95)
    for pat, rep in [(b"a" * 19, b"x"), (b"ab" * 9 + b"a", b"YY"), (src[1000:1300], b"")]:
        serial_env.pop("CR_BITMAP_SERIAL", None)
        par = lib.unified(src, pat, rep)
        serial_env["CR_BITMAP_SERIAL"] = "1"
        ser = lib.unified(src, pat, rep)
        assert par == ser == src.replace(pat, rep), len(pat)


def test_last_kept_from_the_last_tile_with_kept_starts(lib):
    """The last kept start is reported by the one tile after which nothing is kept:
    occurrences only in early tiles, only in the final tile, and none at all."""
    rng = np.random.default_rng(5)
    base = bytes(rng.integers(98, 123, size=40 * TILE + 9).astype(np.uint8))   # no 'a'
    bf = bytes(rng.integers(97, 123, size=300).astype(np.uint8)) + b"!"
    for pat in (b"a" * 17, b"a" * 40, bf):
        L = len(pat)
        for where in ([0], [3 * TILE - 5, 7 * TILE + 1], [len(base) - L], [5, len(base) - L - 3], []):
            b = bytearray(base)
            for pos in where:
                b[pos:pos + L] = pat
            src = bytes(b)
            for rep in (b"", b"R" * 33):
                with Session(lib, src) as s:
                    n2, last = s.apply_seeded(pat, rep, 0)
                    want, want_last = ro.apply_seeded(src, pat, rep, 0)
                    assert s.result() == want and n2 == len(want) and last == want_last, (L, where, len(rep))
