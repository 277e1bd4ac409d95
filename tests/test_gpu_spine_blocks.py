"""Passes large enough for the spine to split into blocks (a per-block summary, a
scan of the block summaries, then every block's groups in parallel): the kept
count and the carry must cross block boundaries exactly, for border-free
patterns, carry tables of 2/4/8/16 states, seeded applies and the fused
multi-pattern pass.  A block holds 16384 groups (border-free, fused) or 1024
(carry tables); group size pinned to one tile (8 KiB) where that keeps the inputs
small, and a pass of exactly one block plus one group."""
import random

import pytest

from oracle import replace_oracle as ro
from cuda_replace_amd import Session

pytestmark = pytest.mark.gpu

MiB = 1 << 20


def _text(n, seed):
    rng = random.Random(seed)
    unit = bytes(rng.choice(b"aab") for _ in range(1 << 16))
    reps = n // len(unit) + 1
    return (unit * reps)[:n]


@pytest.fixture
def one_tile_groups(lib):
    yield
    lib.autotune_reset()


def test_border_free_three_blocks(lib, one_tile_groups):
    n = 300 * MiB                      # 38400 one-tile groups: three blocks
    src = _text(n, 1)
    lib.autotune_set(b"ab", n, 1)
    assert lib.unified(src, b"ab", b"XYZ") == src.replace(b"ab", b"XYZ")
    # for a border-free pattern every occurrence at or past the floor F = seed + L is
    # kept (the identity the oracle's apply_seeded is pinned to), so the expected
    # values come from bytes methods instead of a Python loop over 100M matches
    for seed in (-1, 5, 130 * MiB + 3):
        F = max(0, seed + 2)
        with Session(lib, src) as s:
            n2, last = s.apply_seeded(b"ab", b"", seed)
            want = src[:F] + src[F:].replace(b"ab", b"")
            assert n2 == len(want) and last == src.rfind(b"ab", F)
            assert s.result() == want
            lim = 200 * MiB
            assert s.query_prefix(lim) == (src.count(b"ab", F, lim + 1), src.rfind(b"ab", F, lim + 1))


@pytest.mark.parametrize("pat,size", [(b"aa", 140 * MiB), (b"aba", 70 * MiB), (b"aabaa", 40 * MiB),
                                      (b"a" * 12, 20 * MiB)])
def test_carry_tables_across_blocks(lib, one_tile_groups, pat, size):
    lib.autotune_set(pat, size, 1)     # 17920 / 8960 / 5120 / 2560 one-tile groups: 3-18 blocks
    for src in (b"a" * size, _text(size, len(pat))):
        assert lib.unified(src, pat, b"Q") == src.replace(pat, b"Q"), (pat, size)


def test_fused_pass_across_blocks(lib):
    n = 300 * MiB                      # 19200 two-tile groups: two fused spine blocks
    pairs = [(b"&" + bytes([97 + i]) + b"x;", bytes([65 + i])) for i in range(8)]
    rng = random.Random(3)
    base = bytearray(_text(n, 4))
    for k in range(0, n - 4, 997):
        p = pairs[rng.randrange(8)][0]
        base[k:k + len(p)] = p
    src = bytes(base)
    with Session(lib, src) as s:
        s.apply_batch(pairs)
        assert s.last_batch_fused
        assert s.result() == ro.apply_batch(src, pairs)


@pytest.mark.parametrize("pat,rep,blk", [(b"ab", b"XYZ", 16384), (b"aa", b"Q", 1024), (b"aabaa", b"", 1024)])
def test_one_block_and_one_group(lib, one_tile_groups, pat, rep, blk):
    # blk + 1 one-tile groups: a full spine block and a block of a single group (its
    # only tile a partial one), for border-free and carry-table patterns
    n = blk * 8192 + 100
    src = _text(n, 9)
    lib.autotune_set(pat, n, 1)
    assert lib.unified(src, pat, rep) == src.replace(pat, rep)
