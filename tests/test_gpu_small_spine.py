"""Carry-table passes of at most 1024 groups (gpu): the single-launch small spine.

A self-overlapping pattern of up to 16 bytes runs with carry tables of 2, 4, 8 or 16
states; a pass of at most 1024 groups (16384 for 2- and 4-state tables, 4096 for
8 and 16 states, with several groups per lane) resolves the groups' states in one
launch that scans the group tables wave by wave (cuda_replace.cu:905-968 is the serial
greedy these tables replace).  Inputs are biased {a, b} so long runs of 'a' make
the carry between groups matter; sizes put 1, 2, 63, 64, 65, 128, 1000 and 1024
groups in a pass, with 1025 groups (the blocked spine) beside them.  Every output,
seeded apply and prefix query is checked against the oracle.
"""
import numpy as np
import pytest

from oracle import replace_oracle as ro
from cuda_replace_amd import Session

pytestmark = pytest.mark.gpu

GROUP = 16384   # two 8 KiB tiles: the group size of passes under 4 MiB
GROUPS = [1, 2, 63, 64, 65, 128, 1000, 1024, 1025]
PATS = [(b"aa", b"X"), (b"aaa", b""), (b"abab", b"Q"), (b"a" * 7, b"0123456789"), (b"a" * 8, b"z"),
        (b"a" * 12, b""), (b"a" * 16, b"SIXTEEN"), (b"aabaa", b"@")]


def say(*a):
    print(*a, flush=True)


def biased(n, seed):
    rng = np.random.default_rng(seed)
    return (np.where(rng.random(n) < 0.93, 97, 98).astype(np.uint8)).tobytes()


@pytest.mark.parametrize("pat,rep", PATS, ids=[f"L{len(p)}" for p, _ in PATS])
def test_small_spine_every_group_count(lib, pat, rep):
    assert lib.pattern_mode(pat) == 2, pat
    for g in GROUPS:
        n = g * GROUP - (0 if g % 2 else 777)
        src = biased(n, g * 31 + len(pat))
        say("small spine", pat[:8], len(pat), g, n)
        assert lib.unified(src, pat, rep) == src.replace(pat, rep), (pat, g)


@pytest.mark.parametrize("pat,rep", [(b"aa", b"X"), (b"abab", b"Q"), (b"a" * 12, b"")], ids=["L2", "L4", "L12"])
def test_small_spine_several_groups_per_lane(lib, pat, rep):
    """Passes of 1025 to 16384 groups: every lane of the single spine launch composes
    several consecutive groups (up to 16 for 2- and 4-state tables, 4 for 8 and 16
    states; the blocked spine beyond)."""
    for g in (1025, 2047, 2048, 4095, 4096, 4097, 16383, 16385):
        if len(pat) > 4 and g > 4097:
            continue
        n = g * GROUP - (3 if g % 2 else 0)
        src = biased(n, g + len(pat))
        say("small spine v>1", len(pat), g, n)
        assert lib.unified(src, pat, rep) == src.replace(pat, rep), (pat, g)


def test_small_spine_seeded_and_queries(lib):
    src = biased(70 * GROUP + 5, 4)
    n = len(src)
    for pat, rep in [(b"aa", b"X"), (b"a" * 16, b"#"), (b"abab", b"")]:
        for seed in (-3, 0, 16383, 16384, 500001, n - 3):
            with Session(lib, src) as s:
                n2, last = s.apply_seeded(pat, rep, seed)
                want, want_last = ro.apply_seeded(src, pat, rep, seed)
                assert s.result() == want and n2 == len(want) and last == want_last, (pat, seed)
                for lim in (0, 16384, 16385, 600000, n):
                    assert s.query_prefix(lim) == ro.query_prefix(src, pat, seed, lim), (pat, seed, lim)


def test_small_spine_all_a_closed_form(lib):
    """All-'a' over 1024 groups: every second position kept by aa -> X."""
    n = 1024 * GROUP - 1
    src = b"a" * n
    out = lib.unified(src, b"aa", b"X")
    assert out == b"X" * (n // 2) + b"a"
    out16 = lib.unified(src, b"a" * 16, b"Y")
    assert out16 == b"Y" * (n // 16) + b"a" * (n % 16)
