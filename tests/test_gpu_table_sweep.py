"""The single launch of 2-state carry-table passes (gpu), forced onto small inputs.

A 2-byte pattern of one repeated byte (aa, two spaces, two NULs) is self-overlapping
with two carry states.  From 64 MiB its pass runs as one persistent launch in which
each workgroup counts a 16 KiB group (its carry table), a scanner composes the
tables in group order and the workgroup emits the group one round later from LDS
(the serial greedy it replaces: cuda_replace.cu:905-968).  CR_TABLE_SWEEP_MIN=0
makes every such pass take it, so group and tile boundaries, several rounds per
workgroup, seeds, prefix queries, expansions whose output outgrows the LDS image and
the give-up fallback (CR_SWEEP_SPIN=1) run here; every output is checked against the
oracle and against the three-launch pass (CR_TABLE_SWEEP=0) on the same input.
"""
import os

import numpy as np
import pytest

from oracle import replace_oracle as ro
from cuda_replace_amd import Session

pytestmark = pytest.mark.gpu

GROUP = 16384
PATS = [(b"aa", b"X"), (b"aa", b""), (b"aa", b"0123456789AB"), (b"  ", b" "), (b"\x00\x00", b"Z")]


def say(*a):
    print(*a, flush=True)


@pytest.fixture
def table_env():
    keys = ("CR_TABLE_SWEEP", "CR_TABLE_SWEEP_MIN", "CR_SWEEP_SPIN")
    saved = {k: os.environ.get(k) for k in keys}
    os.environ["CR_TABLE_SWEEP_MIN"] = "0"
    os.environ.pop("CR_TABLE_SWEEP", None)
    yield os.environ
    for k, v in saved.items():
        if v is None:
            os.environ.pop(k, None)
        else:
            os.environ[k] = v


def biased(n, seed, byte, p=0.9):
    """Runs of `byte` broken by other bytes (ab\\0 and space), so chains of overlapping
    occurrences cross segment, tile and group boundaries."""
    rng = np.random.default_rng(seed)
    other = np.frombuffer(b"b\x00c ", dtype=np.uint8)[rng.integers(0, 4, size=n)]
    return np.where(rng.random(n) < p, np.uint8(byte), other).astype(np.uint8).tobytes()


SIZES = [1, 2, 31, 32, 33, 63, 64, 65, 8191, 8192, 8193, 16383, 16384, 16385, 32769, 100001]


def test_every_size_and_pattern(lib, table_env):
    for pat, rep in PATS:
        assert lib.pattern_mode(pat) == 2, pat
        for n in SIZES:
            for src in (biased(n, n + len(rep), pat[0]), pat[:1] * n):
                want = src.replace(pat, rep)
                assert lib.unified(src, pat, rep) == want, (pat, rep, n, src[:4])


def test_several_rounds_per_workgroup(lib, table_env):
    """More groups than resident workgroups: 20 MiB (1280 groups) and 13 MiB + 5."""
    for n in (20 << 20, (13 << 20) + 5):
        src = biased(n, n, ord("a"), 0.93)
        for pat, rep in [(b"aa", b"X"), (b"aa", b"0123456789AB")]:
            say("rounds", n, len(rep))
            want = src.replace(pat, rep)
            assert lib.unified(src, pat, rep) == want, (n, rep)
        all_a = b"a" * n
        assert lib.unified(all_a, b"aa", b"X") == b"X" * (n // 2) + b"a" * (n % 2)


def test_equals_three_launch_pass(lib, table_env):
    src = biased((5 << 20) + 77, 5, ord("a"), 0.95)
    for pat, rep in [(b"aa", b"Q"), (b"aa", b"QRS")]:
        single = lib.unified(src, pat, rep)
        table_env["CR_TABLE_SWEEP"] = "0"
        three = lib.unified(src, pat, rep)
        table_env.pop("CR_TABLE_SWEEP")
        assert single == three == src.replace(pat, rep), rep


def test_seeded_and_queries(lib, table_env):
    src = biased(70 * GROUP + 5, 4, ord("a"), 0.97)
    n = len(src)
    for pat, rep in [(b"aa", b"X"), (b"aa", b"")]:
        for seed in (-3, 0, 1, GROUP - 1, GROUP, 500001, n - 3):
            with Session(lib, src) as s:
                n2, last = s.apply_seeded(pat, rep, seed)
                want, want_last = ro.apply_seeded(src, pat, rep, seed)
                assert s.result() == want and n2 == len(want) and last == want_last, (rep, seed)
                for lim in (0, 8192, GROUP, GROUP + 1, 600000, n):
                    assert s.query_prefix(lim) == ro.query_prefix(src, pat, seed, lim), (rep, seed, lim)


def test_gives_up_and_falls_back(lib, table_env):
    """A spin limit of one poll: the waits give up and the host's three-launch pass
    produces the output."""
    table_env["CR_SWEEP_SPIN"] = "1"
    src = biased(3 << 20, 8, ord("a"), 0.9)
    for rep in (b"X", b"0123"):
        for _ in range(3):
            assert lib.unified(src, b"aa", rep) == src.replace(b"aa", rep), rep
    with Session(lib, src) as s:
        s.apply(b"aa", b"Y")
        assert s.result() == src.replace(b"aa", b"Y")
        assert s.query_prefix(1 << 20) == ro.query_prefix(src, b"aa", None, 1 << 20)
