"""Autotuner on the GPU.  Every carry-table group size (1, 2 or 4 tiles of 8 KiB per
workgroup) and every form of a border-free pass (the 32-byte-per-thread count and
emit with 1, 2 or 4 tiles per group, the 64-byte three-launch pass, the single
launch re-reading or holding each group) must give the oracle's output around the tile and group boundaries of each
size, plain and seeded with prefix queries; and the first pass in a tuned size
bucket must record a choice with a measured time for every candidate (the default
unless another is more than 3% faster), and still return the exact output."""
import random

import pytest

from oracle import replace_oracle as ro
from cuda_replace_amd import Session

pytestmark = pytest.mark.gpu

PATS = [(b"ab", b"X"), (b"b\x00a", b"QRSTUVWXYZ"), (b"aa", b"Y"), (b"abab", b"ZZZZZZ"), (b"aabaa", b"P"),
        (b"a" * 12, b"W"), (b"aba", b"")]
SIZES = [0, 1, 8191, 8192, 8193, 16383, 16384, 16385, 32767, 32768, 32769, 65541, 200_000]


def _src(rng, n):
    return bytes(rng.choice(b"aab\x00") for _ in range(n))


@pytest.fixture
def pinned(lib):
    yield
    lib.autotune_reset()


def _pinnable(lib, pat, gt):
    """1/2/4 pin any tuned class; 8, 16 and 32 (the 64-byte forms) only border-free ones."""
    return gt in (1, 2, 4) or lib.pattern_mode(pat) == 1


@pytest.mark.parametrize("gt", [1, 2, 4, 8, 16, 32])
def test_every_group_size_exact(lib, pinned, gt):
    rng = random.Random(60 + gt)
    for pat, rep in PATS:
        if not _pinnable(lib, pat, gt):
            continue
        lib.autotune_set(pat, 0, gt)           # bucket of small inputs
        for n in SIZES:
            src = _src(rng, n)
            assert lib.unified(src, pat, rep) == src.replace(pat, rep), (gt, pat, n)
        src = b"a" * 70_001
        assert lib.unified(src, pat, rep) == src.replace(pat, rep), (gt, pat, "all-a")


@pytest.mark.parametrize("gt", [1, 2, 4, 8, 16, 32])
def test_group_size_seeded_and_query(lib, pinned, gt):
    rng = random.Random(70 + gt)
    src = _src(rng, 50_000)
    for pat, rep in PATS[:6]:
        if not _pinnable(lib, pat, gt):
            continue
        lib.autotune_set(pat, 0, gt)
        for seed in (-5, 0, 8190, 16383, 32770):
            with Session(lib, src) as s:
                n2, last = s.apply_seeded(pat, rep, seed)
                want, want_last = ro.apply_seeded(src, pat, rep, seed)
                assert s.result() == want and n2 == len(want) and last == want_last, (gt, pat, seed)
                for lim in (0, 8192, 16384, 32768, 49_999):
                    assert s.query_prefix(lim) == ro.query_prefix(src, pat, seed, lim)


def test_first_tuned_pass_measures_every_candidate(lib, pinned):
    lib.autotune_reset()
    rng = random.Random(80)
    src = bytes(rng.choice(b"abcd") for _ in range(1 << 16)) * 96     # 6 MiB
    for pat, rep in [(b"ab", b"X"), (b"aa", b"Y")]:
        want = src.replace(pat, rep)
        assert lib.unified(src, pat, rep) == want      # tunes
        assert lib.unified(src, pat, rep) == want      # uses the choice
    rep_ = lib.autotune_report()
    classes = {e["class"]: e for e in rep_}
    assert set(classes) == {"border_free", "table2"}
    t = classes["table2"]
    assert t["size"] == "4-32MiB" and t["group_tiles"] in (1, 2, 4)
    ms = t["ms"]
    assert all(ms[k] > 0 for k in ("1", "2", "4"))
    best = min(ms[k] for k in ("1", "2", "4"))
    if t["group_tiles"] == 2:   # the default stays unless another size is >3% faster
        assert best >= 0.97 * ms["2"] or ms["2"] == best
    else:
        assert ms[str(t["group_tiles"])] == best and best < 0.97 * ms["2"]
    # 2-state tables also time their single launch; under 64 MiB the three launches
    # stay unless the single launch is more than 3% faster (-1: it gave up)
    sm = ms["single_launch_lds"]
    assert t["form"] in ("single_launch_lds", "three_launch") and (sm > 0 or sm == -1)
    assert (t["form"] == "single_launch_lds") == (sm > 0 and sm < 0.97 * best)
    b = classes["border_free"]
    forms = ("single_launch", "three_launch_64B", "tiles1_32B", "tiles2_32B", "tiles4_32B", "single_launch_lds")
    assert b["size"] == "4-32MiB" and b["form"] in forms
    ms = b["ms"]
    assert all(ms[k] > 0 or (k.startswith("single_launch") and ms[k] == -1) for k in forms)   # -1: gave up
    timed = {k: v for k, v in ms.items() if v > 0}
    dflt = "three_launch_64B"          # the default under 64 MiB
    if b["form"] == dflt:
        assert min(timed.values()) >= 0.97 * ms[dflt] or ms[dflt] == min(timed.values())
    else:
        assert ms[b["form"]] == min(timed.values()) and ms[b["form"]] < 0.97 * ms[dflt]
