"""Group-size autotuner on the GPU: every group size (1, 2 or 4 tiles of 8 KiB per
workgroup) must give the oracle's output, around the tile and group boundaries of
each size, for carry tables of 2/4/8/16 states (border-free patterns run the lean
pass, whose group is fixed at 16 KiB, whatever size is pinned); and the first pass
in a tuned size bucket must record a choice with a measured time for every
candidate (the default size unless another is more than 3% faster), and still
return the exact output."""
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


@pytest.mark.parametrize("gt", [1, 2, 4])
def test_every_group_size_exact(lib, pinned, gt):
    rng = random.Random(60 + gt)
    for pat, rep in PATS:
        lib.autotune_set(pat, 0, gt)           # bucket of small inputs
        for n in SIZES:
            src = _src(rng, n)
            assert lib.unified(src, pat, rep) == src.replace(pat, rep), (gt, pat, n)
        src = b"a" * 70_001
        assert lib.unified(src, pat, rep) == src.replace(pat, rep), (gt, pat, "all-a")


@pytest.mark.parametrize("gt", [1, 2, 4])
def test_group_size_seeded_and_query(lib, pinned, gt):
    rng = random.Random(70 + gt)
    src = _src(rng, 50_000)
    for pat, rep in PATS[:6]:
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
    assert set(classes) == {"table2"}      # border-free passes are not tuned
    for e in rep_:
        assert e["size"] == "4-32MiB" and e["group_tiles"] in (1, 2, 4)
        ms = e["ms"]
        assert all(ms[k] > 0 for k in ("1", "2", "4"))
        best = min(ms.values())
        if e["group_tiles"] == 2:   # the default stays unless another size is >3% faster
            assert best >= 0.97 * ms["2"] or ms["2"] == best
        else:
            assert ms[str(e["group_tiles"])] == best and best < 0.97 * ms["2"]
