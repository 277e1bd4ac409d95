"""The single-pass border-free sweep (gpu), forced onto small inputs.

By default the library runs a border-free pass as one persistent launch (count
phases `ahead` rounds before the emit phases, a scanner workgroup publishing each
group's prefix) only from 64 MiB on; CR_SWEEP_MIN=0 makes every border-free pass a
sweep.  Every test runs both single-launch forms (CR_BF_VARIANT=0 re-stages a group
for its emit, 5 holds it in LDS one round from its count to its emit), so the edge
cases (one group, fewer groups than workers, tile and group
boundaries, seeds, every pattern length class of the count's masked SADs) run here.
CR_SWEEP_AHEAD varies the count-ahead distance, and CR_SWEEP_SPIN=1 makes the
bounded waits give up at once, which must take the host's fallback to the three
launches and still give the oracle's output."""
import os
import random

import pytest

from oracle import replace_oracle as ro
from cuda_replace_amd import Session

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(params=["0", "5"], ids=["restage", "hold"])
def sweep_env(request):
    saved = {k: os.environ.get(k) for k in ("CR_SWEEP", "CR_SWEEP_MIN", "CR_SWEEP_AHEAD", "CR_SWEEP_SPIN",
                                             "CR_BF_VARIANT")}
    os.environ["CR_SWEEP_MIN"] = "0"
    os.environ["CR_BF_VARIANT"] = request.param
    os.environ.pop("CR_SWEEP", None)
    yield os.environ
    for k, v in saved.items():
        if v is None:
            os.environ.pop(k, None)
        else:
            os.environ[k] = v


def text(n, seed, alphabet=b"abcd \x00"):
    rng = random.Random(seed)
    return bytes(rng.choice(alphabet) for _ in range(n))


# border-free patterns of every masked-SAD word count (1..4) and beyond 16 bytes
PATS = [(b"a", b"XY"), (b"ab", b""), (b"abc", b"Q"), (b"abcd", b"0123456789AB"), (b"a\x00b", b"Z"),
        (b"abcdb", b"#"), (b"dcba ab", b"%%"), (b"abcd abcd b\x00c", b""), (b"bcd abcd dcbaa c", b"LONG" * 5),
        (b"abcdd bbb ccc ddd\x00", b"x")]
SIZES = [1, 2, 63, 64, 65, 8191, 16383, 16384, 16385, 40000, 100001]


@pytest.mark.parametrize("ahead", ["0", "1", "2", "5"])
def test_sweep_every_size(lib, sweep_env, ahead):
    sweep_env["CR_SWEEP_AHEAD"] = ahead
    rng = random.Random(int(ahead))
    bases = {n: text(n, rng.randrange(1 << 30)) for n in SIZES}
    for pat, rep in PATS:
        assert lib.pattern_mode(pat) == 1, pat
        for n in SIZES:
            src = bases[n]
            # plant occurrences around the group and tile boundaries
            b = bytearray(src)
            for pos in (0, 8190, 16380, 16383, 32766, n - len(pat)):
                if 0 <= pos <= n - len(pat):
                    b[pos:pos + len(pat)] = pat
            src = bytes(b)
            assert lib.unified(src, pat, rep) == src.replace(pat, rep), (ahead, pat, n)


def test_sweep_many_groups(lib, sweep_env):
    """More groups than resident workers (several rounds per worker)."""
    rng = random.Random(5)
    base = text(1 << 16, 6)
    src = base * 300                                   # 19.7 MiB: 1200 groups
    for pat, rep in [(b"abcd", b"0123456789AB"), (b"ab", b"_"), (b"a\x00", b"")]:
        for ahead in ("1", "3"):
            sweep_env["CR_SWEEP_AHEAD"] = ahead
            say("many groups", pat, ahead)
            assert lib.unified(src, pat, rep) == src.replace(pat, rep), (pat, ahead)


def test_sweep_seeded_and_query(lib, sweep_env):
    src = text(200000, 7)
    for pat, rep in [(b"ab", b"XYZ"), (b"abcd", b""), (b"dcba ab", b"1")]:
        for seed in (-5, 0, 16380, 16390, 120000, 199999):
            with Session(lib, src) as s:
                n2, last = s.apply_seeded(pat, rep, seed)
                want, want_last = ro.apply_seeded(src, pat, rep, seed)
                assert s.result() == want and n2 == len(want) and last == want_last, (pat, seed)
                for lim in (0, 8192, 16384, 16385, 100000, 200000):
                    assert s.query_prefix(lim) == ro.query_prefix(src, pat, seed, lim), (pat, seed, lim)


def test_sweep_gives_up_and_falls_back(lib, sweep_env):
    """A spin limit of one poll: the waits give up, every workgroup returns, and the
    host's three-launch pass produces the output."""
    sweep_env["CR_SWEEP_SPIN"] = "1"
    src = text(3 << 20, 8)
    for pat, rep in [(b"abcd", b"0123456789AB"), (b"ab", b"")]:
        for _ in range(3):
            assert lib.unified(src, pat, rep) == src.replace(pat, rep), pat
    with Session(lib, src) as s:
        s.apply(b"abc", b"Q")
        assert s.result() == src.replace(b"abc", b"Q")
        assert s.query_prefix(1 << 20) == ro.query_prefix(src, b"abc", None, 1 << 20)


def test_sweep_off_matches(lib, sweep_env):
    src = text(70000, 9)
    sweep_env["CR_SWEEP"] = "0"
    a = lib.unified(src, b"abcd", b"!")
    sweep_env.pop("CR_SWEEP")
    b = lib.unified(src, b"abcd", b"!")
    assert a == b == src.replace(b"abcd", b"!")


def test_concurrent_single_launches(lib, sweep_env):
    """Two sessions on two threads run 32 MiB single-launch passes at the same time.
    The grid of each is sized to the whole device, so when both are in flight neither
    may be resident as a whole; a wait then gives up after a few hundred polls (the
    grid's last workgroup has not started) and the pass takes the three launches.
    Every output must be exact."""
    import threading
    import time
    src = text(1 << 16, 11) * 512                      # 32 MiB: 2048 groups per pass
    pairs = [(b"abcd", b"0123456789AB"), (b"ab", b"_")]
    wants = [src.replace(p, r) for p, r in pairs]
    errors = []

    def work(i):
        pat, rep = pairs[i]
        try:
            with Session(lib, src) as s:
                for k in range(6):
                    s.reset(src)
                    s.apply(pat, rep)
                    if s.result() != wants[i]:
                        errors.append((i, k))
        except Exception as e:   # reported below
            errors.append(e)

    t0 = time.time()
    threads = [threading.Thread(target=work, args=(i,)) for i in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    say("concurrent single launches: %.2f s" % (time.time() - t0))
    assert not errors, errors
