"""Device-driven batch rounds (cr_batch_run's default path) on the GPU.

Every round is planned, counted, placed and emitted by persistent kernels over
device-built work lists, and a run shape that repeats is captured as a HIP graph
and replayed.  Each result must equal
applying the document's pairs in order with bytes.replace (oracle apply_batch):
for every counted pattern class and the empty pattern, with documents without
pairs and empty documents mixed in; on graph replays separated by other GPU work
(each replay checked); after a round that overflows the engine's buffers (resumed
once they grew); and for the group size chosen by the autotuner.
"""
import random

import pytest

from oracle import replace_oracle as ro
from cuda_replace_amd import BatchEngine, Session

pytestmark = pytest.mark.gpu

COUNTED = [
    (b"ab", b"X"), (b"b\x00a", b"QRSTUVWXYZ"), (b"a", b""),          # border-free
    (b"aa", b"Y"), (b"aba", b""), (b"abab", b"ZZZZZZ"),               # tables of 2 and 4 states
    (b"aabaa", b"P"), (b"a" * 12, b"W"), (b"\x00\x00", b"\x00"),    # 8 and 16 states, NUL
]
EMPTY = (b"", b"-")
SIZES = [0, 1, 5, 100, 8191, 8192, 8193, 16384, 16385, 32769, 40000, 200000]


def _doc(rng, n):
    return bytes(rng.choice(b"ab\x00") for _ in range(n))


def _check(eng, docs, pairs, path="device"):
    got = eng.run(docs, pairs)
    for i, (d, prs) in enumerate(zip(docs, pairs)):
        assert got[i] == ro.apply_batch(d, prs), (i, len(d), prs)
    if path:
        assert eng.last_stats["path"] == path, eng.last_stats
    return got


def test_device_rounds_every_class(lib):
    rng = random.Random(91)
    docs, pairs = [], []
    for n in SIZES:
        for _ in range(3):
            docs.append(_doc(rng, n))
            k = rng.randint(0, 6)
            prs = [rng.choice(COUNTED) for _ in range(k)]
            if n < 3000 and rng.random() < 0.3:
                prs.insert(rng.randint(0, len(prs)), EMPTY)
            pairs.append(prs)
    with BatchEngine(lib) as eng:
        _check(eng, docs, pairs)
        st = eng.last_stats
        assert not st["graph"] and st["rounds"] == max(len(p) for p in pairs)   # a one-off shape
        _check(eng, docs, pairs)
        assert eng.last_stats["graph"] and not eng.last_stats["graph_replayed"]  # repeated: captured


def test_graph_replays_with_other_work_between(lib):
    rng = random.Random(92)
    docs = [_doc(rng, rng.randrange(0, 60000)) for _ in range(50)]
    pairs = [[rng.choice(COUNTED) for _ in range(rng.randint(1, 5))] for _ in docs]
    with BatchEngine(lib) as eng, BatchEngine(lib, host_rounds=True) as host:
        first = _check(eng, docs, pairs)
        assert _check(eng, docs, pairs) == first
        assert eng.last_stats["graph"] and not eng.last_stats["graph_replayed"]
        for rep_i in range(3):
            # other GPU work between the replays: a host-driven batch and a session pass
            _check(host, docs[:7], pairs[:7], path="host")
            with Session(lib, docs[0]) as s:
                s.apply(b"ab", b"Q")
                assert s.result() == docs[0].replace(b"ab", b"Q")
            again = _check(eng, docs, pairs)
            assert eng.last_stats["graph_replayed"], rep_i
            assert again == first


def test_new_shape_recaptures(lib):
    rng = random.Random(93)
    with BatchEngine(lib) as eng:
        for trial in range(4):
            docs = [_doc(rng, rng.randrange(0, 30000)) for _ in range(rng.randint(1, 20))]
            pairs = [[rng.choice(COUNTED) for _ in range(rng.randint(0, 4))] for _ in docs]
            _check(eng, docs, pairs, path=None)
            if any(pairs):
                assert eng.last_stats["path"] == "device"


def test_overflow_resumes_with_grown_buffers(lib):
    # 1-byte pattern -> 40 bytes: outputs far beyond the first run's buffer sizing;
    # the round that does not fit commits nothing and is resumed after the growth
    docs = [b"a" * 200_000, b"xa" * 3000, b"", b"q" * 9000]
    pairs = [[(b"a", b"Q" * 40)], [(b"a", b"R" * 40), (b"R", b"ss"), (b"s", b"tt")], [(b"a", b"b")], []]
    with BatchEngine(lib) as eng:
        _check(eng, docs, pairs)
        assert eng.last_stats["resumes"] >= 1
        _check(eng, docs, pairs)
        assert eng.last_stats["resumes"] == 0   # the engine kept the grown sizes


def test_bitmap_pairs_use_host_rounds(lib):
    rng = random.Random(94)
    docs = [_doc(rng, 20000) for _ in range(3)]
    pairs = [[(b"ab" * 10 + b"a", b"L")], [(b"ab", b"X")], []]
    with BatchEngine(lib) as eng:
        _check(eng, docs, pairs, path="host")


def test_group_size_tuned_on_first_run(lib):
    rng = random.Random(95)
    docs = [_doc(rng, rng.randrange(100_000, 400_000)) for _ in range(24)]        # ~6 MiB
    pairs = [[rng.choice(COUNTED[:8]) for _ in range(rng.randint(1, 3))] for _ in docs]
    with BatchEngine(lib) as eng:
        _check(eng, docs, pairs)
        first = dict(eng.last_stats)
        _check(eng, docs, pairs)
        second = dict(eng.last_stats)
    assert first["group_tiles"] in (1, 2, 4) and second["group_tiles"] == first["group_tiles"]
    assert not second["tuned"]
