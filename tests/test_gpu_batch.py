"""Batch engine (cr_batch_run) parity on the GPU: every document's result must equal
applying its pairs in order with bytes.replace (oracle.replace_oracle.apply_batch).

Covers every pattern class the engine sorts pairs into (border-free; carry tables
of 2, 4, 8 and 16 states; the bitmap path for long self-overlapping patterns and
patterns over 256 bytes; empty patterns), empty documents, documents without
pairs, sizes around the 8 KiB tile and 16 KiB group boundaries, engine reuse
across runs with different shapes, and the seeded uneven-document workload.
"""
import random

import pytest

from oracle import replace_oracle as ro
from cuda_replace_amd import BatchEngine, workloads

pytestmark = pytest.mark.gpu

LONG_BF = bytes(range(1, 255)) + bytes(range(1, 50))          # 303 bytes, border-free
POOL = [
    (b"ab", b"X"),                      # border-free
    (b"b\x00a", b"QRSTUVWXYZ"),         # border-free, expand
    (b"a", b""),                        # border-free, delete
    (b"aa", b"Y"),                      # table, 2 states
    (b"aba", b""),                      # table, 4 states
    (b"abab", b"ZZZZZZ"),               # table, 4 states, expand
    (b"aabaa", b"P"),                   # table, 8 states
    (b"a" * 12, b"W"),                  # table, 16 states
    (b"ab" * 10 + b"a", b"L"),          # self-overlapping, 21 bytes: bitmap path
    (LONG_BF, b"#"),                    # border-free but 303 bytes: bitmap path
    (b"", b"-"),                        # empty pattern: insert at every gap
    (b"\x00\x00", b"\x00"),             # table with NUL bytes
]
SIZES = [0, 1, 5, 100, 8191, 8192, 8193, 16383, 16384, 16385, 40000, 200000]


def _doc(rng, n):
    body = bytearray(rng.choice(b"ab\x00") for _ in range(n))
    if n > len(LONG_BF) + 10 and rng.random() < 0.5:
        at = rng.randrange(0, n - len(LONG_BF))
        body[at:at + len(LONG_BF)] = LONG_BF
    return bytes(body)


def _pairs(rng, n, k):
    out = []
    for _ in range(k):
        p = rng.choice(POOL)
        if p[0] == b"" and n > 5000:
            continue   # keep inserted lengths small
        out.append(p)
    return out


def _check(engine, docs, pairs):
    got = engine.run(docs, pairs)
    assert len(got) == len(docs)
    for i, (d, prs) in enumerate(zip(docs, pairs)):
        want = ro.apply_batch(d, prs)
        assert got[i] == want, (i, len(d), prs)
    return got


@pytest.fixture(scope="module")
def engine(lib):
    e = BatchEngine(lib)
    yield e
    e.close()


def test_batch_every_class(engine):
    rng = random.Random(17)
    docs, pairs = [], []
    for n in SIZES:
        for _ in range(3):
            docs.append(_doc(rng, n))
            pairs.append(_pairs(rng, n, rng.randint(0, 6)))
    _check(engine, docs, pairs)


def test_batch_each_pair_alone(engine):
    rng = random.Random(18)
    docs = [_doc(rng, n) for n in (3000, 20000, 70000)]
    for p in POOL:
        if p[0] == b"":
            prs = [[p]] * len(docs)
            ds = [d[:2000] for d in docs]
        else:
            prs = [[p]] * len(docs)
            ds = docs
        _check(engine, ds, prs)


def test_batch_reuse_and_shrinking_shapes(engine):
    rng = random.Random(19)
    big = [_doc(rng, rng.randrange(0, 300000)) for _ in range(40)]
    _check(engine, big, [_pairs(rng, len(d), rng.randint(1, 8)) for d in big])
    small = [_doc(rng, rng.randrange(0, 3000)) for _ in range(7)]
    _check(engine, small, [_pairs(rng, len(d), rng.randint(0, 3)) for d in small])
    _check(engine, [], [])


def test_batch_uneven_documents_workload(engine):
    docs, pairs = workloads.uneven_documents(num_docs=192, median=24 << 10)
    _check(engine, docs, pairs)
    st = engine.last_stats
    assert st["rounds"] == max(len(p) for p in pairs)
    assert st["algo_bytes"] > 0 and st["device_ms"] > 0
