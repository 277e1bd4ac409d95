"""Continuous-batching server (cr_server_*) on the GPU: requests submitted together
from several threads are batched, and each result equals the same request run
alone and bytes.replace pair by pair (the oracle)."""
import random
import threading

import pytest

from oracle import replace_oracle as ro
from cuda_replace_amd import Server, workloads

pytestmark = pytest.mark.gpu


def _requests(seed, count):
    rng = random.Random(seed)
    text = workloads.zipf_text(8 << 20, seed=3)
    pool = workloads.pair_pool(text[: 1 << 20], 64, seed=9) + [(b"aa", b"Y"), (b"", b"-"), (b"ab" * 9 + b"a", b"L")]
    reqs = []
    for i in range(count):
        n = rng.choice([0, 17, 4096, 20000, 300000, 2_000_000])
        at = rng.randrange(0, len(text) - n) if n else 0
        doc = text[at:at + n] if i % 7 else bytes(rng.choice(b"ab") for _ in range(min(n, 5000)))
        prs = [rng.choice(pool) for _ in range(rng.randint(0, 6))]
        prs = [p for p in prs if p[0] or len(doc) < 10000]   # keep insertions small
        reqs.append((doc, prs))
    return reqs


def test_server_mix_equals_alone(lib):
    reqs = _requests(41, 48)
    # in the mix: four threads submit at once, results collected afterwards
    with Server(lib, workers=2) as srv:
        tickets = [None] * len(reqs)

        def submit(lo):
            for i in range(lo, len(reqs), 4):
                tickets[i] = srv.submit(*reqs[i])
        ths = [threading.Thread(target=submit, args=(k,)) for k in range(4)]
        for t in ths:
            t.start()
        for t in ths:
            t.join()
        mixed = [t.result() for t in tickets]
        batched = max(t.times["batch_docs"] for t in tickets)
    # alone: one request in flight at a time
    with Server(lib, workers=1) as srv:
        alone = [srv.submit(*r).result() for r in reqs]
    for i, (doc, prs) in enumerate(reqs):
        want = ro.apply_batch(doc, prs)
        assert alone[i] == want, i
        assert mixed[i] == want, i
    assert batched > 1   # the mix really was batched


def test_server_tiny_budget_runs_requests_alone(lib):
    reqs = _requests(42, 12)
    with Server(lib, hbm_budget=1 << 20, workers=2) as srv:
        ts = [srv.submit(*r) for r in reqs]
        outs = [t.result() for t in ts]
        assert all(t.times["batch_docs"] == 1 for t in ts)   # every request exceeds half the budget
    for (doc, prs), got in zip(reqs, outs):
        assert got == ro.apply_batch(doc, prs)


def test_arrivals_join_a_running_worker(lib):
    """One worker: a long request (64 pairs over 8 MiB) is in its rounds when short
    requests arrive; they are admitted into the next round, so they come back
    before the long one, each equal to the oracle."""
    import time
    text = workloads.zipf_text(8 << 20, seed=5)
    pool = workloads.pair_pool(text[: 1 << 20], 64, seed=13)
    long_pairs = [pool[i % len(pool)] for i in range(64)]
    rng = random.Random(3)
    shorts = []
    for _ in range(12):
        at = rng.randrange(0, len(text) - 50000)
        shorts.append((text[at:at + rng.randrange(100, 50000)], [rng.choice(pool) for _ in range(rng.randint(1, 3))]))
    with Server(lib, workers=1) as srv:
        t0 = time.perf_counter()
        lt = srv.submit(text, long_pairs)
        time.sleep(0.005)
        sts = [(time.perf_counter(), srv.submit(d, p)) for d, p in shorts]
        outs = [(at, t, t.result()) for at, t in sts]
        lres = lt.result()
        long_done = t0 + lt.times["latency_ms"] / 1e3
        for (d, p), (at, t, got) in zip(shorts, outs):
            assert got == ro.apply_batch(d, p)
            assert at + t.times["latency_ms"] / 1e3 < long_done, (t.times, lt.times)
            assert t.times["batch_docs"] >= 2, t.times   # joined the live long request
    assert lres == ro.apply_batch(text, long_pairs)
