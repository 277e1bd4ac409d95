"""Fused multi-pattern apply_batch parity on the GPU: when cr_batch_fusable accepts
a pair list, cuda_replace_apply_batch runs it as one pass, and the result must
equal applying the pairs in order with bytes.replace (oracle apply_batch) and
the same session's pair-by-pair path (fuse=False).

Covers sizes around the 8 KiB tile and 16 KiB group boundaries, occurrences
planted across every such boundary (the emit's coverage carry and halo probe),
dense inputs made only of occurrences, heavy expansion (the unstaged emit) and
heavy shrink (negative totals), 64-byte patterns, 256 pairs with colliding hash
keys, and prefix queries after a fused batch (answered for the last pair by a
replay of the pairs).  Every test runs both forms of the fused pass: count, spine
and emit launches (CR_FUSED_SWEEP=0) and the single persistent launch that holds
each group in LDS from its count to its emit (CR_FUSED_SWEEP_MIN=0 forces it onto
every size).
"""
import os
import random

import pytest

from oracle import replace_oracle as ro
from cuda_replace_amd import Session, workloads

pytestmark = pytest.mark.gpu

SIZES = [0, 1, 5, 100, 8191, 8192, 8193, 16383, 16384, 16385, 40000, 200000, 3 << 20]


@pytest.fixture(autouse=True, params=["three_launch", "single_launch"])
def fused_form(request):
    keys = ("CR_FUSED_SWEEP", "CR_FUSED_SWEEP_MIN")
    saved = {k: os.environ.get(k) for k in keys}
    if request.param == "single_launch":
        os.environ.pop("CR_FUSED_SWEEP", None)
        os.environ["CR_FUSED_SWEEP_MIN"] = "0"
    else:
        os.environ["CR_FUSED_SWEEP"] = "0"
    yield request.param
    for k, v in saved.items():
        if v is None:
            os.environ.pop(k, None)
        else:
            os.environ[k] = v
LETTERS = b"abcdefgh"
REP_ALPHA = b"XYZ0123456789"


def entity_list(rng, k, min_body=1, max_body=6, max_rep=6):
    """k distinct patterns b"&" + body + b";" (body over LETTERS) with replacements
    over REP_ALPHA: always fusable."""
    seen, out = set(), []
    while len(out) < k:
        body = bytes(rng.choice(LETTERS) for _ in range(rng.randint(min_body, max_body)))
        pat = b"&" + body + b";"
        if pat in seen:
            continue
        seen.add(pat)
        out.append((pat, bytes(rng.choice(REP_ALPHA) for _ in range(rng.randint(1, max_rep)))))
    return out


def planted(rng, n, pairs, density=0.02, boundaries=()):
    buf = bytearray(rng.choice(LETTERS + b" &;\x00") for _ in range(n))
    pats = [p for p, _ in pairs]
    i = 0
    while i < n:
        if rng.random() < density:
            p = rng.choice(pats)
            if i + len(p) <= n:
                buf[i:i + len(p)] = p
                i += len(p)
                continue
        i += 1
    for b in boundaries:   # an occurrence ending 1..L-1 bytes past a boundary
        p = rng.choice(pats)
        at = b - rng.randint(1, len(p) - 1)
        if 0 <= at and at + len(p) <= n:
            buf[at:at + len(p)] = p
    return bytes(buf)


def run(lib, src, pairs, fuse=True):
    with Session(lib, src) as s:
        n = s.apply_batch(pairs, fuse=fuse)
        out = s.result()
        assert n == len(out)
        return out, s.last_batch_fused


def check(lib, src, pairs):
    assert lib.batch_fusable(pairs)
    want = ro.apply_batch(src, pairs)
    got, fused = run(lib, src, pairs)
    assert fused
    assert got == want, (len(src), pairs[:3])
    seq, fused = run(lib, src, pairs, fuse=False)
    assert not fused and seq == want
    return got


def test_fused_sizes_and_boundaries(lib):
    rng = random.Random(41)
    for n in SIZES:
        pairs = entity_list(rng, rng.randint(2, 12))
        bounds = [k * 8192 for k in range(1, n // 8192 + 1)][:600]
        check(lib, planted(rng, n, pairs, boundaries=bounds), pairs)


def test_fused_dense_occurrences_only(lib):
    rng = random.Random(42)
    for n_occ in (1, 3, 5000, 60000):
        pairs = entity_list(rng, 6, max_body=3)
        src = b"".join(rng.choice(pairs)[0] for _ in range(n_occ))
        check(lib, src, pairs)


def test_fused_heavy_expand_and_shrink(lib):
    rng = random.Random(43)
    # 2-byte patterns to 64-byte replacements: tiles grow ~32x, the emit writes directly
    pairs = [(b"&" + bytes([c]), bytes(rng.choice(REP_ALPHA) for _ in range(64))) for c in b"abcd"]
    src = b"".join(rng.choice([b"&a", b"&b", b"&c", b"&d", b"x"]) for _ in range(100000))
    check(lib, src, pairs)
    # 64-byte patterns to 1-byte replacements: the total length change is negative
    pairs = [(b"&" + bytes([c]) * 62 + b";", b"Z") for c in b"abcdef"]
    src = b"".join(rng.choice([p for p, _ in pairs] + [b"q"]) for _ in range(40000))
    out = check(lib, src, pairs)
    assert len(out) < len(src) // 10


def test_fused_many_pairs_colliding_keys(lib):
    rng = random.Random(44)
    # 256 patterns sharing their first 5 bytes: one hash key, a 256-long probe chain
    pairs = [(b"&abcd" + bytes([97 + i // 16, 97 + i % 16]) + b";", bytes([48 + i % 10, 65 + i // 10 % 26]))
             for i in range(256)]
    src = planted(rng, 300000, pairs, density=0.05)
    check(lib, src, pairs)
    pairs = entity_list(rng, 256, min_body=2)
    check(lib, planted(rng, 500000, pairs, density=0.05), pairs)


def test_query_after_fused_batch(lib):
    rng = random.Random(45)
    pairs = entity_list(rng, 5)
    src = planted(rng, 100000, pairs, density=0.05)
    mid = ro.apply_batch(src, pairs[:-1])
    last_pat = pairs[-1][0]
    limits = [0, 1, 4000, 50000, len(mid), len(mid) + 10]
    with Session(lib, src) as s:
        s.apply_batch(pairs)
        assert s.last_batch_fused
        for lim in limits:
            assert s.query_prefix(lim) == ro.query_prefix(mid, last_pat, None, min(lim, len(mid))), lim
        assert s.result() == ro.apply_batch(src, pairs)
        # a plain apply after the fused batch: queries follow it
        out = s.result()
        s.apply(b"XY", b"")
        for lim in (0, 777, len(out)):
            assert s.query_prefix(lim) == ro.query_prefix(out, b"XY", None, min(lim, len(out)))
        assert s.result() == out.replace(b"XY", b"")


def test_fused_then_sequential_batches_on_one_session(lib):
    rng = random.Random(46)
    src = planted(rng, 70000, entity_list(rng, 4), density=0.03)
    with Session(lib, src) as s:
        want = src
        for _ in range(4):
            pairs = entity_list(rng, rng.randint(2, 6))
            s.apply_batch(pairs)
            assert s.last_batch_fused
            want = ro.apply_batch(want, pairs)
            s.apply_batch([(b"X", b"xx"), (b"xx", b"Q")])   # not fusable: pair by pair
            assert not s.last_batch_fused
            want = ro.apply_batch(want, [(b"X", b"xx"), (b"xx", b"Q")])
            assert s.result() == want


def test_fused_entity_workload(lib):
    pairs = workloads.entity_pairs()
    src = workloads.entity_text(64 << 20, pairs)
    check(lib, src, pairs)


def test_fused_mixed_lengths(lib):
    # 2- to 64-byte patterns in one list: a 1-byte hash key, pattern bytes compared
    # 16 at a time past the first 16.  b"&h" is fusable with the b"&...;" patterns
    # because no body starts with h.
    rng = random.Random(47)
    pairs, seen = [(b"&h", b"H2")], set()
    for L in list(range(3, 65)) + [17, 31, 33, 48, 63, 64]:
        while True:
            pat = b"&" + bytes(rng.choice(LETTERS[:7]) for _ in range(L - 2)) + b";"
            if pat not in seen:
                break
        seen.add(pat)
        pairs.append((pat, bytes(rng.choice(REP_ALPHA) for _ in range(rng.randint(1, 64)))))
    for n in (1000, 70000, 1 << 20):
        check(lib, planted(rng, n, pairs, density=0.03, boundaries=range(8192, n, 8192)), pairs)


def test_repeated_and_changed_pair_lists(lib):
    # a session keeps the last list's fusability check and table: a repeated list
    # reuses them, a list differing in one byte (or one more pair, or one that is not
    # fusable) must not
    rng = random.Random(48)
    pairs = entity_list(rng, 6)
    changed = pairs[:-1] + [(pairs[-1][0], pairs[-1][1][:-1] + b"Z" if pairs[-1][1][-1:] != b"Z" else b"Y")]
    longer = pairs + [(b"&zz;", b"W")]
    unfusable = pairs + [(b"W", b"&a;")]
    seq = [pairs, pairs, changed, pairs, longer, unfusable, pairs, pairs]
    want_fused = [True, True, True, True, True, False, True, True]
    with Session(lib, b"") as s:
        for k, (prs, wf) in enumerate(zip(seq, want_fused)):
            src = planted(rng, rng.choice([0, 5000, 70000]), pairs + [(b"&zz;", b"")], density=0.05)
            s.reset(src)
            s.apply_batch(prs)
            assert s.last_batch_fused == wf, k
            assert s.result() == ro.apply_batch(src, prs), k


def test_pair_list_mutated_between_calls(lib):
    # the wrapper reuses the pointer arrays of an equal pair list: a list changed in
    # place, pairs given as lists or bytearrays, and the original list again must
    # each produce their own result
    rng = random.Random(49)
    pairs = entity_list(rng, 5)
    src = planted(rng, 40000, pairs, density=0.05)
    with Session(lib, b"") as s:
        mutated = list(pairs)
        for want_pairs in (pairs, pairs):
            s.reset(src)
            s.apply_batch(want_pairs)
            assert s.result() == ro.apply_batch(src, want_pairs)
        mutated[0] = (mutated[0][0], mutated[0][1] + b"Q")
        for prs in (mutated, [list(p) for p in pairs], [(bytearray(p), bytearray(r)) for p, r in pairs], pairs):
            want = ro.apply_batch(src, [(bytes(p), bytes(r)) for p, r in prs])
            s.reset(src)
            s.apply_batch(prs)
            assert s.result() == want
        edit = [(bytearray(p), bytearray(r)) for p, r in pairs]
        s.reset(src)
        s.apply_batch(edit)
        edit[1][1][0:1] = b"V"                      # changed after the call, before the next
        s.reset(src)
        s.apply_batch(edit)
        assert s.result() == ro.apply_batch(src, [(bytes(p), bytes(r)) for p, r in edit])


def test_unbounded_expansion_reads_the_size_first(lib):
    # 2-byte patterns to 64-byte replacements over 16 MiB: the worst case (32 x the
    # input) is beyond what a pass may allocate up front, so the exact size is read
    # between the spine and the emit
    rng = random.Random(49)
    pairs = [(b"&" + bytes([c]), bytes(rng.choice(REP_ALPHA) for _ in range(64))) for c in b"abcd"]
    src = planted(rng, 1 << 20, pairs, density=0.01) * 16
    check(lib, src, pairs)
