"""Fused multi-pattern apply_batch: the library's fusability check
(cr_batch_fusable, host code, no GPU call) against the property it must
guarantee.

When the check accepts a pair list, applying the pairs in order with
bytes.replace (the reference's apply_batch, cuda_replace.cu:2465-2527) must equal
one left-to-right scan that replaces every occurrence of every pattern -- the
single pass the GPU runs.  The GPU pass replaces every occurrence without a
greedy, so it also needs the occurrences of all patterns to be pairwise
disjoint.  The brute force below checks both on every input of up to 6 bytes over
the patterns' alphabet and on random longer inputs built from pattern pieces; the
counterexamples show that each rejected condition is needed (a check missing it
would accept a list whose fused result differs or whose occurrences overlap).
"""
import itertools
import random

import pytest

from oracle import replace_oracle as ro
from cuda_replace_amd import workloads


def fused_scan(src: bytes, pairs) -> bytes:
    """One left-to-right pass: at each position, the first pattern that occurs is
    replaced and skipped; otherwise the byte is copied."""
    table = {p: r for p, r in reversed(pairs)}
    lens = sorted({len(p) for p, _ in pairs})
    out, i = bytearray(), 0
    while i < len(src):
        for L in lens:
            piece = src[i:i + L]
            if len(piece) == L and piece in table:
                out += table[piece]
                i += L
                break
        else:
            out.append(src[i])
            i += 1
    return bytes(out)


def occurrences_disjoint(src: bytes, pairs) -> bool:
    spans = sorted((i, i + len(p)) for p, _ in pairs for i in range(len(src) - len(p) + 1)
                   if src[i:i + len(p)] == p)
    return all(b[0] >= a[1] for a, b in zip(spans, spans[1:]))


POOL = [b"ab", b"ac", b"acb", b"abc", b"abbc", b"accb", b"abcc", b"acbb", b"abcb", b"ad", b"adb", b"abd"]
REPS = [b"X", b"YY", b"Z", b"XYZ", b"b", b"a"]


def test_fused_scan_is_a_scan():
    # the model itself on hand-checked cases
    assert fused_scan(b"xabyacz", [(b"ab", b"1"), (b"ac", b"22")]) == b"x1y22z"
    assert fused_scan(b"", [(b"ab", b"1")]) == b""
    assert fused_scan(b"aab", [(b"ab", b"Q")]) == b"aQ"


def test_accepted_lists_equal_sequential_brute_force(lib):
    rng = random.Random(11)
    accepted = 0
    for trial in range(400):
        k = rng.randint(2, 4)
        pats = rng.sample(POOL, k)
        pairs = [(p, rng.choice(REPS)) for p in pats]
        if not lib.batch_fusable(pairs):
            continue
        accepted += 1
        alpha = sorted(set(b"".join(pats)) | {ord("X")})
        for n in range(0, 6 if trial < 40 else 4):
            for t in itertools.product(alpha, repeat=n):
                src = bytes(t)
                assert fused_scan(src, pairs) == ro.apply_batch(src, pairs), (pairs, src)
                assert occurrences_disjoint(src, pairs), (pairs, src)
        for _ in range(200):
            src = b"".join(rng.choice(pats + [b"a", b"b", b"c", b"X", b"d"]) for _ in range(rng.randint(0, 30)))
            assert fused_scan(src, pairs) == ro.apply_batch(src, pairs), (pairs, src)
            assert occurrences_disjoint(src, pairs), (pairs, src)
    assert accepted >= 40   # the brute force is not vacuous


# (pairs, input, the condition the list violates)
COUNTEREXAMPLES = [
    ([(b"ab", b"a"), (b"ac", b"Z")], b"abc", "replacement holds a pattern byte"),
    ([(b"abc", b"X"), (b"ab", b"Y")], b"abc", "pattern is a prefix of another"),
    ([(b"ac", b"X"), (b"abac", b"Y")], b"abac", "pattern inside another"),
    ([(b"abd", b"Y"), (b"acab", b"X")], b"acabd", "suffix of one is a prefix of another"),
    ([(b"aa", b"X"), (b"ab", b"Y")], b"aaab", "pattern with a border"),
    ([(b"ab", b"X"), (b"ab", b"Y")], b"ab", "duplicate pattern"),
    ([(b"ab", b""), (b"ac", b"Z")], b"aabc", "empty replacement"),
]


@pytest.mark.parametrize("pairs,src,why", COUNTEREXAMPLES, ids=[c[2] for c in COUNTEREXAMPLES])
def test_each_condition_is_needed(lib, pairs, src, why):
    differs = fused_scan(src, pairs) != ro.apply_batch(src, pairs)
    assert differs or not occurrences_disjoint(src, pairs), why
    assert not lib.batch_fusable(pairs), why


def test_design_limits_rejected(lib):
    # fusable in principle, outside the kernel's limits: the pairs run one by one
    assert not lib.batch_fusable([(b"ab", b"X")])                          # one pair
    assert not lib.batch_fusable([(b"ab", b"X"), (b"cd", b"Y")])           # no shared first byte
    assert not lib.batch_fusable([(b"a", b"X"), (b"bc", b"Y")])            # 1-byte pattern
    long_pat = b"a" + bytes(range(130, 194))                                 # 65 bytes, border-free
    assert not lib.batch_fusable([(long_pat, b"X"), (b"ab", b"Y")])
    assert lib.batch_fusable([(long_pat[:64], b"X"), (b"ab", b"Y")])
    assert not lib.batch_fusable([(b"ab", b"X" * 65), (b"ac", b"Y")])      # replacement over 64 bytes
    many = [(b"&" + bytes([97 + i // 26, 97 + i % 26]) + b";", b"Q") for i in range(257)]
    assert lib.batch_fusable(many[:256]) and not lib.batch_fusable(many)


def test_entity_workload_is_fusable(lib):
    pairs = workloads.entity_pairs()
    assert len(pairs) == 128 and lib.batch_fusable(pairs)
    text = workloads.entity_text(1 << 20, pairs)
    assert fused_scan(text[:200_000], pairs) == ro.apply_batch(text[:200_000], pairs)
