"""Pins the CPU oracle (oracle/replace_oracle.py, oracle/cr_oracle.c).

Independent checks, none of which retypes the oracle's own formula:
* CPython ``bytes.replace`` / ``bytes.find`` / ``bytes.count`` on the
  reference's generated cases, known answers and regression inputs;
* the seeded form ``S[:F] + S[F:].replace(P, R)`` by brute force over every
  short input on ``{a, b, NUL}`` with every seed;
* closed forms on periodic inputs;
* planted mistakes in copies of the oracle must fail these checks.
"""
import hashlib
import itertools
import json
import os
import random

import pytest

from oracle import replace_oracle as ro
from tests.cases import KNOWN_ANSWERS, REFERENCE_CASES, REGRESSIONS, STREAMING_REGRESSIONS

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = json.load(open(os.path.join(HERE, "golden", "vectors.json")))
ALPHA = b"ab\x00"


def sha(b):
    return hashlib.sha256(b).hexdigest()


@pytest.fixture(scope="module")
def clib():
    if not os.path.exists(ro.C_LIB_PATH):
        import subprocess
        subprocess.run(["make", "-C", os.path.dirname(ro.C_LIB_PATH) + "/.."], check=True)
    return ro.c_lib()


def seeded_contract(src, pat, rep, seed):
    """CPython form of a seeded apply: bytes below F = max(0, seed + L) are
    copied, the rest is bytes.replace; kept starts from a find loop."""
    L = len(pat)
    f = 0 if seed is None else max(0, seed + L)
    tail = src[f:]
    ks, i = [], tail.find(pat)
    while i != -1:
        ks.append(f + i)
        i = tail.find(pat, i + L)
    return src[:f] + tail.replace(pat, rep), ks


def small_inputs():
    for n in range(0, 6):
        for t in itertools.product(ALPHA, repeat=n):
            yield bytes(t)
    rng = random.Random(7)
    for _ in range(600):
        n = rng.randint(6, 16)
        yield bytes(rng.choice(ALPHA) for _ in range(n))


PATTERNS = [bytes(t) for L in (1, 2, 3) for t in itertools.product(ALPHA, repeat=L)]


# ---------------------------------------------------------------- known answers
@pytest.mark.parametrize("name,src,pat,rep,expected", KNOWN_ANSWERS)
def test_known_answers(clib, name, src, pat, rep, expected):
    assert ro.replace(src, pat, rep) == expected
    assert ro.apply_seeded(src, pat, rep)[0] == expected
    assert ro.c_apply_seeded(src, pat, rep)[0] == expected


def test_golden_known_answers_match_cases():
    for rec, (name, src, pat, rep, expected) in zip(GOLDEN["known_answers"], KNOWN_ANSWERS):
        assert rec["name"] == name and bytes.fromhex(rec["out"]) == expected


@pytest.mark.parametrize("case", REFERENCE_CASES, ids=[c[0] for c in REFERENCE_CASES])
def test_reference_cases_c_oracle(clib, case):
    name, make, pat, rep = case
    rec = next(r for r in GOLDEN["reference_cases"] if r["name"] == name)
    data = make()
    assert sha(data) == rec["sha_in"]
    out, kept, last = ro.c_apply_seeded(data, pat, rep)
    assert len(out) == rec["n_out"] and sha(out) == rec["sha_out"]
    # border-free patterns: bytes.count is also leftmost non-overlapping
    assert kept == rec["count"] or pat in (b"aaa", b"xx")
    assert ro.c_kept_count(data, pat) == kept


@pytest.mark.parametrize("case", REGRESSIONS, ids=[c[0] for c in REGRESSIONS])
def test_regressions(clib, case):
    name, src, pat, rep = case
    rec = next(r for r in GOLDEN["regressions"] if r["name"] == name)
    out_py = ro.apply_seeded(src, pat, rep)[0]
    out_c = ro.c_apply_seeded(src, pat, rep)[0]
    assert sha(out_py) == rec["sha_out"] == sha(out_c)
    assert len(out_py) == rec["n_out"]


def test_golden_seeded(clib):
    for rec in GOLDEN["seeded"]:
        src, pat, rep = (bytes.fromhex(rec[k]) for k in ("src", "pat", "rep"))
        out, last = ro.apply_seeded(src, pat, rep, rec["seed"])
        assert out.hex() == rec["out"] and last == rec["last"]
        assert ro.kept_positions(src, pat, rec["seed"]) == rec["kept"]
        c_out, c_kept, c_last = ro.c_apply_seeded(src, pat, rep, rec["seed"])
        assert c_out.hex() == rec["out"] and c_last == rec["last"] and c_kept == len(rec["kept"])


# ----------------------------------------------------------------- brute force
def brute_force_failures(apply_fn, kept_fn=None, limit=None):
    """Counts disagreements of apply_fn(src, pat, rep, seed) -> (out, last) with
    the CPython contract over every small input, pattern and seed."""
    bad = 0
    for src in small_inputs():
        n = len(src)
        for pat in PATTERNS:
            L = len(pat)
            rep = b"R" * ((n + L) % 3)
            for seed in [None] + list(range(-L - 2, n + 2)):
                want, ks = seeded_contract(src, pat, rep, seed)
                out, last = apply_fn(src, pat, rep, seed)
                if out != want or last != (ks[-1] if ks else -1):
                    bad += 1
                if kept_fn is not None and kept_fn(src, pat, seed) != ks:
                    bad += 1
                if limit and bad >= limit:
                    return bad
    return bad


def test_python_oracle_brute_force():
    assert brute_force_failures(ro.apply_seeded, ro.kept_positions) == 0


def test_c_oracle_brute_force(clib):
    def c_apply(src, pat, rep, seed):
        out, _, last = ro.c_apply_seeded(src, pat, rep, seed)
        return out, last
    assert brute_force_failures(c_apply) == 0


def test_query_prefix_brute_force(clib):
    rng = random.Random(3)
    for src in small_inputs():
        for pat in PATTERNS[:12]:
            seed = rng.choice([None, -1, 0, 1, 3])
            _, ks = seeded_contract(src, pat, b"", seed)
            for limit in range(0, len(src) + 3):
                want = [k for k in ks if k < min(limit, len(src))]
                exp = (len(want), want[-1] if want else -1)
                assert ro.query_prefix(src, pat, seed, limit) == exp
                assert ro.c_query_prefix(src, pat, seed, limit) == exp


# ----------------------------------------------------------------- closed forms
@pytest.mark.parametrize("L", [1, 2, 3, 4, 7, 16, 17, 100])
def test_periodic_closed_form(clib, L):
    pat = b"a" * L
    for n in list(range(0, 40)) + [8191, 8192, 8193, 16387, 16388, 3 * 8192 + 5]:
        src = b"a" * n
        out, kept, last = ro.c_apply_seeded(src, pat, b"XY")
        assert kept == n // L
        assert out == b"XY" * (n // L) + b"a" * (n % L)
        assert last == (L * (n // L - 1) if n >= L else -1)


def test_empty_pattern_closed_form(clib):
    for src in [b"", b"a", b"abc", b"\x00\x00"]:
        for rep in [b"", b"X", b"XY"]:
            want = src.replace(b"", rep)
            assert ro.apply_seeded(src, b"", rep)[0] == want
            assert ro.c_apply_seeded(src, b"", rep)[0] == want


# ----------------------------------------------------------------- streaming
@pytest.mark.parametrize("case", STREAMING_REGRESSIONS, ids=[c[0] for c in STREAMING_REGRESSIONS])
def test_streaming_regressions(case):
    name, src, pat, rep, chunk = case
    assert ro.streaming(src, [(pat, rep)], chunk) == src.replace(pat, rep)


def test_streaming_brute_force():
    rng = random.Random(11)
    for _ in range(400):
        n = rng.randint(0, 24)
        src = bytes(rng.choice(ALPHA) for _ in range(n))
        pairs = [(bytes(rng.choice(ALPHA) for _ in range(rng.randint(0, 3))),
                  bytes(rng.choice(b"XY") for _ in range(rng.randint(0, 3))))
                 for _ in range(rng.randint(1, 3))]
        want = ro.apply_batch(src, pairs)
        for chunk in (1, 2, 3, 5, 8, 64):
            assert ro.streaming(src, pairs, chunk) == want, (src, pairs, chunk)


# ----------------------------------------------------------------- planted mistakes
def _mutant(kind):
    def apply_fn(src, pat, rep, seed):
        L = len(pat)
        last = seed
        kept = []
        i = src.find(pat)
        while i != -1:
            if kind == "strict_gap":
                ok = last is None or i - last > L
            elif kind == "seed_ignored":
                ok = last is None or last == seed or i - last >= L
            else:
                ok = last is None or i - last >= L
            if ok:
                kept.append(i)
                last = i
            i = src.find(pat, i + 1)
        out, prev = bytearray(), 0
        for k in kept:
            out += src[prev:k] + rep
            prev = k + (L - 1 if kind == "short_skip" else L)
        out += src[prev:]
        return bytes(out), (kept[-1] if kept else -1)
    return apply_fn


@pytest.mark.parametrize("kind", ["strict_gap", "seed_ignored", "short_skip"])
def test_planted_mistakes_are_caught(kind):
    assert brute_force_failures(_mutant(kind), limit=1) > 0


def test_unmutated_mutant_harness_passes():
    assert brute_force_failures(_mutant("none"), limit=1) == 0
