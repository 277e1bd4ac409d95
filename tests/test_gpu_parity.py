"""Parity of the HIP library with the oracle, through the C ABI (gpu).

Bit-exact comparisons against CPython bytes.replace and the oracle's seeded /
query restatement, on the reference's generated cases, regression inputs,
random small inputs over a three-letter alphabet, tile-boundary stress inputs
and the seeded benchmark inputs (sha256 fixtures)."""
import hashlib
import json
import os
import random
import sys
import threading

import pytest

from oracle import replace_oracle as ro
from tests.cases import KNOWN_ANSWERS, REFERENCE_CASES, REGRESSIONS, STREAMING_REGRESSIONS

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = json.load(open(os.path.join(HERE, "golden", "vectors.json")))
TILE = 8192


def sha(b):
    return hashlib.sha256(b).hexdigest()


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def seeded_contract(src, pat, rep, seed):
    L = len(pat)
    f = 0 if seed is None else max(0, seed + L)
    out, last = ro.apply_seeded(src, pat, rep, seed)
    assert out == src[:f] + src[f:].replace(pat, rep)
    return out, last


@pytest.fixture(scope="module")
def sess(lib):
    from cuda_replace_amd import Session
    s = Session(lib, b"")
    yield s
    s.close()


# ------------------------------------------------------------ known answers
@pytest.mark.parametrize("name,src,pat,rep,expected", KNOWN_ANSWERS)
def test_known_answers(lib, name, src, pat, rep, expected):
    say("known", name)
    from cuda_replace_amd import Session
    assert lib.unified(src, pat, rep) == expected
    with Session(lib, src) as s:
        assert s.apply(pat, rep) == len(expected)
        assert s.result() == expected


@pytest.mark.parametrize("case", REGRESSIONS, ids=[c[0] for c in REGRESSIONS])
def test_regressions(lib, case):
    name, src, pat, rep = case
    say("regression", name)
    rec = next(r for r in GOLDEN["regressions"] if r["name"] == name)
    out = lib.unified(src, pat, rep)
    assert len(out) == rec["n_out"] and sha(out) == rec["sha_out"]


@pytest.mark.parametrize("case", REFERENCE_CASES, ids=[c[0] for c in REFERENCE_CASES])
def test_reference_cases(lib, case):
    name, make, pat, rep = case
    say("reference case", name)
    rec = next(r for r in GOLDEN["reference_cases"] if r["name"] == name)
    out = lib.unified(make(), pat, rep)
    assert len(out) == rec["n_out"] and sha(out) == rec["sha_out"]


def test_golden_seeded(sess):
    for rec in GOLDEN["seeded"]:
        src, pat, rep = (bytes.fromhex(rec[k]) for k in ("src", "pat", "rep"))
        say("seeded", rec["src"], rec["pat"], rec["seed"])
        sess.reset(src)
        n, last = sess.apply_seeded(pat, rep, rec["seed"])
        assert sess.result().hex() == rec["out"] and last == rec["last"] and n == len(rec["out"]) // 2
        for lim in range(0, len(src) + 2):
            want = [k for k in rec["kept"] if k < min(lim, len(src))]
            assert sess.query_prefix(lim) == (len(want), want[-1] if want else -1)


# ------------------------------------------------------------ random small inputs
def test_random_small_seeded(sess):
    rng = random.Random(1234)
    alpha = b"ab\x00"
    pats = [b"a", b"b", b"ab", b"aa", b"aba", b"abab", b"aab", b"a\x00", b"aaaa", b"abaab", b"aaaaaaaa",
            b"ab" * 8, b"a" * 17, b"ab" * 9 + b"a", b"\x00\x00", b"ba" * 3 + b"b", b""]
    say("random small seeded")
    for it in range(1500):
        n = rng.choice([rng.randint(0, 40), rng.randint(0, 300)])
        src = bytes(rng.choice(alpha) for _ in range(n))
        pat = rng.choice(pats)
        rep = bytes(rng.choice(b"XYZ") for _ in range(rng.randint(0, 6)))
        seed = rng.choice([None, None, -1, 0, 1, 2, n // 2, n + 3])
        want, want_last = seeded_contract(src, pat, rep, seed) if pat else (src.replace(pat, rep), n)
        sess.reset(src)
        if seed is None:
            got_n = sess.apply(pat, rep)
            got_last = None
        else:
            got_n, got_last = sess.apply_seeded(pat, rep, seed)
        got = sess.result()
        assert got == want, (it, src, pat, rep, seed)
        assert got_n == len(want)
        if got_last is not None and pat:
            assert got_last == want_last
        if pat:
            for lim in (0, 1, n // 3, n // 2, n - 1, n, n + 5):
                assert sess.query_prefix(lim) == ro.query_prefix(src, pat, seed, lim), (it, lim)


@pytest.mark.parametrize("pat", [b"aa", b"aaa", b"aaaa", b"a" * 7, b"a" * 16, b"a" * 17, b"ab", b"a" * 100])
def test_periodic_tile_boundaries(sess, pat):
    L = len(pat)
    say("periodic", L)
    for n in [TILE - 1, TILE, TILE + 1, 2 * TILE + 3, 5 * TILE - 2, 64 * TILE + 11, 300 * TILE + 5]:
        src = b"a" * n
        sess.reset(src)
        got_n = sess.apply(pat, b"XY")
        k = n // L if pat == b"a" * L else src.count(pat)
        want = src.replace(pat, b"XY")
        assert got_n == len(want)
        out = sess.result()
        assert out == want, (L, n)
        if pat == b"a" * L:
            assert k == n // L


def test_matches_straddling_tiles(sess):
    rng = random.Random(99)
    say("straddle")
    for pat in [b"NEEDLE", b"abcab", b"qarn", b"xyzzy" * 3, b"aab"]:
        L = len(pat)
        buf = bytearray(b"." * (12 * TILE))
        for t in range(1, 12):
            off = t * TILE - rng.randint(1, L - 1) if L > 1 else t * TILE
            buf[off:off + L] = pat
        src = bytes(buf)
        for rep in (b"", b"#", b"!" * (L + 5)):
            sess.reset(src)
            sess.apply(pat, rep)
            assert sess.result() == src.replace(pat, rep), (pat, rep)


def test_modes_on_text(sess, lib):
    from cuda_replace_amd import workloads
    say("modes on text")
    text = workloads.zipf_text(3 << 20, 5)
    frag = workloads.text_fragment(text[: 1 << 20], 5)
    pats = [frag, b"e", b" ", b"ee", b"eee", b"the", b"a a", b"\n", b"zz", frag + b" ", b"ab" * 9,
            b"q" * 20, bytes(range(97, 97 + 26)) * 11]
    for pat in pats:
        for rep in (b"", b"1", b"12345678901234567890"):
            sess.reset(text)
            sess.apply(pat, rep)
            assert sess.result() == text.replace(pat, rep), (pat, rep, lib.pattern_mode(pat))


# ------------------------------------------------------------ batch / query / streaming
def test_apply_batch(sess):
    rng = random.Random(7)
    say("apply_batch")
    for _ in range(40):
        n = rng.randint(0, 50000)
        src = bytes(rng.choice(b"ab&;x") for _ in range(n))
        pairs = [(bytes(rng.choice(b"ab&;x") for _ in range(rng.randint(0, 5))),
                  bytes(rng.choice(b"XYZab") for _ in range(rng.randint(0, 6))))
                 for _ in range(rng.randint(1, 8))]
        sess.reset(src)
        n_out = sess.apply_batch(pairs)
        want = ro.apply_batch(src, pairs)
        assert n_out == len(want)
        assert sess.result() == want


def test_query_prefix_large(sess):
    rng = random.Random(5)
    say("query large")
    src = bytes(rng.choice(b"aab") for _ in range(5 * TILE + 77))
    for pat in (b"ab", b"aa", b"aab", b"a" * 17):
        for seed in (None, 3):
            sess.reset(src)
            if seed is None:
                sess.apply(pat, b"Z")
            else:
                sess.apply_seeded(pat, b"Z", seed)
            ks = ro.kept_positions(src, pat, seed)
            for lim in [0, 1, TILE - 1, TILE, TILE + 1, 3 * TILE + 5, len(src), len(src) + 9]:
                want = [k for k in ks if k < min(lim, len(src))]
                assert sess.query_prefix(lim) == (len(want), want[-1] if want else -1), (pat, seed, lim)


@pytest.mark.parametrize("case", STREAMING_REGRESSIONS, ids=[c[0] for c in STREAMING_REGRESSIONS])
def test_streaming_regressions(lib, case):
    from cuda_replace_amd import gpu_replace_streaming
    name, src, pat, rep, chunk = case
    say("streaming", name)
    assert gpu_replace_streaming(lib, src, [(pat, rep)], chunk) == src.replace(pat, rep)


def test_streaming_text(lib):
    from cuda_replace_amd import gpu_replace_streaming
    say("streaming text")
    text = (b"hello world &amp; hello &lt;x&gt; aaaa " * 5000)
    pairs = [(b"&amp;", b"&"), (b"hello", b"hi"), (b"aa", b"b"), (b"&lt;", b"<"), (b"", b"|")]
    want = ro.apply_batch(text, pairs)
    for chunk in (7, 4096, 100000, 1 << 24):
        assert gpu_replace_streaming(lib, text, pairs, chunk) == want


# ------------------------------------------------------------ threads
def test_threads_distinct_and_shared_sessions(lib):
    from cuda_replace_amd import Session
    say("threads")
    errors = []
    shared = Session(lib, b"")
    lock_free_results = []

    def worker(i):
        try:
            rng = random.Random(i)
            with Session(lib, b"") as s:
                for _ in range(20):
                    src = bytes(rng.choice(b"abc") for _ in range(rng.randint(0, 30000)))
                    pat = rng.choice([b"ab", b"aa", b"abc", b"cc"])
                    s.reset(src)
                    s.apply(pat, b"Q")
                    if s.result() != src.replace(pat, b"Q"):
                        errors.append(("own", i))
            for _ in range(10):
                lock_free_results.append(lib.unified(b"xaax" * 1000, b"aa", b"b") == b"xbx" * 1000)
                shared.length()
        except Exception as e:  # noqa: BLE001
            errors.append(repr(e))

    ts = [threading.Thread(target=worker, args=(i,)) for i in range(8)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    shared.close()
    assert not errors
    assert all(lock_free_results)


# ------------------------------------------------------------ benchmark inputs
@pytest.mark.parametrize("name", ["uniform_abcd_64mib_ab_X", "zipf_text_64mib_expand", "zipf_text_64mib_shrink"])
def test_config_fixtures(lib, name):
    from cuda_replace_amd import Session, workloads
    say("config", name)
    rec = next(r for r in GOLDEN["configs"] if r["name"] == name)
    if name.startswith("uniform"):
        src = workloads.uniform_abcd(rec["n_in"], 0)
    else:
        src = workloads.zipf_text(rec["n_in"])
    assert sha(src) == rec["sha_in"]
    with Session(lib, src) as s:
        n = s.apply(bytes.fromhex(rec["pat"]), bytes.fromhex(rec["rep"]))
        out = s.result()
    assert n == rec["n_out"] and sha(out) == rec["sha_out"]


def test_all_a_64mib_closed_form(lib):
    from cuda_replace_amd import Session
    say("all-a 64 MiB")
    n = 64 << 20
    with Session(lib, b"a" * n) as s:
        assert s.apply(b"aa", b"X") == n // 2
        assert s.result() == b"X" * (n // 2)
        s.reset(b"a" * (n + 1))
        assert s.apply(b"aa", b"X") == n // 2 + 1
        assert s.result() == b"X" * (n // 2) + b"a"


# ------------------------------------------------------------ count first, then emit
@pytest.mark.parametrize("pat", [b"a", b"aa", b"aba"])
def test_expansion_beyond_the_upfront_bound(lib, pat):
    # 16 MiB with a 64-byte replacement: the worst case (up to 64 x the input) is
    # beyond what a pass allocates up front, so count and spine run first, the
    # exact size is read back, and the emit runs later over the candidate masks
    # the count left behind; once plain and once seeded
    from cuda_replace_amd import Session
    rng = random.Random(131)
    unit = bytes(rng.choice(b"abcdefgh") for _ in range(1 << 16))
    src = unit * 256
    rep = bytes(rng.choice(b"XYZ") for _ in range(64))
    assert lib.unified(src, pat, rep) == src.replace(pat, rep)
    with Session(lib, src) as s:
        seed = 5_000_003
        n2, last = s.apply_seeded(pat, rep, seed)
        want, want_last = seeded_contract(src, pat, rep, seed)
        assert s.result() == want and n2 == len(want) and last == want_last


def test_modes_alternate_on_one_buffer(lib):
    # passes of every mode applied one after another to the same buffer (no reset):
    # the bitmap mode's bitmap and the count's candidate masks share one scratch
    # buffer, and each pass must see its own; a prefix query after each pass
    from cuda_replace_amd import Session
    rng = random.Random(141)
    src = bytes(rng.choice(b"ab \n") for _ in range(3 << 20))
    steps = [(b"ab", b"X"), (b"ab" * 9 + b"a", b"Q"), (b"aa", b"a"), (b"X", b"ba"), (b"b" * 20, b""),
             (b"aba", b"Z"), (b" \n", b"ab")]
    want = src
    with Session(lib, src) as s:
        for pat, rep in steps:
            s.apply(pat, rep)
            prev = want
            want = want.replace(pat, rep)
            assert s.result() == want, (pat, lib.pattern_mode(pat))
            lim = len(prev) // 2
            assert s.query_prefix(lim) == ro.query_prefix(prev, pat, None, lim), pat
