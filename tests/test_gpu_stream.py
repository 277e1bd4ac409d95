"""Pipelined streaming (cr_stream_replace) parity on the GPU: the result of every
pair list must equal bytes.replace applied pair by pair to the whole input, for
every chunk size, including chunks shorter than the pattern (a kept match then
covers several chunk ends), self-overlapping patterns on the table and bitmap
paths, empty patterns and empty inputs."""
import random

import pytest

from oracle import replace_oracle as ro
from cuda_replace_amd import gpu_replace_streaming, workloads

pytestmark = pytest.mark.gpu

PATS = [(b"ab", b"X"), (b"aa", b"Y"), (b"aba", b""), (b"aaaa", b"QQQQQQQ"), (b"abaab", b"#"),
        (b"ab" * 10 + b"a", b"L"), (b"", b"-"), (b"a", b"bb"), (b"b\x00b\x00b", b"\x00")]


def test_stream_small_inputs_all_chunk_sizes(lib):
    rng = random.Random(31)
    for trial in range(40):
        n = rng.choice([0, 1, 2, 7, 30, 100, 257])
        src = bytes(rng.choice(b"ab\x00") for _ in range(n))
        pairs = [rng.choice(PATS) for _ in range(rng.randint(1, 3))]
        want = ro.apply_batch(src, pairs)
        for chunk in (1, 2, 3, 5, 8, 64, 1000):
            got = gpu_replace_streaming(lib, src, pairs, chunk)
            assert got == want, (trial, src, pairs, chunk)


def test_stream_tile_sized_chunks(lib):
    rng = random.Random(32)
    src = bytes(rng.choice(b"ab") for _ in range(200_000))
    for pairs in ([(b"aa", b"Y")], [(b"ab" * 10 + b"a", b"L"), (b"ba", b"")], [(b"abab", b"Z" * 9)]):
        want = ro.apply_batch(src, pairs)
        for chunk in (8191, 8192, 16385, 50_000):
            assert gpu_replace_streaming(lib, src, pairs, chunk) == want, (pairs, chunk)


def test_stream_text_pipelined(lib):
    text = workloads.zipf_text(48 << 20)
    frag = workloads.text_fragment(text[: 4 << 20])
    pairs = [(frag, workloads.expand_replacement()), (b"e ", b"E_"), (b"\n", b"")]
    want = ro.apply_batch(text, pairs)
    st = {}
    got = gpu_replace_streaming(lib, text, pairs, 8 << 20, stats=st)
    assert got == want
    assert st["chunks"] >= 3 * 6 and st["bytes_in"] >= 3 * len(text) * 0.9
