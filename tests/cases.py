"""Input definitions shared by the fixture script and the tests.

* ``REFERENCE_CASES``: the six generated cases of the reference's benchmark
  (test/test.py:44-99), same data, pattern and replacement.
* ``KNOWN_ANSWERS``: the README example (README.md:40-42) and the wrapper's
  self-test (cuda_replace_wrapper.py:379-392), with outputs as printed there.
* ``REGRESSIONS``: inputs on which the reference as read diverges from
  bytes.replace (fused-tile stale carry, end-of-buffer overhang, negative seed,
  streaming straddle, empty pattern); the expected value is bytes.replace.
"""

MB = 1024 * 1024


def _needle():
    base = b"x" * (10 * MB)
    return base[:len(base) // 2] + b"NEEDLE" + base[len(base) // 2:]


REFERENCE_CASES = [
    # name, data factory, pattern, replacement       (test/test.py:44-99)
    ("dense_small_1mb", lambda: b"hello " * (MB // 6), b"hello", b"hi"),
    ("sparse_large_10mb", _needle, b"NEEDLE", b"FOUND"),
    ("expansion_5mb", lambda: b"abc" * (5 * MB // 3), b"abc", b"ABCDEF"),
    ("medium_20mb", lambda: (b"normal text " * 100 + b"REPLACE ") * (20 * 1024 // 12), b"REPLACE", b"DONE"),
    ("large_dense_50mb", lambda: b"aaabaaabaaab" * (50 * MB // 12), b"aaa", b"X"),
    ("huge_sparse_100mb", lambda: b"x" * (100 * MB), b"xx", b"y"),
]

KNOWN_ANSWERS = [
    # README.md:40-42
    ("readme_hello", b"hello world hello universe", b"hello", b"goodbye",
     b"goodbye world goodbye universe"),
    # cuda_replace_wrapper.py:379-392 (printed, expected from bytes.replace)
    ("wrapper_selftest", b"aaabaaabaaa", b"aaa", b"X", b"XbXbX"),
]

# (name, src, pat, rep) -- expected output is src.replace(pat, rep)
REGRESSIONS = [
    # stale carry between 8 KiB fused tiles for a self-overlapping pattern
    ("stale_tile_carry_aaa", b"a" * 16388, b"aaa", b"X"),
    ("stale_tile_carry_aaa_long", b"a" * (3 * 8192 + 7), b"aaa", b"XY"),
    # pattern with a trailing NUL must not match past the end of the buffer
    ("nul_overhang_shrink", b"xa", b"a\x00", b"Z"),
    ("nul_overhang_expand", b"xa", b"a\x00", b"ZZZ"),
    # empty pattern inserts the replacement at every gap
    ("empty_pattern", b"abc", b"", b"X"),
    ("empty_pattern_empty_src", b"", b"", b"X"),
    ("empty_pattern_empty_rep", b"abc", b"", b""),
    # pattern longer than the input
    ("pattern_longer", b"ab", b"abc", b"X"),
    ("empty_src", b"", b"a", b"XYZ"),
]

# streaming straddle: (name, src, pat, rep, chunk_bytes)
STREAMING_REGRESSIONS = [
    ("straddle_aaaa_chunk3", b"aaaa", b"aa", b"X", 3),
    ("straddle_ab_chunk1", b"ab", b"ab", b"X", 1),
    ("wrapper_selftest_chunk4", b"aaabaaabaaa", b"aaa", b"X", 4),
]
