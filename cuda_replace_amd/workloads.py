"""Seeded synthetic inputs for the benchmark configurations of the replace path.

Every generator is deterministic for a given seed (numpy PCG64 streams), so the
GPU box regenerates byte-identical inputs and the committed sha256 fixtures in
``tests/golden/`` apply there too.

Workloads (sizes are the full benchmark sizes; tests pass smaller ``n``):

* ``uniform_abcd``  -- 64 MiB, each byte uniform over ``b"abcd"`` (seed 0);
  ``b"ab" -> b"X"`` (shrink, ~1/16 match density).
* ``all_a``         -- ``b"a" * n``; ``b"aa" -> b"X"`` (maximal overlap).
* ``zipf_text``     -- lowercase words (50k-word seeded vocabulary, lengths
  2-12, Zipf(1.1) rank frequencies), separated by space (90%) or newline (10%).
* ``text_fragment`` -- the 4-byte fragment of a frequent vocabulary word whose
  density in a 4 MiB sample is closest to 1%; the expand replacement is 12
  seeded uppercase bytes, the shrink replacement one byte.
* ``uneven_documents`` -- 8192 documents of the text with log-normal sizes
  (median 64 KiB, sigma 1.5, clipped to 1 KiB - 32 MiB), each with 1-16 pairs
  drawn from ``pair_pool``: patterns are 1-16 byte slices of the text at seeded
  positions (so they occur, some self-overlapping), replacements 0-32 seeded
  bytes of ``A-Z0-9``; a pair whose replacement would grow the sample by more
  than 5% gets a replacement no longer than its pattern.
* ``entity_pairs`` / ``entity_text`` -- the 128-pair session batch: patterns
  ``b"&" + 3-6 seeded lowercase letters + b";"`` (distinct), replacements 1-4
  seeded bytes of ``A-Z0-9``, applied in order; the text gets non-overlapping
  pattern occurrences starting at ~0.5% of its positions (pattern uniform over
  the list) and NUL bytes at ~0.01% of them.
"""
from __future__ import annotations

import functools

import numpy as np

KiB = 1 << 10
MiB = 1 << 20
GiB = 1 << 30

_VOCAB_SIZE = 50_000
_ZIPF_S = 1.1
_SEP_SPACE_FRAC = 0.9
_INV_BITS = 22


def uniform_abcd(n: int = 64 * MiB, seed: int = 0) -> bytes:
    rng = np.random.default_rng(seed)
    lut = np.frombuffer(b"abcd", dtype=np.uint8)
    return lut[rng.integers(0, 4, size=n, dtype=np.uint8)].tobytes()


def all_a(n: int) -> bytes:
    return b"a" * n


@functools.lru_cache(maxsize=4)
def _vocab(seed: int):
    rng = np.random.default_rng([seed, 0x5EED])
    lens = rng.integers(2, 13, size=_VOCAB_SIZE)              # 2..12 letters
    letters = rng.integers(ord("a"), ord("z") + 1, size=(_VOCAB_SIZE, 13), dtype=np.uint8)
    ranks = np.arange(1, _VOCAB_SIZE + 1, dtype=np.float64)
    w = ranks ** (-_ZIPF_S)
    cdf = np.cumsum(w)
    cdf /= cdf[-1]
    # inverse-CDF table: a uniform 22-bit draw u maps to the first rank whose
    # cumulative weight exceeds (u + 0.5) / 2^22
    grid = (np.arange(1 << _INV_BITS, dtype=np.float64) + 0.5) / (1 << _INV_BITS)
    inv = np.minimum(np.searchsorted(cdf, grid, side="right"), _VOCAB_SIZE - 1).astype(np.int32)
    return lens, letters, inv


def zipf_text(n: int, seed: int = 2, chunk_words: int = 4_000_000) -> bytes:
    """Exactly ``n`` bytes of seeded Zipf-frequency text."""
    lens, letters, inv = _vocab(seed)
    rng = np.random.default_rng([seed, 0x7E47])
    out = bytearray()
    col = np.arange(13)
    while len(out) < n:
        m = chunk_words
        idx = inv[rng.integers(0, 1 << _INV_BITS, size=m, dtype=np.int32)]
        wl = lens[idx]
        mat = letters[idx]                                   # (m, 13)
        sep = np.where(rng.integers(0, 10, size=m, dtype=np.int8) < 9, 32, 10).astype(np.uint8)
        mat[np.arange(m), wl] = sep
        mask = col[None, :] <= wl[:, None]
        out += mat[mask].tobytes()
    del out[n:]
    return bytes(out)


def vocab_words(seed: int = 2, top: int = 400):
    """The ``top`` most frequent vocabulary words (rank order)."""
    lens, letters, _ = _vocab(seed)
    return [letters[i, :lens[i]].tobytes() for i in range(top)]


def text_fragment(sample: bytes, seed: int = 2, target: float = 0.01) -> bytes:
    """4-byte fragment of a frequent word with density closest to ``target``."""
    best, best_err = None, None
    seen = set()
    for w in vocab_words(seed):
        for j in range(len(w) - 3):
            f = w[j:j + 4]
            if f in seen:
                continue
            seen.add(f)
            d = sample.count(f) / max(1, len(sample))
            err = abs(d - target)
            if best_err is None or err < best_err:
                best, best_err = f, err
    return best


def expand_replacement(seed: int = 2, length: int = 12) -> bytes:
    rng = np.random.default_rng([seed, 0xE7A0])
    return rng.integers(ord("A"), ord("Z") + 1, size=length, dtype=np.uint8).tobytes()


def text_config(n: int, seed: int = 2):
    """(text, fragment, expand_rep, shrink_rep) for the 1 GiB text workload."""
    text = zipf_text(n, seed)
    frag = text_fragment(text[: 4 * MiB], seed)
    return text, frag, expand_replacement(seed), b"#"


_REP_ALPHABET = np.frombuffer(b"ABCDEFGHIJKLMNOPQRSTUVWXYZ0123456789", dtype=np.uint8)


def pair_pool(sample: bytes, count: int = 256, seed: int = 4, max_growth: float = 0.05):
    """Seeded (pattern, replacement) pool over ``sample`` (see module docstring)."""
    rng = np.random.default_rng([seed, 0xBA7C])
    pool = []
    while len(pool) < count:
        L = int(rng.integers(1, 17))
        pos = int(rng.integers(0, len(sample) - L))
        pat = sample[pos:pos + L]
        r = int(rng.integers(0, 33))
        density = sample.count(pat) / max(1, len(sample))
        if density * max(0, r - L) > max_growth:
            r = int(rng.integers(0, L + 1))
        rep = _REP_ALPHABET[rng.integers(0, len(_REP_ALPHABET), size=r)].tobytes()
        pool.append((pat, rep))
    return pool


def uneven_documents(num_docs: int = 8192, median: int = 64 * KiB, sigma: float = 1.5, lo: int = KiB,
                     hi: int = 32 * MiB, seed: int = 4, text_seed: int = 2, max_pairs: int = 16,
                     pool_size: int = 256):
    """(docs, pairs): the batched-documents workload (see module docstring)."""
    rng = np.random.default_rng([seed, 0xD0C5])
    sizes = np.clip(np.rint(rng.lognormal(np.log(median), sigma, num_docs)), lo, hi).astype(np.int64)
    text = zipf_text(int(sizes.sum()), text_seed)
    offs = np.concatenate([[0], np.cumsum(sizes)])
    docs = [text[int(offs[i]):int(offs[i + 1])] for i in range(num_docs)]
    pool = pair_pool(text[:4 * MiB], pool_size, seed)
    counts = rng.integers(1, max_pairs + 1, size=num_docs)
    pairs = [[pool[int(k)] for k in rng.integers(0, pool_size, size=int(c))] for c in counts]
    return docs, pairs


def entity_pairs(count: int = 128, seed: int = 3):
    """``count`` distinct (entity pattern, replacement) pairs (see module docstring)."""
    rng = np.random.default_rng([seed, 0xE471])
    seen, pairs = set(), []
    while len(pairs) < count:
        k = int(rng.integers(3, 7))
        pat = b"&" + rng.integers(ord("a"), ord("z") + 1, size=k, dtype=np.uint8).tobytes() + b";"
        if pat in seen:
            continue
        seen.add(pat)
        r = int(rng.integers(1, 5))
        pairs.append((pat, _REP_ALPHABET[rng.integers(0, len(_REP_ALPHABET), size=r)].tobytes()))
    return pairs


def entity_text(n: int, pairs, density: float = 0.005, nul_density: float = 0.0001, seed: int = 3,
                text_seed: int = 2) -> bytes:
    """``n`` bytes of the text with the pairs' patterns and NULs planted (see module docstring)."""
    buf = np.frombuffer(zipf_text(n, text_seed), dtype=np.uint8).copy()
    rng = np.random.default_rng([seed, 0x7E47E])
    nuls = rng.integers(0, max(1, n), size=int(n * nul_density))
    buf[nuls] = 0
    step = int(round(1 / density))
    longest = max(len(p) for p, _ in pairs)
    if n >= step + longest:
        # one occurrence per step-byte slot, at a seeded offset that keeps it inside the slot
        starts = np.arange(0, n - step - longest + 1, step, dtype=np.int64)
        starts += rng.integers(0, step - longest, size=len(starts))
        which = rng.integers(0, len(pairs), size=len(starts))
        for k, (pat, _) in enumerate(pairs):
            at = starts[which == k]
            if len(at):
                buf[at[:, None] + np.arange(len(pat))[None, :]] = np.frombuffer(pat, dtype=np.uint8)
    return buf.tobytes()
