/*
 * cr_oracle.c -- CPU restatement of the byte-pattern replace semantics.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library in
 * cuda_replace_amd/.  Only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py may load it.  The product library never links or calls it.
 *
 * Semantics restated (all citations are into the reference's cuda_replace.cu):
 *   - match definition with bounds check ......... cuda_replace.cu:119-129
 *   - leftmost greedy with an initial "last" seed . cuda_replace.cu:938-968
 *     (keep candidate i iff i - last >= L, then last = i)
 *   - output layout: a kept start emits the replacement and skips L source
 *     bytes, every other uncovered byte is copied ... cuda_replace.cu:1391-1467
 *     (shrink) and cuda_replace.cu:1520-1559 (expand); shrink and expand are
 *     one operation (cuda_replace.cu:2097-2159).
 *   - prefix query: count and last of kept starts below a limit, in the
 *     coordinates of the source of the most recent apply .. cuda_replace.cu:1825-1897
 *     (limit clamped to the source length).
 *   - empty pattern: Python bytes.replace semantics (the replacement is
 *     inserted at all n+1 gaps).  The reference returns the input unchanged
 *     (cuda_replace.cu:2104-2109); the stated contract is bytes.replace.
 *
 * Pinned by tests/test_oracle.py against CPython's bytes.replace / bytes.find
 * on the reference's own generated cases (test/test.py:44-99, README.md:40-42,
 * cuda_replace_wrapper.py:379-392), by brute force over every short input on a
 * three-letter alphabet with every seed, and by the closed forms for periodic
 * inputs.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define CRO_NO_SEED INT64_MIN

/* cuda_replace.cu:119-129: pattern fits and every byte equals. */
static int cro_is_match(const uint8_t *s, int64_t n, int64_t pos,
                        const uint8_t *p, int64_t L)
{
    if (L > n || pos + L > n)
        return 0;
    return memcmp(s + pos, p, (size_t)L) == 0;
}

/* Seed -> first position where a candidate may be kept.  A seed "last" blocks
 * every candidate i with i - last < L, i.e. every i < last + L
 * (cuda_replace.cu:947-953, negative seeds honoured as in the legacy path). */
static int64_t cro_floor(int64_t seed, int64_t L)
{
    if (seed == CRO_NO_SEED)
        return 0;
    if (seed > INT64_MAX - L)
        return INT64_MAX;
    int64_t f = seed + L;
    return f < 0 ? 0 : f;
}

/* Greedy kept starts (cuda_replace.cu:938-968).  Writes up to cap positions
 * into kept (may be NULL), returns the total number kept.  L must be > 0. */
int64_t cro_kept(const uint8_t *s, int64_t n, const uint8_t *p, int64_t L,
                 int64_t seed, int64_t *kept, int64_t cap)
{
    if (L <= 0)
        return 0;
    int64_t count = 0;
    int have_last = seed != CRO_NO_SEED;
    int64_t last = seed;
    for (int64_t i = 0; i + L <= n; ++i) {
        if (!cro_is_match(s, n, i, p, L))
            continue;
        if (!have_last || i - last >= L) {
            if (kept && count < cap)
                kept[count] = i;
            ++count;
            last = i;
            have_last = 1;
            /* Keep scanning from i+1: the i - last test above rejects the
             * overlapping candidates exactly as the reference loop does. */
        }
    }
    return count;
}

/* Output length for a given kept count (cuda_replace.cu:2052: n + kept*diff). */
static int64_t cro_out_len(int64_t n, int64_t kept, int64_t L, int64_t r)
{
    return n + kept * (r - L);
}

/* One apply.  Returns the output length, or -1 if out_cap is too small.
 * *n_kept and *last_kept (max kept start or -1) are optional. */
int64_t cro_apply(const uint8_t *s, int64_t n, const uint8_t *p, int64_t L,
                  const uint8_t *r, int64_t rl, int64_t seed,
                  uint8_t *out, int64_t out_cap,
                  int64_t *n_kept, int64_t *last_kept)
{
    if (L == 0) {
        /* bytes.replace(b"", R): R before every byte and once at the end. */
        int64_t total = n + (n + 1) * rl;
        if (n_kept)
            *n_kept = n + 1;
        if (last_kept)
            *last_kept = n;
        if (total > out_cap)
            return -1;
        int64_t o = 0;
        for (int64_t i = 0; i <= n; ++i) {
            memcpy(out + o, r, (size_t)rl);
            o += rl;
            if (i < n)
                out[o++] = s[i];
        }
        return total;
    }
    int64_t f = cro_floor(seed, L);
    int64_t o = 0, kept = 0, last = -1;
    int64_t i = 0;
    while (i < n) {
        if (i >= f && cro_is_match(s, n, i, p, L)) {
            if (o + rl > out_cap)
                return -1;
            memcpy(out + o, r, (size_t)rl);
            o += rl;
            last = i;
            ++kept;
            i += L; /* the next L-1 positions are covered */
        } else {
            if (o + 1 > out_cap)
                return -1;
            out[o++] = s[i++];
        }
    }
    if (n_kept)
        *n_kept = kept;
    if (last_kept)
        *last_kept = last;
    return o;
}

/* Output length without writing (used to size buffers). */
int64_t cro_apply_len(const uint8_t *s, int64_t n, const uint8_t *p, int64_t L,
                      int64_t rl, int64_t seed)
{
    if (L == 0)
        return n + (n + 1) * rl;
    int64_t k = cro_kept(s, n, p, L, seed, NULL, 0);
    return cro_out_len(n, k, L, rl);
}

/* Prefix query over the kept starts of an apply of (p, seed) on s
 * (cuda_replace.cu:1825-1897, limit clamped to n). */
void cro_query_prefix(const uint8_t *s, int64_t n, const uint8_t *p, int64_t L,
                      int64_t seed, int64_t limit,
                      int64_t *count, int64_t *last)
{
    if (limit > n)
        limit = n;
    int64_t c = 0, lst = -1;
    if (L == 0) {
        /* every gap 0..n is a kept start */
        c = limit > 0 ? limit : 0;
        lst = limit > 0 ? limit - 1 : -1;
    } else {
        int64_t f = cro_floor(seed, L);
        int64_t i = 0;
        while (i < limit) {
            if (i >= f && cro_is_match(s, n, i, p, L)) {
                ++c;
                lst = i;
                i += L;
            } else {
                ++i;
            }
        }
    }
    if (count)
        *count = c;
    if (last)
        *last = lst;
}
