// cr_pattern.h -- host-side classification of a pattern into a suppression mode.
#pragma once
#include "cr_internal.h"

#include <vector>

namespace cr {

constexpr int kModeEmpty = 4;   // empty pattern (bytes.replace inserts at every gap)

// A pattern is border-free iff its longest proper border (KMP failure value) is 0;
// then two occurrences can never overlap, so every candidate is kept.
inline bool border_free(const uint8_t *pat, int L)
{
    if (L <= 1)
        return true;
    std::vector<int> f((size_t)L, 0);
    int k = 0;
    for (int i = 1; i < L; ++i) {
        while (k > 0 && pat[i] != pat[k])
            k = f[(size_t)k - 1];
        if (pat[i] == pat[k])
            ++k;
        f[(size_t)i] = k;
    }
    return f[(size_t)L - 1] == 0;
}

// Border-free patterns up to kMaxSinglePassPat bytes use counts; self-overlapping
// ones up to kMaxTablePat bytes use carry tables of *table_size (the next power of
// two >= L) states; everything else takes the bitmap path.
inline int pattern_mode(const uint8_t *pat, int L, int *table_size)
{
    if (L == 0)
        return kModeEmpty;
    const bool bf = border_free(pat, L);
    if (bf && L <= kMaxSinglePassPat)
        return kModeBorderFree;
    if (!bf && L <= kMaxTablePat) {
        int lt = 2;
        while (lt < L)
            lt <<= 1;
        if (table_size)
            *table_size = lt;
        return kModeTable;
    }
    return kModeBitmap;
}

// First min(L, 4) pattern bytes little-endian, and their byte mask.
inline void pattern_head(const uint8_t *pat, int L, uint32_t *pat4, uint32_t *mask4)
{
    uint32_t p = 0, m = 0;
    for (int i = 0; i < 4 && i < L; ++i) {
        p |= (uint32_t)pat[i] << (8 * i);
        m |= 0xffu << (8 * i);
    }
    *pat4 = p;
    *mask4 = m;
}

// Masked-SAD words of a pattern of up to 32 bytes: c is the smallest byte value that
// does not occur in the pattern, *qx is c in every byte and qref[q] holds pattern
// bytes 4q..4q+3 xor c (never zero), with zero bytes past the pattern.  A masked SAD
// of (data xor *qx) against qref[q] skips exactly the bytes past the pattern.
inline void pattern_qsad(const uint8_t *pat, int L, uint32_t qref[8], uint32_t *qx)
{
    bool used[256] = {false};
    const int m = L < 32 ? L : 32;
    for (int i = 0; i < m; ++i)
        used[pat[i]] = true;
    int c = 0;
    while (c < 255 && used[c])
        ++c;
    *qx = (uint32_t)c * 0x01010101u;
    for (int q = 0; q < 8; ++q) {
        uint32_t w = 0;
        for (int i = 0; i < 4; ++i) {
            const int k = 4 * q + i;
            if (k < m)
                w |= (uint32_t)(pat[k] ^ (uint8_t)c) << (8 * i);
        }
        qref[q] = w;
    }
}

} // namespace cr
