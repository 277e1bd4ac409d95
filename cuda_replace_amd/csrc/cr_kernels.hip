// cr_kernels.hip -- HIP kernels for leftmost non-overlapping byte replace on gfx950.
//
// The source is cut into 8 KiB tiles; kGroupTiles consecutive tiles form a group,
// handled by one 256-thread workgroup (4 wave64, 32 source bytes per thread); runs of
// consecutive groups form chunks (by default one chunk holds the whole pass).  Each
// chunk is processed by three launches on the session stream (reduce-then-scan):
//   count  stages the group in LDS (global_load_lds, 16 B per lane), builds each
//          thread's 32-position candidate mask with byte-aligned dword compares, and
//          writes the group's summary: its kept count for a border-free pattern, its
//          carry table (carry in -> carry out, kept) for a self-overlapping one;
//   spine  one workgroup scans the chunk's summaries into each group's exclusive
//          state (kept starts before it, carry into it);
//   emit   stages the group again, resolves the kept starts of each tile from the
//          group's state, assembles the tile's output in LDS and writes it with
//          16-byte stores.
// No workgroup ever waits on another: launch boundaries order the chunks.
//
// Replaces the reference's mark / suppress / scan / scatter chain
// (cuda_replace.cu:737-866 fused mark+suppress, 905-968 serial greedy, 237-334 three-stage
// scan, 1391-1608 shrink-pad / expansion scatter, 2097-2159 apply_once).
//
// Suppression modes:
//   border-free pattern (no proper prefix equals a suffix): occurrences never overlap,
//     so every candidate at or above the seed floor is kept; the carry into a group is
//     recomputed from the L-1 bytes left of it.
//   self-overlapping pattern, L <= 16: per-thread tables over the L carry states,
//     composed in position order (wave64 shuffle scans, then across groups).
//   otherwise: candidate bitmap, one-wave serial greedy, then the emit kernel scatters.
#include "cr_internal.h"

#include <algorithm>
#include <map>
#include <mutex>
#include <type_traits>

#include <limits.h>
#include <stdio.h>
#include <stdlib.h>

namespace cr {

constexpr int kWave = 64;
constexpr int kWaves = kThreads / kWave;
constexpr int kNegInf = INT_MIN / 4;
// Replacements up to this length are staged in LDS; longer ones are read from HBM.
constexpr int kRepLds = 256;

__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }

// Greedy over a 32-position mask with the first `c` positions blocked.
// Returns the kept mask; *last gets the last kept position or kNegInf.
__device__ __forceinline__ uint32_t greedy32(uint32_t m, int c, int L, int *last)
{
    uint32_t mm = c >= 32 ? 0u : (m & (0xffffffffu << c));
    uint32_t kept = 0;
    int lst = kNegInf;
    while (mm) {
        int q = __builtin_ctz(mm);
        kept |= 1u << q;
        lst = q;
        int s = q + L;
        mm = s >= 32 ? 0u : (mm & (0xffffffffu << s));
    }
    *last = lst;
    return kept;
}

// The same greedy for a self-overlapping 2-byte pattern (its two bytes are equal,
// so occurrences overlap by one): in every run of consecutive candidates the kept
// starts are the candidates at an even distance from the run's first one.  Adding
// the odd-positioned run starts to x carries through exactly the runs that start
// at odd positions.
__device__ __forceinline__ uint32_t greedy32_pair(uint32_t m, int c, int *last)
{
    const uint32_t x = c >= 32 ? 0u : (m & (0xffffffffu << c));
    const uint32_t starts = x & ~(x << 1);
    const uint32_t odd_runs = ((x + (starts & 0xAAAAAAAAu)) ^ x) & x;
    const uint32_t kept = ((x & ~odd_runs) & 0x55555555u) | (odd_runs & 0xAAAAAAAAu);
    *last = kept ? 31 - __builtin_clz(kept) : kNegInf;
    return kept;
}

// ---- carry tables (self-overlapping patterns) -------------------------------
// entry = carry_out << 24 | kept_count ; carry < 16, count <= 8192 inside a tile
template <int LT>
struct Table {
    uint32_t e[LT];
};

// Entry idx (< LT) of a table: a binary tree of selects on the index bits, level
// by level.  Written as a chain of equality selects, the compiler turns the lookup
// into an indexed array in scratch memory for LT >= 4.
template <int LT, class T>
__device__ __forceinline__ T tree_select(const T (&e)[LT], uint32_t idx)
{
    T v[LT];
#pragma unroll
    for (int i = 0; i < LT; ++i)
        v[i] = e[i];
    constexpr int kLevels = LT >= 16 ? 4 : LT >= 8 ? 3 : LT >= 4 ? 2 : LT >= 2 ? 1 : 0;
#pragma unroll
    for (int b = 0; b < kLevels; ++b) {
        // a bitwise select: a select of two table entries is also turned into an
        // indexed load from a stack copy of the table
        const T m = (T)0 - (T)((idx >> b) & 1u);
#pragma unroll
        for (int i = 0; i < (LT >> (b + 1)); ++i)
            v[i] = v[2 * i] ^ ((v[2 * i] ^ v[2 * i + 1]) & m);
    }
    return v[0];
}

template <int LT>
__device__ __forceinline__ uint32_t tsel(const Table<LT> &t, uint32_t idx)
{
    return tree_select<LT, uint32_t>(t.e, idx);
}

// apply `first` then `second`
template <int LT>
__device__ __forceinline__ Table<LT> tcompose(const Table<LT> &first, const Table<LT> &second)
{
    Table<LT> r;
#pragma unroll
    for (int c = 0; c < LT; ++c) {
        uint32_t a = first.e[c];
        uint32_t b = tsel<LT>(second, a >> 24);
        r.e[c] = (b & 0xff000000u) | ((a & 0xffffffu) + (b & 0xffffffu));
    }
    return r;
}

template <int LT>
__device__ __forceinline__ Table<LT> tidentity()
{
    Table<LT> r;
#pragma unroll
    for (int c = 0; c < LT; ++c)
        r.e[c] = (uint32_t)c << 24;
    return r;
}

// ---- LDS / HBM copy helpers ---------------------------------------------------
// 16 bytes starting at byte offset `off` of an LDS buffer (4-byte aligned base):
// five aligned dword reads and a byte align.
__device__ __forceinline__ uint4 lds_read16(const uint8_t *base, int off)
{
    const uint32_t *d = reinterpret_cast<const uint32_t *>(base + (off & ~3));
    const uint32_t sh = (uint32_t)(off & 3);
    const uint32_t d0 = d[0], d1 = d[1], d2 = d[2], d3 = d[3], d4 = d[4];
    return make_uint4(__builtin_amdgcn_alignbyte(d1, d0, sh), __builtin_amdgcn_alignbyte(d2, d1, sh),
                      __builtin_amdgcn_alignbyte(d3, d2, sh), __builtin_amdgcn_alignbyte(d4, d3, sh));
}

// Stores the first `len` (< 16) bytes of v at dst (any alignment), exactly.
__device__ __forceinline__ void store_tail(uint8_t *dst, uint4 v, int len)
{
    uint64_t lo = (uint64_t)v.x | ((uint64_t)v.y << 32);
    const uint64_t hi = (uint64_t)v.z | ((uint64_t)v.w << 32);
    if (len & 8) {
        __builtin_memcpy(dst, &lo, 8);
        dst += 8;
        lo = hi;
    }
    if (len & 4) {
        const uint32_t w = (uint32_t)lo;
        __builtin_memcpy(dst, &w, 4);
        dst += 4;
        lo >>= 32;
    }
    if (len & 2) {
        const uint16_t h = (uint16_t)lo;
        __builtin_memcpy(dst, &h, 2);
        dst += 2;
        lo >>= 16;
    }
    if (len & 1)
        *dst = (uint8_t)lo;
}

// Copies len bytes from LDS (byte offset src) to HBM at dst: unaligned 16-byte
// stores (full speed on gfx950), then an exact tail.  Reads up to 19 bytes past the
// source range.
__device__ __forceinline__ void store_from_lds(uint8_t *dst, const uint8_t *lds, int src, int len)
{
    for (; len >= 16; len -= 16, src += 16, dst += 16) {
        const uint4 v = lds_read16(lds, src);
        __builtin_memcpy(dst, &v, 16);
    }
    if (len > 0)
        store_tail(dst, lds_read16(lds, src), len);
}

// Unaligned 16-byte LDS accesses: gfx950 runs with unaligned LDS access enabled, so
// these compile to single ds_read_b128 / ds_write_b128 at any byte address.
__device__ __forceinline__ uint4 lds_load16(const uint8_t *p)
{
    uint4 v;
    __builtin_memcpy(&v, p, 16);
    return v;
}
__device__ __forceinline__ void lds_store16(uint8_t *p, uint4 v) { __builtin_memcpy(p, &v, 16); }

// Copies len bytes between LDS buffers (any alignment): 16-byte moves, then an exact
// tail.  Reads up to 15 bytes past the source range.
__device__ __forceinline__ void lds_move(uint8_t *dst, const uint8_t *src, int len)
{
    for (; len >= 16; len -= 16, src += 16, dst += 16)
        lds_store16(dst, lds_load16(src));
    if (len > 0)
        store_tail(dst, lds_load16(src), len);
}

// Copies len bytes HBM -> HBM with unaligned 16-byte accesses (long replacements).
__device__ __forceinline__ void store_from_global(uint8_t *dst, const uint8_t *src, int64_t len)
{
    for (; len >= 16; len -= 16, src += 16, dst += 16) {
        uint4 v;
        __builtin_memcpy(&v, src, 16);
        __builtin_memcpy(dst, &v, 16);
    }
    for (int64_t i = 0; i < len; ++i)
        dst[i] = src[i];
}

// 64-bit carry tables for composing many group tables: entry = carry_out << 40 | count.
template <int LT>
struct Table64 {
    uint64_t e[LT];
};

template <int LT>
__device__ __forceinline__ uint64_t tsel64(const Table64<LT> &t, uint32_t idx)
{
    return tree_select<LT, uint64_t>(t.e, idx);
}

// apply `first` then `second`
template <int LT>
__device__ __forceinline__ Table64<LT> tcompose64(const Table64<LT> &first, const Table64<LT> &second)
{
    Table64<LT> r;
#pragma unroll
    for (int c = 0; c < LT; ++c) {
        const uint64_t a = first.e[c];
        const uint64_t b = tsel64<LT>(second, (uint32_t)(a >> kCarryShift));
        r.e[c] = (b & ~kCountMask) | ((a & kCountMask) + (b & kCountMask));
    }
    return r;
}

template <int LT>
__device__ __forceinline__ Table64<LT> tidentity64()
{
    Table64<LT> r;
#pragma unroll
    for (int c = 0; c < LT; ++c)
        r.e[c] = (uint64_t)c << kCarryShift;
    return r;
}

template <int LT>
__device__ __forceinline__ Table64<LT> widen(const Table<LT> &t)
{
    Table64<LT> r;
#pragma unroll
    for (int c = 0; c < LT; ++c)
        r.e[c] = ((uint64_t)(t.e[c] >> 24) << kCarryShift) | (uint64_t)(t.e[c] & 0xffffffu);
    return r;
}

// Ordered composition across the wave: lane 0 ends with lane0 ∘ lane1 ∘ ... ∘ lane63.
template <int LT>
__device__ __forceinline__ Table64<LT> wave_compose(Table64<LT> t, int lane)
{
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        Table64<LT> o;
#pragma unroll
        for (int c = 0; c < LT; ++c)
            o.e[c] = __shfl_down(t.e[c], d, kWave);
        if ((lane & (2 * d - 1)) == 0 && lane + d < kWave)
            t = tcompose64<LT>(t, o);
    }
    return t;
}

template <int LT>
__device__ __forceinline__ Table64<LT> wave_broadcast_last(const Table64<LT> &t)
{
    Table64<LT> r;
#pragma unroll
    for (int c = 0; c < LT; ++c)
        r.e[c] = __shfl(t.e[c], kWave - 1, kWave);
    return r;
}

template <int LT>
__device__ __forceinline__ Table64<LT> wave_broadcast0(const Table64<LT> &t)
{
    Table64<LT> r;
#pragma unroll
    for (int c = 0; c < LT; ++c)
        r.e[c] = __shfl(t.e[c], 0, kWave);
    return r;
}

// Ordered reduction of 32-bit carry tables across the wave: lane 0 ends with
// lane0 ∘ lane1 ∘ ... ∘ lane63.
template <int LT>
__device__ __forceinline__ Table<LT> wave_reduce_table(Table<LT> t, int lane)
{
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        Table<LT> o;
#pragma unroll
        for (int c = 0; c < LT; ++c)
            o.e[c] = __shfl_down(t.e[c], d, kWave);
        if ((lane & (2 * d - 1)) == 0 && lane + d < kWave)
            t = tcompose<LT>(t, o);
    }
    return t;
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v)
{
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1)
        v += __shfl_xor(v, d, kWave);
    return v;
}

// Wave64 inclusive scans with DPP: row shifts by 1, 2, 4 and 8 within each row of 16
// lanes, then the last lane of a row broadcast to the rows after it (row_bcast15 to
// rows 1 and 3, row_bcast31 to rows 2 and 3).  A lane whose DPP source is disabled or
// out of its row receives the operation's identity.
__device__ __forceinline__ uint32_t wave_incl_add_u32(uint32_t x)
{
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, true);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, true);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, true);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, true);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false);
    return x;
}
__device__ __forceinline__ int wave_incl_max_i32(int x)
{
    x = max(x, __builtin_amdgcn_update_dpp(INT_MIN, x, 0x111, 0xf, 0xf, false));
    x = max(x, __builtin_amdgcn_update_dpp(INT_MIN, x, 0x112, 0xf, 0xf, false));
    x = max(x, __builtin_amdgcn_update_dpp(INT_MIN, x, 0x114, 0xf, 0xf, false));
    x = max(x, __builtin_amdgcn_update_dpp(INT_MIN, x, 0x118, 0xf, 0xf, false));
    x = max(x, __builtin_amdgcn_update_dpp(INT_MIN, x, 0x142, 0xa, 0xf, false));
    x = max(x, __builtin_amdgcn_update_dpp(INT_MIN, x, 0x143, 0xc, 0xf, false));
    return x;
}
__device__ __forceinline__ int wave_shr1_i32(int x, int first)
{
    return __builtin_amdgcn_update_dpp(first, x, 0x138, 0xf, 0xf, false);
}

// The same scan over 32-bit carry tables: lane i ends with lane0 ∘ ... ∘ lanei.
// The identity table (carry c -> c, nothing kept) fills the lanes without a source.
template <int LT, int CTRL, int ROWS>
__device__ __forceinline__ Table<LT> dpp_compose_step(const Table<LT> &t)
{
    Table<LT> o;
#pragma unroll
    for (int c = 0; c < LT; ++c)
        o.e[c] = (uint32_t)__builtin_amdgcn_update_dpp((int)((uint32_t)c << 24), (int)t.e[c], CTRL, ROWS, 0xf, false);
    return tcompose<LT>(o, t);
}

template <int LT>
__device__ __forceinline__ Table<LT> wave_incl_table(Table<LT> t)
{
    t = dpp_compose_step<LT, 0x111, 0xf>(t);
    t = dpp_compose_step<LT, 0x112, 0xf>(t);
    t = dpp_compose_step<LT, 0x114, 0xf>(t);
    t = dpp_compose_step<LT, 0x118, 0xf>(t);
    t = dpp_compose_step<LT, 0x142, 0xa>(t);
    t = dpp_compose_step<LT, 0x143, 0xc>(t);
    return t;
}

// The table of the lane before (wave_shr1), the identity on lane 0.
template <int LT>
__device__ __forceinline__ Table<LT> wave_shr1_table(const Table<LT> &t)
{
    Table<LT> o;
#pragma unroll
    for (int c = 0; c < LT; ++c)
        o.e[c] = (uint32_t)__builtin_amdgcn_update_dpp((int)((uint32_t)c << 24), (int)t.e[c], 0x138, 0xf, 0xf, false);
    return o;
}

// ---- group staging ----------------------------------------------------------
// LDS image of a group of GT tiles: [kHalo bytes before | GT tiles | kHalo bytes after].
template <int GT>
struct GroupShape {
    static constexpr int kBytes = GT * kTile;
    static constexpr int kIn = kHalo + kBytes + kHalo;
};

// Issues the group's loads straight into LDS with global_load_lds (no registers):
// every wave streams 1 KiB per instruction of the body; wave 0 fetches the left
// halo and wave 1 the right halo (4 bytes per lane).  The caller waits with
// s_waitcnt + a barrier.  The source is a 16-byte aligned device buffer, so the
// granule that holds the last valid byte is readable; bytes at or past n are never
// matched or emitted.
template <int GT>
__device__ __forceinline__ void load_group(const ReplaceParams &p, int64_t t0, int tid, bool left_halo,
                                           uint8_t *s_in)
{
    constexpr int kGroupBytes = GroupShape<GT>::kBytes;
    typedef __attribute__((address_space(3))) void lds_void;
    const int64_t start = t0 * kTile;
    const int64_t avail = p.n - start;
    const int wave = tid / kWave;
    const int lane = tid & (kWave - 1);
#pragma unroll
    for (int k = 0; k < kGroupBytes / (kWaves * 1024); ++k) {
        const int off = k * (kWaves * 1024) + wave * 1024;
        if (off + lane * 16 < avail)
            __builtin_amdgcn_global_load_lds((const void *)(p.src + start + off + lane * 16),
                                             (lds_void *)(s_in + kHalo + off), 16, 0, 0);
    }
    if (wave == 0 && left_halo && start >= kHalo)
        __builtin_amdgcn_global_load_lds((const void *)(p.src + start - kHalo + lane * 4), (lds_void *)(s_in), 4,
                                         0, 0);
    if (wave == 1 && start + kGroupBytes + lane * 4 < p.n)
        __builtin_amdgcn_global_load_lds((const void *)(p.src + start + kGroupBytes + lane * 4),
                                         (lds_void *)(s_in + kHalo + kGroupBytes), 4, 0, 0);
}

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Occurrence mask of the 32 positions whose bytes start at s_seg (16-byte aligned
// LDS) for a pattern of 4(NQ-1) < L <= 4 NQ bytes: the masked quad SADs of
// bf_mask64_qsad (below) over eight windows of four positions, against the reference words
// qref and the absent byte qx (pattern_qsad).  Reads s_seg[0, 32 + 4 NQ).
template <int NQ>
__device__ __forceinline__ uint32_t mask32_qsad(const uint32_t *qref, uint32_t qx, const uint8_t *s_seg)
{
    uint32_t d[16];
    const u32x4 *q4 = reinterpret_cast<const u32x4 *>(s_seg);
#pragma unroll
    for (int k = 0; k < (8 + NQ + 3) / 4; ++k) {
        const u32x4 v = q4[k];
        d[4 * k] = v.x ^ qx;
        d[4 * k + 1] = v.y ^ qx;
        d[4 * k + 2] = v.z ^ qx;
        d[4 * k + 3] = v.w ^ qx;
    }
    uint32_t m = 0u;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        uint64_t acc = 0;
#pragma unroll
        for (int q = 0; q < NQ; ++q)
            acc = __builtin_amdgcn_mqsad_pk_u16_u8(((uint64_t)d[k + q + 1] << 32) | d[k + q], qref[q], acc);
        const uint32_t tl = (0x80008000u - (uint32_t)acc) & 0x80008000u;
        const uint32_t th = (0x80008000u - (uint32_t)(acc >> 32)) & 0x80008000u;
        m |= (((tl >> 15) | (tl >> 30) | (th >> 13) | (th >> 28)) & 0xfu) << (4 * k);
    }
    return m;
}

// Positions of [gseg, gseg+32) where an occurrence may start: floor <= pos <= n - L.
__device__ __forceinline__ uint32_t range_mask32(const ReplaceParams &p, int64_t gseg)
{
    const int64_t lo = p.floor - gseg;
    const int64_t hi = p.n - (int64_t)p.L - gseg;
    if (hi < 0 || lo > 31 || lo > hi)
        return 0u;
    const int a = lo < 0 ? 0 : (int)lo;
    const int b = hi > 31 ? 31 : (int)hi;
    return (b == 31 ? 0xffffffffu : ((1u << (b + 1)) - 1u)) & (0xffffffffu << a);
}

// The masked-SAD words of the pattern in s_pat (L <= 16; as pattern_qsad on the
// host), computed by every wave for itself: lane v tests byte value v, and the
// smallest value absent from the pattern (below 17) is the first clear bit of the
// ballot.
struct QsadWords {
    uint32_t qref[4];
    uint32_t qx;
};

__device__ __forceinline__ QsadWords qsad_words_lds(const uint8_t *s_pat, int L)
{
    const uint32_t lane = threadIdx.x & (kWave - 1);
    bool present = false;
    for (int i = 0; i < L; ++i)
        present |= s_pat[i] == lane;
    const uint32_t c = (uint32_t)__builtin_ctzll(~__builtin_amdgcn_ballot_w64(present));
    QsadWords q;
    q.qx = c * 0x01010101u;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        uint32_t v = 0;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int k = 4 * w + b;
            if (k < L)
                v |= (uint32_t)(s_pat[k] ^ (uint8_t)c) << (8 * b);
        }
        q.qref[w] = v;
    }
    return q;
}

// Positions of [gseg, gseg+32) (bytes from s_seg, 16-byte aligned LDS) holding the
// byte p.pat4 & 0xff, within [floor, n - L]: the zero-byte test and flag gathering
// of first_byte64 on eight dwords.
__device__ __forceinline__ uint32_t first_byte32(const ReplaceParams &p, const uint8_t *s_seg, int64_t gseg)
{
    const uint32_t f4 = (p.pat4 & 0xffu) * 0x01010101u;
    const uint4 *q = reinterpret_cast<const uint4 *>(s_seg);
    uint32_t m = 0u;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const uint4 v = q[k];
        const uint32_t d[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const uint32_t t = d[i] ^ f4;
            const uint32_t z = ~(((t & 0x7f7f7f7fu) + 0x7f7f7f7fu) | t) & 0x80808080u;
            m |= ((z * 0x00204081u) >> 28) << (4 * (4 * k + i));
        }
    }
    return m & range_mask32(p, gseg);
}

// Candidate mask of the 32 positions [gseg, gseg+32) whose bytes start at s_seg
// (16-byte aligned LDS) for a pattern of up to 16 bytes, from masked quad SADs:
// every position costs the same whatever the data (the byte loop of cand_mask runs
// once per head match, all of them on dense input).
__device__ __forceinline__ uint32_t cand_mask_qsad(const ReplaceParams &p, const QsadWords &qw, const uint8_t *s_seg,
                                                   int64_t gseg)
{
    const uint32_t rm = range_mask32(p, gseg);
    if (!rm)
        return 0u;
    switch ((p.L + 3) >> 2) {
    case 1: return rm & mask32_qsad<1>(qw.qref, qw.qx, s_seg);
    case 2: return rm & mask32_qsad<2>(qw.qref, qw.qx, s_seg);
    case 3: return rm & mask32_qsad<3>(qw.qref, qw.qx, s_seg);
    default: return rm & mask32_qsad<4>(qw.qref, qw.qx, s_seg);
    }
}

// Candidate mask of the 32 positions [gseg, gseg+32) whose bytes start at
// s_tile + seg in LDS: bit j set iff the pattern occurs at gseg + j and
// floor <= gseg + j <= n - L.  Reads up to L - 1 bytes past the segment.
__device__ __forceinline__ uint32_t cand_mask(const ReplaceParams &p, const uint8_t *s_tile, const uint8_t *s_pat,
                                              int seg, int64_t gseg)
{
    const int L = p.L;
    const uint32_t *w32 = reinterpret_cast<const uint32_t *>(s_tile + seg);
    uint32_t d[9];
#pragma unroll
    for (int k = 0; k < 9; ++k)
        d[k] = w32[k];
    const uint32_t pat4 = p.pat4, mask4 = p.mask4;
    // bit j <- window at j equals the pattern head; m + m + cmp is one v_addc per
    // position (the compare's lane mask is the carry in)
    uint32_t m = 0;
#pragma unroll
    for (int j = 31; j >= 0; --j) {
        const uint32_t w = (j % 4 == 0) ? d[j / 4] : __builtin_amdgcn_alignbyte(d[j / 4 + 1], d[j / 4], j % 4);
        m = m + m + (uint32_t)(((w ^ pat4) & mask4) == 0u);
    }
    m &= range_mask32(p, gseg);
    if (L > 4 && m) {
        uint32_t mm = m;
        m = 0;
        const uint8_t *sp = s_tile + seg;
        while (mm) {
            const int j = __builtin_ctz(mm);
            mm &= mm - 1;
            bool ok = true;
            for (int k = 4; k < L; ++k) {
                if (sp[j + k] != s_pat[k]) {
                    ok = false;
                    break;
                }
            }
            m |= (uint32_t)ok << j;
        }
    }
    return m;
}

// Per-thread carry table over carry-in c in [0, LT): the greedy over 32 candidate
// positions with the first c blocked gives (carry into the next segment, kept).
template <int LT>
__device__ __forceinline__ Table<LT> segment_table(uint32_t m, int L)
{
    Table<LT> t;
#pragma unroll
    for (int c = 0; c < LT; ++c) {
        int lst;
        // two carry states: the pattern is 2 bytes long (L <= LT) and self-overlapping
        const uint32_t km = LT == 2 ? greedy32_pair(m, c, &lst) : greedy32(m, c, L, &lst);
        int co = lst + L - 32;
        co = co < 0 ? 0 : co;
        t.e[c] = ((uint32_t)co << 24) | (uint32_t)__builtin_popcount(km);
    }
    return t;
}

// ---- count: one summary per group ------------------------------------------
// Border-free pattern: every candidate at or above the floor is kept, so the
// group's kept count is a popcount.  Self-overlapping pattern: the group's carry
// table (carry in -> carry out, kept), composed tile by tile in position order.
// Workgroups never wait on each other.
template <int MODE, int LT, int GT>
__device__ __forceinline__ void count_group(const ReplaceParams &p, int64_t g, uint32_t *summary)
{
    constexpr int kGroupTiles = GT;
    __shared__ __align__(16) uint8_t s_in[GroupShape<GT>::kIn];
    __shared__ uint8_t s_pat[kMaxSinglePassPat];
    __shared__ uint32_t s_red[kWaves];
    __shared__ uint32_t s_wt[kWaves][LT];

    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wave = tid / kWave;
    const int64_t t0 = g * kGroupTiles;
    if (t0 >= p.num_tiles)
        return;
    for (int i = tid; i < p.L; i += kThreads)
        s_pat[i] = p.pat[i];
    load_group<GT>(p, t0, tid, false, s_in);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();

    const int seg = tid * kBytesPerThread;
    if constexpr (MODE == kModeBorderFree) {
        uint32_t cnt = 0;
#pragma unroll
        for (int i = 0; i < kGroupTiles; ++i) {
            const int64_t ti = t0 + i;
            if (ti < p.num_tiles) {
                const uint32_t m = cand_mask(p, s_in + kHalo + i * kTile, s_pat, seg, ti * kTile + seg);
                if (p.masks)
                    p.masks[ti * (kTile / 32) + tid] = m;
                cnt += __builtin_popcount(m);
            }
        }
        cnt = wave_incl_add_u32(cnt);
        if (lane == kWave - 1)
            s_red[wave] = cnt;
        __syncthreads();
        if (tid == 0)
            summary[0] = s_red[0] + s_red[1] + s_red[2] + s_red[3];
    } else {
        // the summary is the composition of every 32-byte segment's table in position
        // order, so each thread takes GT consecutive segments of the group and
        // composes them itself: one wave reduction and one exchange per group.
        // Segments at or past n have no candidates (cand_mask's range), and the
        // carry out of the input's last group is never used.
        const int b0 = tid * kBytesPerThread * GT;
        const int64_t g0 = t0 * kTile;
        const QsadWords qw = qsad_words_lds(s_pat, p.L);
        // segment masks of tiles that exist are kept for the emit
        auto seg_mask = [&](int b) -> uint32_t {
            const uint32_t m = cand_mask_qsad(p, qw, s_in + kHalo + b, g0 + b);
            if (p.masks && t0 + b / kTile < p.num_tiles)
                p.masks[(g0 + b) / 32] = m;
            return m;
        };
        Table<LT> mine = segment_table<LT>(seg_mask(b0), p.L);
#pragma unroll
        for (int k = 1; k < GT; ++k)
            mine = tcompose<LT>(mine, segment_table<LT>(seg_mask(b0 + k * kBytesPerThread), p.L));
        const Table<LT> t = wave_incl_table<LT>(mine);   // lane 63: the wave's composition
        if (lane == kWave - 1) {
#pragma unroll
            for (int c = 0; c < LT; ++c)
                s_wt[wave][c] = t.e[c];
        }
        __syncthreads();
        Table<LT> grp = tidentity<LT>();
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            Table<LT> wt;
#pragma unroll
            for (int c = 0; c < LT; ++c)
                wt.e[c] = s_wt[w][c];
            grp = tcompose<LT>(grp, wt);
        }
        if (tid < LT)
            summary[tid] = tsel<LT>(grp, (uint32_t)tid);
    }
}

template <int MODE, int LT, int GT>
__global__ void __launch_bounds__(kThreads)
count_kernel(ReplaceParams p, int64_t first_group)
{
    constexpr int W = MODE == kModeTable ? LT : 1;
    const int64_t g = first_group + blockIdx.x;
    count_group<MODE, LT, GT>(p, g, p.gsum + g * W);
}

// ---- spine: exclusive kept state of every group of a chunk -------------------
// One workgroup of kSpineThreads walks the chunk's group summaries in passes of
// kSpineThreads * V groups.  Wave w owns the pass's groups [w*64V, (w+1)*64V) and
// lane l their groups w*64V + 64j + l (j < V), so every load and store of the wave
// is one coalesced run.  Per j the wave scans its 64 summaries in order (shuffles),
// carrying the running state to j+1; the wave totals are then composed across
// waves through LDS and each group's exclusive state (kept before it, carry into
// it) is written.  The state after the chunk goes to chunks[chunk + 1]; the last
// chunk also publishes the pass's kept total.
constexpr int kSpineThreads = 1024;
constexpr int kSpineWaves = kSpineThreads / kWave;

template <int W>
struct SpineShape {
    static constexpr int V = W == 1 ? 16 : (16 / W > 1 ? 16 / W : 1);   // groups per lane and pass
};

// Groups per lane and step of a blocked spine: carry tables take one group per lane
// (a block is a single short table step per wave, so a pass spreads over many
// workgroups: 2-state all-'a' pass 0.774 -> 0.766 ms); kept counts keep 16 per lane
// (1024-group blocks measured no faster there).
template <int MODE, int LT>
constexpr int spine_block_v()
{
    return MODE == kModeTable ? 1 : SpineShape<1>::V;
}
template <int MODE, int LT>
constexpr int64_t spine_block_groups()
{
    return (int64_t)kWave * spine_block_v<MODE, LT>() * kSpineWaves;
}

// BLOCKS: one workgroup per block of spine_block_groups() groups (block = chunk =
// blockIdx.x, its entering state from chunks[]), the pass total published by the
// block scan instead.
template <int MODE, int LT, bool BLOCKS = false>
__global__ void __launch_bounds__(kSpineThreads)
spine_kernel(ReplaceParams p, int64_t first_group, int64_t groups, int64_t chunk, int last_chunk)
{
    constexpr int W = MODE == kModeTable ? LT : 1;   // words per group summary
    constexpr int V = BLOCKS ? spine_block_v<MODE, LT>() : SpineShape<W>::V;
    constexpr int64_t kBlock = (int64_t)kWave * V;   // groups per wave and pass
    constexpr int64_t kPass = kBlock * kSpineWaves;
    __shared__ uint64_t s_wt[kSpineWaves][LT];
    __shared__ int64_t s_wsum[kSpineWaves];

    if constexpr (BLOCKS) {
        chunk = blockIdx.x;
        const int64_t total = groups;
        first_group = chunk * kPass;
        groups = total - first_group < kPass ? total - first_group : kPass;
    }
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wave = tid / kWave;
    int64_t k = 0;    // kept before the current pass
    int carry = 0;    // carry into the current pass
    if (chunk > 0) {
        k = p.chunks[chunk].k_base;
        carry = p.chunks[chunk].carry;
    }

    for (int64_t base = 0; base < groups; base += kPass) {
        const int64_t wb = base + (int64_t)wave * kBlock;   // wave's first group (chunk-relative)
        if constexpr (MODE == kModeFused) {
            // signed length changes (replacement - pattern bytes) per group: at most
            // 8192 occurrences x 63 bytes, so a wave's V * 64 groups sum within int32
            int32_t v[V], ex[V];
            int32_t run = 0;
#pragma unroll
            for (int j = 0; j < V; ++j) {
                const int64_t gi = wb + (int64_t)j * kWave + lane;
                v[j] = gi < groups ? (int32_t)p.gsum[first_group + gi] : 0;
            }
#pragma unroll
            for (int j = 0; j < V; ++j) {
                int32_t inc = v[j];
#pragma unroll
                for (int d = 1; d < kWave; d <<= 1) {
                    const int32_t u = __shfl_up(inc, d, kWave);
                    if (lane >= d)
                        inc += u;
                }
                ex[j] = run + inc - v[j];
                run += __shfl(inc, kWave - 1, kWave);
            }
            if (lane == 0)
                s_wsum[wave] = run;
            __syncthreads();
            int64_t before = 0, total = 0;
            for (int w = 0; w < kSpineWaves; ++w) {
                if (w < wave)
                    before += s_wsum[w];
                total += s_wsum[w];
            }
#pragma unroll
            for (int j = 0; j < V; ++j) {
                const int64_t gi = wb + (int64_t)j * kWave + lane;
                if (gi < groups) {
                    ChunkState cs;
                    cs.k_base = k + before + ex[j];
                    cs.carry = 0;
                    cs.pad = 0;
                    p.chunks_pre[first_group + gi] = cs;
                }
            }
            k += total;
        } else if constexpr (MODE != kModeTable) {
            // kept counts: per j an inclusive wave scan, offset by the running sum
            uint32_t v[V], ex[V];
            uint32_t run = 0;
#pragma unroll
            for (int j = 0; j < V; ++j) {
                const int64_t gi = wb + (int64_t)j * kWave + lane;
                v[j] = gi < groups ? p.gsum[first_group + gi] : 0u;
            }
#pragma unroll
            for (int j = 0; j < V; ++j) {
                uint32_t inc = v[j];
#pragma unroll
                for (int d = 1; d < kWave; d <<= 1) {
                    const uint32_t u = __shfl_up(inc, d, kWave);
                    if (lane >= d)
                        inc += u;
                }
                ex[j] = run + inc - v[j];
                run += __shfl(inc, kWave - 1, kWave);
            }
            if (lane == 0)
                s_wsum[wave] = run;
            __syncthreads();
            int64_t before = 0, total = 0;
            for (int w = 0; w < kSpineWaves; ++w) {
                if (w < wave)
                    before += s_wsum[w];
                total += s_wsum[w];
            }
#pragma unroll
            for (int j = 0; j < V; ++j) {
                const int64_t gi = wb + (int64_t)j * kWave + lane;
                if (gi < groups) {
                    ChunkState cs;
                    cs.k_base = k + before + ex[j];
                    cs.carry = 0;
                    cs.pad = 0;
                    p.chunks_pre[first_group + gi] = cs;
                }
            }
            k += total;
        } else {
            // carry tables, two sweeps over j: first the wave's total table, then,
            // once the state entering the wave's block is known, each group's
            // exclusive state from an ordered wave scan; the summaries are loaded once
            Table<LT> tv[V];
#pragma unroll
            for (int j = 0; j < V; ++j) {
                const int64_t gi = wb + (int64_t)j * kWave + lane;
                tv[j] = tidentity<LT>();
                if (gi < groups) {
#pragma unroll
                    for (int c = 0; c < LT; ++c)
                        tv[j].e[c] = p.gsum[(first_group + gi) * LT + c];
                }
            }
            auto wave_scan = [&](Table64<LT> inc) {
#pragma unroll
                for (int d = 1; d < kWave; d <<= 1) {
                    Table64<LT> nb;
#pragma unroll
                    for (int c = 0; c < LT; ++c)
                        nb.e[c] = __shfl_up(inc.e[c], d, kWave);
                    if (lane >= d)
                        inc = tcompose64<LT>(nb, inc);
                }
                return inc;
            };
            Table64<LT> run = tidentity64<LT>();
#pragma unroll
            for (int j = 0; j < V; ++j)
                run = tcompose64<LT>(run, wave_broadcast0<LT>(wave_compose<LT>(widen<LT>(tv[j]), lane)));
            if (lane == 0) {
#pragma unroll
                for (int c = 0; c < LT; ++c)
                    s_wt[wave][c] = run.e[c];
            }
            __syncthreads();
            Table64<LT> wpre = tidentity64<LT>(), tot = tidentity64<LT>();
            for (int w = 0; w < kSpineWaves; ++w) {
                Table64<LT> wt;
#pragma unroll
                for (int c = 0; c < LT; ++c)
                    wt.e[c] = s_wt[w][c];
                if (w < wave)
                    wpre = tcompose64<LT>(wpre, wt);
                tot = tcompose64<LT>(tot, wt);
            }
            const uint64_t ew = tsel64<LT>(wpre, (uint32_t)carry);
            int64_t kw = k + (int64_t)(ew & kCountMask);   // state entering the j-block
            uint32_t cw = (uint32_t)(ew >> kCarryShift);
#pragma unroll
            for (int j = 0; j < V; ++j) {
                const int64_t gi = wb + (int64_t)j * kWave + lane;
                const Table64<LT> inc = wave_scan(widen<LT>(tv[j]));
                uint64_t e = (uint64_t)cw << kCarryShift;   // lane 0: nothing before it
#pragma unroll
                for (int c = 0; c < LT; ++c) {
                    const uint64_t u = __shfl_up(inc.e[c], 1, kWave);
                    if (lane > 0 && (uint32_t)c == cw)
                        e = u;
                }
                if (gi < groups) {
                    ChunkState cs;
                    cs.k_base = kw + (int64_t)(e & kCountMask);
                    cs.carry = (int)(e >> kCarryShift);
                    cs.pad = 0;
                    p.chunks_pre[first_group + gi] = cs;
                }
                const uint64_t el = tsel64<LT>(wave_broadcast_last<LT>(inc), cw);
                kw += (int64_t)(el & kCountMask);
                cw = (uint32_t)(el >> kCarryShift);
            }
            const uint64_t te = tsel64<LT>(tot, (uint32_t)carry);
            k += (int64_t)(te & kCountMask);
            carry = (int)(te >> kCarryShift);
        }
        __syncthreads();   // the wave slots are reused by the next pass
    }
    if (tid == 0 && !BLOCKS) {
        ChunkState cs;
        cs.k_base = k;
        cs.carry = carry;
        cs.pad = 0;
        p.chunks[chunk + 1] = cs;
        if (last_chunk)   // 48-bit two's complement payload (a fused pass's total may be negative)
            p.stats->k_total = ((uint64_t)p.epoch << kEpochShift) | ((uint64_t)k & ((1ull << kEpochShift) - 1));
    }
}

// Spine blocks, first launch: block b's summary over its groups -- the kept count
// (signed length change for a fused pass) or the composed carry table (entry c:
// carry out << kCarryShift | kept, for carry in c) -- into block_tabs[b * kMaxTablePat].
template <int MODE, int LT>
__global__ void __launch_bounds__(kSpineThreads) spine_reduce_kernel(ReplaceParams p, int64_t groups)
{
    constexpr int64_t kPass = spine_block_groups<MODE, LT>();
    __shared__ uint64_t s_wt[kSpineWaves][LT];
    __shared__ int64_t s_wsum[kSpineWaves];
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
    const int64_t g0 = (int64_t)blockIdx.x * kPass;
    const int64_t gn = groups - g0 < kPass ? groups - g0 : kPass;
    if constexpr (MODE != kModeTable) {
        int64_t sum = 0;
        for (int64_t i = tid; i < gn; i += kSpineThreads)
            sum += MODE == kModeFused ? (int64_t)(int32_t)p.gsum[g0 + i] : (int64_t)p.gsum[g0 + i];
#pragma unroll
        for (int d = kWave / 2; d > 0; d >>= 1)
            sum += __shfl_xor(sum, d, kWave);
        if (lane == 0)
            s_wsum[wave] = sum;
        __syncthreads();
        if (tid == 0) {
            int64_t t = 0;
            for (int w = 0; w < kSpineWaves; ++w)
                t += s_wsum[w];
            p.block_tabs[blockIdx.x * kMaxTablePat] = (uint64_t)t;
        }
    } else {
        // each wave composes a contiguous run of the block in order, then the waves'
        // tables are composed in wave order
        constexpr int64_t kPerWave = kPass / kSpineWaves;
        Table64<LT> run = tidentity64<LT>();
        const int64_t w0 = (int64_t)wave * kPerWave;
        for (int64_t b = 0; b < kPerWave; b += kWave) {
            const int64_t gi = w0 + b + lane;
            Table<LT> t = tidentity<LT>();
            if (gi < gn) {
#pragma unroll
                for (int c = 0; c < LT; ++c)
                    t.e[c] = p.gsum[(g0 + gi) * LT + c];
            }
            run = tcompose64<LT>(run, wave_broadcast0<LT>(wave_compose<LT>(widen<LT>(t), lane)));
        }
        if (lane == 0) {
#pragma unroll
            for (int c = 0; c < LT; ++c)
                s_wt[wave][c] = run.e[c];
        }
        __syncthreads();
        if (tid == 0) {
            Table64<LT> tot = tidentity64<LT>();
            for (int w = 0; w < kSpineWaves; ++w) {
                Table64<LT> wt;
#pragma unroll
                for (int c = 0; c < LT; ++c)
                    wt.e[c] = s_wt[w][c];
                tot = tcompose64<LT>(tot, wt);
            }
#pragma unroll
            for (int c = 0; c < LT; ++c)
                p.block_tabs[blockIdx.x * kMaxTablePat + c] = tot.e[c];
        }
    }
}

// Spine blocks, second launch (one wave, one walking thread): the state entering
// every block into chunks[b], the state after the last into chunks[nblocks], and
// the pass total.
template <int MODE, int LT>
__global__ void __launch_bounds__(kWave) spine_top_kernel(ReplaceParams p, int64_t nblocks)
{
    // the walk is serial (each block's entering carry selects its table entry), so
    // the wave first stages 64 blocks' summaries in LDS with independent loads and
    // lane 0 walks them there, instead of one dependent HBM round trip per block
    constexpr int W = MODE == kModeTable ? LT : 1;
    __shared__ uint64_t s_tab[kWave * W];
    const int lane = threadIdx.x;
    int64_t k = 0;
    uint32_t carry = 0;
    for (int64_t b0 = 0; b0 < nblocks; b0 += kWave) {
        const int cnt = nblocks - b0 < kWave ? (int)(nblocks - b0) : kWave;
        if (lane < cnt) {
#pragma unroll
            for (int c = 0; c < W; ++c)
                s_tab[lane * W + c] = p.block_tabs[(b0 + lane) * kMaxTablePat + c];
        }
        __syncthreads();
        if (lane == 0) {
            for (int i = 0; i < cnt; ++i) {
                ChunkState cs;
                cs.k_base = k;
                cs.carry = (int)carry;
                cs.pad = 0;
                p.chunks[b0 + i] = cs;
                if constexpr (MODE != kModeTable) {
                    k += (int64_t)s_tab[i];
                } else {
                    const uint64_t e = s_tab[i * W + carry];
                    k += (int64_t)(e & kCountMask);
                    carry = (uint32_t)(e >> kCarryShift);
                }
            }
        }
        __syncthreads();   // s_tab is refilled for the next 64 blocks
    }
    if (lane == 0) {
        ChunkState cs;
        cs.k_base = k;
        cs.carry = (int)carry;
        cs.pad = 0;
        p.chunks[nblocks] = cs;
        p.stats->k_total = ((uint64_t)p.epoch << kEpochShift) | ((uint64_t)k & ((1ull << kEpochShift) - 1));
    }
}

// Carry-table spine of a pass of up to kSpineThreads * v groups in one launch of
// 64 * ceil(groups / (64 v)) threads: lane l takes the v consecutive groups from
// l * v and composes their tables.  Each wave runs one ordered inclusive scan of its
// lanes' tables, with only as many steps as it has lanes in use; wave 0 then scans
// the waves' totals the same way and hands each wave its entering state; a lane's
// entering state is its left neighbour's inclusive table at that state, and its
// groups' states follow by walking their tables again.  A small document's pass (a
// few groups) costs a few compositions, and a pass of thousands of groups one
// launch instead of a reduce, a block scan and a spine launch.
template <int LT>
__global__ void __launch_bounds__(kSpineThreads) spine_small_kernel(ReplaceParams p, int64_t groups, int v)
{
    __shared__ uint64_t s_wt[kSpineWaves][LT];
    __shared__ int64_t s_k[kSpineWaves];
    __shared__ uint32_t s_c[kSpineWaves];
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wave = tid / kWave;
    const int nwaves = (int)(blockDim.x / kWave);
    const int64_t g0 = (int64_t)tid * v;   // this lane's first group
    Table64<LT> mine = tidentity64<LT>();
    for (int i = 0; i < v; ++i) {
        const int64_t gi = g0 + i;
        if (gi >= groups)
            break;
        Table<LT> t;
#pragma unroll
        for (int c = 0; c < LT; ++c)
            t.e[c] = p.gsum[gi * LT + c];
        mine = tcompose64<LT>(mine, widen<LT>(t));
    }
    const int64_t per_wave = (int64_t)kWave * v;
    const int64_t left = groups - (int64_t)wave * per_wave;
    const int64_t lanes_used = (left + v - 1) / v;
    const int act = lanes_used < kWave ? (int)lanes_used : kWave;   // >= 1: the launch covers `groups`
    Table64<LT> inc = mine;
    for (int d = 1; d < act; d <<= 1) {
        Table64<LT> nb;
#pragma unroll
        for (int c = 0; c < LT; ++c)
            nb.e[c] = __shfl_up(inc.e[c], d, kWave);
        if (lane >= d)
            inc = tcompose64<LT>(nb, inc);
    }
    if (lane == act - 1) {
#pragma unroll
        for (int c = 0; c < LT; ++c)
            s_wt[wave][c] = inc.e[c];
    }
    __syncthreads();
    if (wave == 0) {
        Table64<LT> w = tidentity64<LT>();
        if (lane < nwaves) {
#pragma unroll
            for (int c = 0; c < LT; ++c)
                w.e[c] = s_wt[lane][c];
        }
        for (int d = 1; d < nwaves; d <<= 1) {
            Table64<LT> nb;
#pragma unroll
            for (int c = 0; c < LT; ++c)
                nb.e[c] = __shfl_up(w.e[c], d, kWave);
            if (lane >= d)
                w = tcompose64<LT>(nb, w);
        }
        // the pass enters with no kept start and carry 0: wave w starts where the
        // inclusive scan of waves 0..w-1 leaves carry 0
        uint64_t before = 0;
#pragma unroll
        for (int c = 0; c < LT; ++c) {
            const uint64_t u = __shfl_up(w.e[c], 1, kWave);
            if (c == 0 && lane > 0)
                before = u;
        }
        if (lane < nwaves) {
            s_k[lane] = (int64_t)(before & kCountMask);
            s_c[lane] = (uint32_t)(before >> kCarryShift);
        }
        if (lane == nwaves - 1) {
            const uint64_t tot = w.e[0];
            ChunkState cs;
            cs.k_base = (int64_t)(tot & kCountMask);
            cs.carry = (int)(tot >> kCarryShift);
            cs.pad = 0;
            p.chunks[1] = cs;
            p.stats->k_total = ((uint64_t)p.epoch << kEpochShift) |
                               ((uint64_t)cs.k_base & ((1ull << kEpochShift) - 1));
        }
    }
    __syncthreads();
    const uint32_t cw = s_c[wave];
    uint64_t e = (uint64_t)cw << kCarryShift;   // lane 0: nothing of this wave before it
#pragma unroll
    for (int c = 0; c < LT; ++c) {
        const uint64_t u = __shfl_up(inc.e[c], 1, kWave);
        if (lane > 0 && (uint32_t)c == cw)
            e = u;
    }
    int64_t k = s_k[wave] + (int64_t)(e & kCountMask);
    uint32_t carry = (uint32_t)(e >> kCarryShift);
    for (int i = 0; i < v; ++i) {
        const int64_t gi = g0 + i;
        if (gi >= groups)
            break;
        ChunkState cs;
        cs.k_base = k;
        cs.carry = (int)carry;
        cs.pad = 0;
        p.chunks_pre[gi] = cs;
        Table<LT> t;
#pragma unroll
        for (int c = 0; c < LT; ++c)
            t.e[c] = p.gsum[gi * LT + c];
        const uint32_t en = tsel<LT>(t, carry);
        k += (int64_t)(en & 0xffffffu);
        carry = en >> 24;
    }
}

// ---- emit: resolve kept starts and write each tile's output ----------------
// The group's kept prefix and carry come from the spine's record for the group.
// The bitmap mode instead reads each tile's prefix and carry from the records of
// the serial greedy, and the kept starts from the bitmap.
// g: the group within p's buffer; pre: its state (kept before it, carry into it),
// unused in bitmap mode; per-tile records are written when p.info is set.
template <int MODE, int LT, int GT>
__device__ __forceinline__ void emit_group(const ReplaceParams &p, int64_t g, const ChunkState *pre)
{
    constexpr int kGroupTiles = GT;
    __shared__ __align__(16) uint8_t s_in[GroupShape<GT>::kIn];
    __shared__ __align__(16) uint8_t s_out[kStage];
    __shared__ uint8_t s_pat[kMaxSinglePassPat];
    __shared__ __align__(16) uint8_t s_rep[kRepLds + 32];
    __shared__ int s_wsum[kWaves];
    __shared__ int s_wmax[kWaves];
    __shared__ int s_halo_last;
    __shared__ uint32_t s_wtab[kWaves][LT];

    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wave = tid / kWave;
    const int L = p.L;
    const int r = p.r;
    const int64_t n = p.n;
    const int64_t nt = p.num_tiles;
    const int64_t diff = (int64_t)r - (int64_t)L;
    const int64_t t0 = g * kGroupTiles;
    if (t0 >= nt)
        return;

    if (MODE != kModeBitmap) {
        for (int i = tid; i < L; i += kThreads)
            s_pat[i] = p.pat[i];
    }
    if (p.write_output) {
        for (int i = tid; i < r && i < kRepLds; i += kThreads)
            s_rep[i] = p.rep[i];
    }
    if (tid == 0)
        s_halo_last = kNegInf;

    // ---- kept starts before the group and carry into it (from the spine) --------
    load_group<GT>(p, t0, tid, MODE == kModeBorderFree, s_in);
    int64_t k_run = 0;   // kept starts before the current tile
    int c_run = 0;       // carry into the current tile
    if (MODE != kModeBitmap) {
        const ChunkState st = *pre;
        k_run = st.k_base;
        c_run = st.carry;
    }
    // the count's candidate masks of my segments, loaded with the group
    uint32_t cm[GT];
    const bool have_masks = MODE != kModeBitmap && p.masks != nullptr;
#pragma unroll
    for (int i = 0; i < GT; ++i)
        cm[i] = (have_masks && t0 + i < nt) ? p.masks[(t0 + i) * (kTile / 32) + tid] : 0u;
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    const uint4 rep_v = lds_load16(s_rep);   // the replacement's first 16 bytes
    // carry tables without the count's masks (batch rounds) rebuild them from SADs
    QsadWords qw{};
    if (MODE == kModeTable && !have_masks)
        qw = qsad_words_lds(s_pat, L);

    const int seg = tid * kBytesPerThread;
    for (int i = 0; i < kGroupTiles; ++i) {
        const int64_t ti = t0 + i;
        if (ti >= nt)
            break;
        const int64_t start = ti * kTile;
        const int valid = (int)((n - start) < kTile ? (n - start) : kTile);
        const bool last_tile = ti == nt - 1;
        const int seg_valid = valid - seg;
        const int jend = seg_valid < 32 ? (seg_valid > 0 ? seg_valid : 0) : 32;
        const int in_off = kHalo + i * kTile;   // LDS offset of the tile's first byte

        uint32_t m = 0;
        if constexpr (MODE == kModeBitmap) {
            m = p.bitmap[ti * (kTile / 32) + tid];
            k_run = p.info[ti].k_excl;
            c_run = p.info[ti].carry_in;
        } else if (have_masks) {
#pragma unroll
            for (int k = 0; k < GT; ++k)
                m = k == i ? cm[k] : m;
        } else if constexpr (MODE == kModeTable) {
            m = cand_mask_qsad(p, qw, s_in + in_off + seg, start + seg);
        } else {
            m = cand_mask(p, s_in + in_off, s_pat, seg, start + seg);
        }
        if constexpr (MODE == kModeBorderFree) {
            if (i == 0) {
                // carry from a candidate in the L-1 bytes left of the group; it is kept
                // (border-free occurrences never overlap) unless below the floor
                if (ti > 0 && tid < L - 1) {
                    const int q = -1 - tid;
                    const int64_t gp = start + q;
                    if (gp >= p.floor && gp + L <= n) {
                        const uint8_t *sp = s_in + in_off + q;
                        bool ok = true;
                        for (int k = 0; k < L; ++k) {
                            if (sp[k] != s_pat[k]) {
                                ok = false;
                                break;
                            }
                        }
                        if (ok)
                            atomicMax(&s_halo_last, q);
                    }
                }
                __syncthreads();
                c_run = s_halo_last + L > 0 ? s_halo_last + L : 0;
            }
        }
        const int c_t = c_run;

        int c_i = 0;          // leading bytes of my segment covered by an earlier kept match
        int k_loc = 0;        // kept starts in the tile before my segment
        int tile_k = 0;       // kept starts in the tile
        int tile_last = kNegInf;
        int c_next = 0;       // carry into the next tile
        uint32_t kept = m;

        if constexpr (MODE == kModeBorderFree || MODE == kModeBitmap) {
            const int kc = __builtin_popcount(m);
            const int lastq = m ? (seg + 31 - __builtin_clz(m)) : kNegInf;
            const int ks = (int)wave_incl_add_u32((uint32_t)kc);
            const int lq = wave_incl_max_i32(lastq);
            const int lq_ex = wave_shr1_i32(lq, kNegInf);
            if (lane == kWave - 1) {
                s_wsum[wave] = ks;
                s_wmax[wave] = lq;
            }
            __syncthreads();
            int wsum = 0, wmax = c_t - L;   // a virtual kept start covering [.., c_t)
            for (int w = 0; w < kWaves; ++w) {
                if (w < wave) {
                    wsum += s_wsum[w];
                    wmax = s_wmax[w] > wmax ? s_wmax[w] : wmax;
                }
                tile_k += s_wsum[w];
                tile_last = s_wmax[w] > tile_last ? s_wmax[w] : tile_last;
            }
            k_loc = wsum + ks - kc;
            const int prev_last = lq_ex > wmax ? lq_ex : wmax;
            c_i = prev_last + L - seg;
            if (c_i < 0)
                c_i = 0;
            if constexpr (MODE == kModeBorderFree) {
                c_next = tile_last + L - kTile;
                if (c_next < 0 || last_tile)
                    c_next = 0;
            } else {
                c_next = last_tile ? 0 : p.info[ti + 1].carry_in;
            }
        } else {
            // carry tables: wave-inclusive scan, then across the four waves
            const Table<LT> t = segment_table<LT>(m, L);
            const Table<LT> inc = wave_incl_table<LT>(t);
            const Table<LT> ex = wave_shr1_table<LT>(inc);
            if (lane == kWave - 1) {
#pragma unroll
                for (int c = 0; c < LT; ++c)
                    s_wtab[wave][c] = inc.e[c];
            }
            __syncthreads();
            // the tile's entry state c_t through the wave totals: the state and kept
            // count entering my wave, and leaving the tile (each step one LDS read of
            // the entry at the current state, not a composition of whole tables)
            uint32_t st = (uint32_t)c_t, kw = 0, s_mine = 0, k_mine = 0;
#pragma unroll
            for (int w = 0; w < kWaves; ++w) {
                if (w == wave) {
                    s_mine = st;
                    k_mine = kw;
                }
                const uint32_t e = s_wtab[w][st];
                kw += e & 0xffffffu;
                st = e >> 24;
            }
            const uint32_t me = tsel<LT>(ex, s_mine);
            c_i = (int)(me >> 24);
            k_loc = (int)(k_mine + (me & 0xffffffu));
            c_next = last_tile ? 0 : (int)st;
            int lst;
            kept = LT == 2 ? greedy32_pair(m, c_i, &lst) : greedy32(m, c_i, L, &lst);
            const int lastq = kept ? seg + lst : kNegInf;
            // kept starts in the tile from the wave totals; the last one only for the
            // tile's record
            tile_k = (int)kw;
            if (p.info) {
                const int lq = __builtin_amdgcn_readlane(wave_incl_max_i32(lastq), kWave - 1);
                if (lane == 0)
                    s_wmax[wave] = lq;
                __syncthreads();
                for (int w = 0; w < kWaves; ++w)
                    tile_last = s_wmax[w] > tile_last ? s_wmax[w] : tile_last;
            }
        }

        // ---- per-tile record (prefix queries) ------------------------------------
        if (tid == 0 && MODE != kModeBitmap && p.info) {
            TileInfo ti_rec;
            ti_rec.k_excl = k_run;
            ti_rec.last_abs = tile_k > 0 ? start + tile_last : -1;
            ti_rec.carry_in = c_t;
            ti_rec.mode = MODE;
            p.info[ti] = ti_rec;
        }

        // ---- output ------------------------------------------------------------------
        // Each thread places its source runs and replacements into the tile's output
        // image in LDS (unaligned 16-byte LDS moves, exact tails), then the workgroup
        // writes the image out with contiguous 16-byte stores.  A tile whose output
        // does not fit the stage (or a replacement longer than kRepLds) is written
        // straight to HBM by each thread instead.
        if (p.write_output) {
            const int64_t out_len = (int64_t)valid + c_next - c_t + (int64_t)tile_k * diff;
            uint8_t *const out_first = p.dst + start + c_t + k_run * diff;   // tile's first output byte
            if (r <= kRepLds && out_len + 32 <= kStage) {
                int o = (seg + c_i - c_t) + k_loc * (int)diff;
                int j = c_i;
                uint32_t M = j >= 32 ? 0u : (kept & (0xffffffffu << j));
                while (j < jend) {
                    const int q = M ? __builtin_ctz(M) : jend;
                    if (q > j) {
                        lds_move(s_out + o, s_in + in_off + seg + j, q - j);
                        o += q - j;
                    }
                    if (q >= jend)
                        break;
                    if (r == 1 && !(q + L < 32 && ((M >> (q + L)) & 1u))) {
                        // a 1-byte replacement with no kept start right after it
                        s_out[o] = (uint8_t)rep_v.x;
                        o += 1;
                        j = q + L;
                    } else if (r == 1) {
                        // a 1-byte replacement: this and the kept starts right after it
                        // (every L bytes) become one run of replacement bytes
                        int k = 1;
                        while (q + k * L < 32 && ((M >> (q + k * L)) & 1u))
                            ++k;
                        const uint32_t b4 = (rep_v.x & 0xffu) * 0x01010101u;
                        const uint4 fill = {b4, b4, b4, b4};
                        int w = 0;
                        for (; w + 16 <= k; w += 16)
                            lds_store16(s_out + o + w, fill);
                        if (w < k)
                            store_tail(s_out + o + w, fill, k - w);
                        o += k;
                        j = q + k * L;
                    } else {
                        if (r < 16)
                            store_tail(s_out + o, rep_v, r);
                        else if (r == 16)
                            lds_store16(s_out + o, rep_v);
                        else
                            lds_move(s_out + o, s_rep, r);
                        o += r;
                        j = q + L;
                    }
                    M = j >= 32 ? 0u : (M & (0xffffffffu << j));
                }
                __syncthreads();
                const int len = (int)out_len;
                for (int c = tid * 16; c < len; c += kThreads * 16) {
                    const uint4 v = *reinterpret_cast<const uint4 *>(s_out + c);
                    if (c + 16 <= len)
                        __builtin_memcpy(out_first + c, &v, 16);
                    else
                        store_tail(out_first + c, v, len - c);
                }
            } else {
                uint8_t *out = out_first + (int64_t)(seg + c_i - c_t) + (int64_t)k_loc * diff;
                int j = c_i;
                uint32_t M = j >= 32 ? 0u : (kept & (0xffffffffu << j));
                while (j < jend) {
                    const int q = M ? __builtin_ctz(M) : jend;
                    if (q > j) {
                        store_from_lds(out, s_in, in_off + seg + j, q - j);
                        out += q - j;
                    }
                    if (q >= jend)
                        break;
                    if (r <= kRepLds)
                        store_from_lds(out, s_rep, 0, r);
                    else
                        store_from_global(out, p.rep, r);
                    out += r;
                    j = q + L;
                    M = j >= 32 ? 0u : (M & (0xffffffffu << j));
                }
            }
        }
        k_run += tile_k;
        c_run = c_next;
        __syncthreads();   // s_out, s_wsum and s_wtab are reused by the next tile
    }
}

// Waves per SIMD the emit is compiled for (1: the compiler's choice).  The 8- and
// 16-state emits are held to 96 VGPRs so that five fit instead of four.
template <int MODE, int LT, int GT>
constexpr int emit_waves_per_eu()
{
    return MODE == kModeTable && LT >= 8 && GT <= 2 ? 5 : 1;
}

template <int MODE, int LT, int GT>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(emit_waves_per_eu<MODE, LT, GT>())))
emit_kernel(ReplaceParams p, int64_t first_group)
{
    const int64_t g = first_group + blockIdx.x;
    emit_group<MODE, LT, GT>(p, g, p.chunks_pre + g);
}

// ---- lean border-free pass: 64 source bytes per thread ------------------------
// The pass of a border-free pattern (no proper prefix equals a suffix: occurrences
// never overlap, so every occurrence at or above the seed floor is kept).  A
// workgroup takes a 16 KiB group; thread t owns source bytes [64t, 64t + 64) of it.
//   bf_count  stages the group, builds each thread's 64-position candidate mask and
//             writes the group's kept count (one word);
//   spine     the same spine as the other modes (group states);
//   bf_emit   stages the group again, rebuilds the masks, scans kept counts and last
//             kept starts across the workgroup (wave64 DPP scans), and writes every
//             thread's pieces (source runs and replacements) into an LDS image of
//             the group's output, which the workgroup then stores with 16-byte
//             stores.
// A piece copy moves 16-byte chunks with its last chunk ending exactly at the run
// end (the chunks overlap; same bytes), so only runs shorter than 16 bytes take the
// 8/4/2/1-byte forms.
constexpr int kBfSeg = 64;                            // source bytes per thread
constexpr int kBfGroup = kThreads * kBfSeg;           // 16384 source bytes per workgroup
static_assert(kBfGroup == kGroupBytes, "the spine's group size");
constexpr int kBfOutCap = kBfGroup + 4096;            // LDS image of a group's output

// wave64 inclusive scans with DPP (row_shr 1/2/4/8 inside rows of 16 lanes, then
// row_bcast 15 and 31 across rows).  Lanes without a source read the identity.


// previous lane's value (wave_shr 1); lane 0 reads `first`

__device__ __forceinline__ uint64_t lds_load8(const uint8_t *p)
{
    uint64_t v;
    __builtin_memcpy(&v, p, 8);
    return v;
}
__device__ __forceinline__ void lds_store8(uint8_t *p, uint64_t v) { __builtin_memcpy(p, &v, 8); }
__device__ __forceinline__ uint32_t lds_load4(const uint8_t *p)
{
    uint32_t v;
    __builtin_memcpy(&v, p, 4);
    return v;
}
__device__ __forceinline__ void lds_store4(uint8_t *p, uint32_t v) { __builtin_memcpy(p, &v, 4); }
__device__ __forceinline__ uint16_t lds_load2(const uint8_t *p)
{
    uint16_t v;
    __builtin_memcpy(&v, p, 2);
    return v;
}
__device__ __forceinline__ void lds_store2(uint8_t *p, uint16_t v) { __builtin_memcpy(p, &v, 2); }

// Copies len (0..64) bytes between LDS buffers at any alignment, exactly: 16-byte
// chunks with the last one ending at len, or two overlapping 8/4/2-byte moves.
__device__ __forceinline__ void bf_copy_run(uint8_t *dst, const uint8_t *src, int len)
{
    if (len >= 16) {
        for (int k = 0; k < len - 16; k += 16)
            lds_store16(dst + k, lds_load16(src + k));
        lds_store16(dst + len - 16, lds_load16(src + len - 16));
    } else if (len >= 8) {
        const uint64_t a = lds_load8(src), b = lds_load8(src + len - 8);
        lds_store8(dst, a);
        lds_store8(dst + len - 8, b);
    } else if (len >= 4) {
        const uint32_t a = lds_load4(src), b = lds_load4(src + len - 4);
        lds_store4(dst, a);
        lds_store4(dst + len - 4, b);
    } else if (len >= 2) {
        const uint16_t a = lds_load2(src), b = lds_load2(src + len - 2);
        lds_store2(dst, a);
        lds_store2(dst + len - 2, b);
    } else if (len == 1) {
        *dst = *src;
    }
}

// The replacement split once into the 8/4/2/1-byte stores of its length (r < 16);
// longer replacements are copied from LDS (r <= kRepLds) or HBM.
struct BfRep {
    uint64_t w8;
    uint32_t w4;
    uint32_t w2;
    uint32_t w1;
    int o4, o2, o1;
};

__device__ __forceinline__ BfRep bf_rep_parts(uint4 v, int r)
{
    BfRep b;
    uint64_t lo = (uint64_t)v.x | ((uint64_t)v.y << 32);
    const uint64_t hi = (uint64_t)v.z | ((uint64_t)v.w << 32);
    b.w8 = lo;
    int off = 0;
    if (r & 8) {
        lo = hi;
        off = 8;
    }
    b.o4 = off;
    b.w4 = (uint32_t)lo;
    if (r & 4) {
        lo >>= 32;
        off += 4;
    }
    b.o2 = off;
    b.w2 = (uint32_t)lo & 0xffffu;
    if (r & 2) {
        lo >>= 16;
        off += 2;
    }
    b.o1 = off;
    b.w1 = (uint32_t)lo & 0xffu;
    return b;
}

__device__ __forceinline__ void bf_put_rep_lds(uint8_t *dst, const BfRep &b, uint4 v, const uint8_t *s_rep, int r)
{
    if (r < 16) {
        if (r & 8)
            lds_store8(dst, b.w8);
        if (r & 4)
            lds_store4(dst + b.o4, b.w4);
        if (r & 2)
            lds_store2(dst + b.o2, (uint16_t)b.w2);
        if (r & 1)
            dst[b.o1] = (uint8_t)b.w1;
    } else if (r == 16) {
        lds_store16(dst, v);
    } else {
        bf_copy_run(dst, s_rep, r <= 64 ? r : 64);
        for (int k = 64; k < r; k += 64)
            bf_copy_run(dst + k, s_rep + k, r - k < 64 ? r - k : 64);
    }
}

// Copies len (0..64) bytes from LDS to HBM (any alignment), exactly: the 16-byte
// chunks are read first (the last one ending at len), then stored; a run shorter than
// 16 bytes is stored from one read as 8/4/2/1-byte pieces.
__device__ __forceinline__ void bf_store_run(uint8_t *dst, const uint8_t *src, int len)
{
    if (len >= 16) {
        const int e = len - 16;
        const uint4 a = lds_load16(src), b = lds_load16(src + (16 < e ? 16 : e)),
                    c = lds_load16(src + (32 < e ? 32 : e)), d = lds_load16(src + e);
        __builtin_memcpy(dst, &a, 16);
        __builtin_memcpy(dst + (16 < e ? 16 : e), &b, 16);
        __builtin_memcpy(dst + (32 < e ? 32 : e), &c, 16);
        __builtin_memcpy(dst + e, &d, 16);
    } else if (len > 0) {
        store_tail(dst, lds_load16(src), len);
    }
}

__device__ __forceinline__ void bf_put_rep_global(uint8_t *dst, const BfRep &b, uint4 v, const uint8_t *s_rep,
                                                  const uint8_t *g_rep, int r)
{
    if (r < 16) {
        if (r & 8)
            __builtin_memcpy(dst, &b.w8, 8);
        if (r & 4)
            __builtin_memcpy(dst + b.o4, &b.w4, 4);
        if (r & 2) {
            const uint16_t h = (uint16_t)b.w2;
            __builtin_memcpy(dst + b.o2, &h, 2);
        }
        if (r & 1)
            dst[b.o1] = (uint8_t)b.w1;
    } else if (r == 16) {
        __builtin_memcpy(dst, &v, 16);
    } else if (r <= kRepLds) {
        for (int k = 0; k < r; k += 64)
            bf_store_run(dst + k, s_rep + k, r - k < 64 ? r - k : 64);
    } else {
        store_from_global(dst, g_rep, r);
    }
}

// Candidate mask of the 64 positions whose bytes start at s_seg (16-byte aligned
// LDS): bit j set iff the pattern occurs at gpos + j and floor <= gpos + j <= n - L.
// FULL4: L >= 4, so the first four pattern bytes are compared unmasked.  Reads
// s_seg[0, 68) and, for L > 4, the candidates' bytes up to L - 1 past them.
template <bool FULL4>
__device__ __forceinline__ uint64_t bf_cand64(const ReplaceParams &p, const uint8_t *s_seg, const uint8_t *s_pat,
                                              int64_t gpos)
{
    uint32_t d[17];
    const uint4 *q = reinterpret_cast<const uint4 *>(s_seg);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint4 v = q[k];
        d[4 * k] = v.x;
        d[4 * k + 1] = v.y;
        d[4 * k + 2] = v.z;
        d[4 * k + 3] = v.w;
    }
    d[16] = *reinterpret_cast<const uint32_t *>(s_seg + 64);
    const uint32_t pat4 = p.pat4, mask4 = p.mask4;
    uint32_t lo = 0, hi = 0;
#pragma unroll
    for (int j = 31; j >= 0; --j) {
        const uint32_t w = (j % 4 == 0) ? d[j / 4] : __builtin_amdgcn_alignbyte(d[j / 4 + 1], d[j / 4], j % 4);
        lo = lo + lo + (uint32_t)(FULL4 ? w == pat4 : ((w ^ pat4) & mask4) == 0u);
    }
#pragma unroll
    for (int j = 63; j >= 32; --j) {
        const uint32_t w = (j % 4 == 0) ? d[j / 4] : __builtin_amdgcn_alignbyte(d[j / 4 + 1], d[j / 4], j % 4);
        hi = hi + hi + (uint32_t)(FULL4 ? w == pat4 : ((w ^ pat4) & mask4) == 0u);
    }
    uint64_t m = ((uint64_t)hi << 32) | lo;
    const int64_t lo_lim = p.floor - gpos;
    const int64_t hi_lim = p.n - (int64_t)p.L - gpos;
    if (lo_lim > 0 || hi_lim < 63) {
        if (hi_lim < 0 || lo_lim > 63 || lo_lim > hi_lim) {
            m = 0;
        } else {
            const int a = lo_lim < 0 ? 0 : (int)lo_lim;
            const int b = hi_lim > 63 ? 63 : (int)hi_lim;
            m &= (b == 63 ? ~0ull : ((2ull << b) - 1ull)) & (~0ull << a);
        }
    }
    if (p.L > 4 && m) {
        uint64_t mm = m;
        m = 0;
        while (mm) {
            const int j = (int)__builtin_ctzll(mm);
            mm &= mm - 1;
            bool ok = true;
            for (int k = 4; k < p.L; ++k) {
                if (s_seg[j + k] != s_pat[k]) {
                    ok = false;
                    break;
                }
            }
            m |= (uint64_t)ok << j;
        }
    }
    return m;
}

// Number of occurrences among the 64 positions whose bytes start at s_seg (16-byte
// aligned LDS), for a pattern of 4(NQ-1) < L <= 4 NQ bytes, all positions in range.
// v_mqsad_pk_u16_u8 gives the masked byte SADs of the four windows starting at
// bytes 0..3 of a 64-bit operand against a reference word; chained over the NQ
// pattern words (the data xor qx, see pattern_qsad) a window's sum is 0 exactly when
// the pattern occurs there, and (0x8000 - sum) has bit 15 set only then (sums stay
// below 16 * 255).  Reads s_seg[0, 64 + 4 NQ).
template <int NQ>
__device__ __forceinline__ uint32_t bf_count64_qsad(const ReplaceParams &p, const uint8_t *s_seg)
{
    uint32_t d[20];
    const u32x4 *q4 = reinterpret_cast<const u32x4 *>(s_seg);
    const uint32_t x = p.qx;
#pragma unroll
    for (int k = 0; k < (16 + NQ + 3) / 4; ++k) {
        const u32x4 v = q4[k];
        d[4 * k] = v.x ^ x;
        d[4 * k + 1] = v.y ^ x;
        d[4 * k + 2] = v.z ^ x;
        d[4 * k + 3] = v.w ^ x;
    }
    uint32_t cnt = 0;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        uint64_t acc = 0;
#pragma unroll
        for (int q = 0; q < NQ; ++q)
            acc = __builtin_amdgcn_mqsad_pk_u16_u8(((uint64_t)d[k + q + 1] << 32) | d[k + q], p.qref[q], acc);
        const uint32_t lo = (uint32_t)acc, hi = (uint32_t)(acc >> 32);
        cnt += __builtin_popcount((0x80008000u - lo) & 0x80008000u) +
               __builtin_popcount((0x80008000u - hi) & 0x80008000u);
    }
    return cnt;
}

template <int NQ>
__device__ __forceinline__ uint32_t bf_count_seg(const ReplaceParams &p, const uint8_t *s_seg, const uint8_t *s_pat,
                                                 int64_t gpos)
{
    // a segment wholly inside [floor, n - L] takes the SAD count; the boundary
    // segments the per-position masks
    if (gpos >= p.floor && gpos + 63 <= p.n - p.L)
        return bf_count64_qsad<NQ>(p, s_seg);
    return (uint32_t)__builtin_popcountll(bf_cand64<(NQ > 1)>(p, s_seg, s_pat, gpos));
}

// The candidate mask of the 64 positions whose bytes start at s_seg (all in range),
// from the same masked quad SADs as bf_count64_qsad: bit i set where the window
// at position i sums to 0.  (0x80008000 - w) has bit 15 set exactly where the low
// 16-bit sum of w is 0 and bit 31 where the high one is.
template <int NQ>
__device__ __forceinline__ uint64_t bf_mask64_qsad(const ReplaceParams &p, const uint8_t *s_seg)
{
    uint32_t d[20];
    const u32x4 *q4 = reinterpret_cast<const u32x4 *>(s_seg);
    const uint32_t x = p.qx;
#pragma unroll
    for (int k = 0; k < (16 + NQ + 3) / 4; ++k) {
        const u32x4 v = q4[k];
        d[4 * k] = v.x ^ x;
        d[4 * k + 1] = v.y ^ x;
        d[4 * k + 2] = v.z ^ x;
        d[4 * k + 3] = v.w ^ x;
    }
    uint32_t half[2] = {0u, 0u};
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        uint64_t acc = 0;
#pragma unroll
        for (int q = 0; q < NQ; ++q)
            acc = __builtin_amdgcn_mqsad_pk_u16_u8(((uint64_t)d[k + q + 1] << 32) | d[k + q], p.qref[q], acc);
        const uint32_t tl = (0x80008000u - (uint32_t)acc) & 0x80008000u;
        const uint32_t th = (0x80008000u - (uint32_t)(acc >> 32)) & 0x80008000u;
        const uint32_t nib = ((tl >> 15) | (tl >> 30) | (th >> 13) | (th >> 28)) & 0xfu;
        half[k / 8] |= nib << (4 * (k % 8));
    }
    return ((uint64_t)half[1] << 32) | half[0];
}

template <int NQ>
__device__ __forceinline__ uint64_t bf_mask_seg(const ReplaceParams &p, const uint8_t *s_seg, const uint8_t *s_pat,
                                                int64_t gpos)
{
    if (gpos >= p.floor && gpos + 63 <= p.n - p.L)
        return bf_mask64_qsad<NQ>(p, s_seg);
    return bf_cand64<(NQ > 1)>(p, s_seg, s_pat, gpos);
}

template <bool FULL4>
__global__ void __launch_bounds__(kThreads)
bf_count_kernel(ReplaceParams p, int64_t first_group)
{
    __shared__ __align__(16) uint8_t s_in[GroupShape<kGroupTiles>::kIn];
    __shared__ uint8_t s_pat[kMaxSinglePassPat];
    __shared__ uint32_t s_red[kWaves];
    const int64_t g = first_group + blockIdx.x;
    const int tid = threadIdx.x;
    const int64_t t0 = g * kGroupTiles;
    if (t0 >= p.num_tiles)
        return;
    if (p.L > 4)
        for (int i = tid; i < p.L; i += kThreads)
            s_pat[i] = p.pat[i];
    load_group<kGroupTiles>(p, t0, tid, false, s_in);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    const int64_t gpos = t0 * kTile + (int64_t)tid * kBfSeg;
    uint32_t cnt = 0;
    if (gpos < p.n) {
        const uint8_t *seg = s_in + kHalo + tid * kBfSeg;
        if (p.L <= 4)
            cnt = bf_count_seg<1>(p, seg, s_pat, gpos);
        else if (p.L <= 8)
            cnt = bf_count_seg<2>(p, seg, s_pat, gpos);
        else if (p.L <= 12)
            cnt = bf_count_seg<3>(p, seg, s_pat, gpos);
        else if (p.L <= 16)
            cnt = bf_count_seg<4>(p, seg, s_pat, gpos);
        else
            cnt = (uint32_t)__builtin_popcountll(bf_cand64<FULL4>(p, seg, s_pat, gpos));
    }
    cnt = wave_incl_add_u32(cnt);
    if ((tid & (kWave - 1)) == kWave - 1)
        s_red[tid / kWave] = cnt;
    __syncthreads();
    if (tid == 0)
        p.gsum[g] = s_red[0] + s_red[1] + s_red[2] + s_red[3];
}

// One group's source, prefetched into registers: lane l of wave w holds the 16-byte
// granules w*4096 + 1024k + 16l (k < 4) of the group (coalesced loads), wave 0 the
// 256 bytes left of the group and wave 1 the 256 bytes right of it (4 per lane).
// (clang vector types and plain scalars: with HIP's uint4 and the ChunkState struct
// the stage was kept in scratch memory across the loop iteration)
struct BfStage {
    u32x4 v0, v1, v2, v3;
    uint32_t halo;
    int64_t k_base;         // the group's ChunkState.k_base (WITH_STATE)
};

template <bool WITH_STATE>
__device__ __forceinline__ void bf_prefetch(const ReplaceParams &p, int64_t g, int tid, BfStage &b)
{
    // every load is unconditional (a granule past n reads the group's first granule
    // instead), so no branch makes the compiler wait for the loads where they join
    const int64_t G0 = g * (int64_t)kBfGroup;
    const int64_t avail = p.n - G0;
    const int wave = tid / kWave, lane = tid & (kWave - 1);
    const uint8_t *const base = p.src + G0;
    const int off = wave * 4096 + lane * 16;
    b.v0 = *reinterpret_cast<const u32x4 *>(base + (off < avail ? off : 0));
    b.v1 = *reinterpret_cast<const u32x4 *>(base + (off + 1024 < avail ? off + 1024 : 0));
    b.v2 = *reinterpret_cast<const u32x4 *>(base + (off + 2048 < avail ? off + 2048 : 0));
    b.v3 = *reinterpret_cast<const u32x4 *>(base + (off + 3072 < avail ? off + 3072 : 0));
    int64_t hoff = lane * 4;                       // wave 0: left halo, wave 1: right halo
    if (wave == 0)
        hoff = G0 >= kHalo ? hoff - kHalo : hoff;
    else if (kBfGroup + lane * 4 < avail)
        hoff = kBfGroup + lane * 4;
    b.halo = *reinterpret_cast<const uint32_t *>(base + hoff);
    if constexpr (WITH_STATE)
        b.k_base = p.chunks_pre[g].k_base;
}

__device__ __forceinline__ void bf_stage_to_lds(const BfStage &b, int tid, uint8_t *s_in)
{
    const int wave = tid / kWave, lane = tid & (kWave - 1);
    uint8_t *const d = s_in + kHalo + wave * 4096 + lane * 16;
    *reinterpret_cast<u32x4 *>(d) = b.v0;
    *reinterpret_cast<u32x4 *>(d + 1024) = b.v1;
    *reinterpret_cast<u32x4 *>(d + 2048) = b.v2;
    *reinterpret_cast<u32x4 *>(d + 3072) = b.v3;
    if (wave == 0)
        *reinterpret_cast<uint32_t *>(s_in + lane * 4) = b.halo;
    if (wave == 1)
        *reinterpret_cast<uint32_t *>(s_in + kHalo + kBfGroup + lane * 4) = b.halo;
}

// LDS of a border-free emit workgroup.
struct BfLds {
    uint8_t *in;         // the group with halos (GroupShape<kGroupTiles>::kIn)
    uint8_t *out;        // the group's output image (kBfOutCap)
    const uint8_t *pat;
    const uint8_t *rep;
    uint32_t *wsum;      // kWaves
    int *wmax;           // kWaves
    int *halo_last;      // set to INT_MIN before the barrier that publishes the staged group
    int *tile;           // 4
};

// Carry into the group at G0: an occurrence starting in the L-1 bytes left of it
// (kept unless below the floor, since occurrences never overlap), its start relative
// to G0 raised into *halo_last by the threads below L-1.
__device__ __forceinline__ void bf_halo_carry(const ReplaceParams &p, const uint8_t *s_in, const uint8_t *s_pat,
                                              int64_t G0, int *halo_last)
{
    const int tid = threadIdx.x;
    if (G0 > 0 && tid < p.L - 1) {
        const int qq = -1 - tid;
        const int64_t gp = G0 + qq;
        if (gp >= p.floor && gp + p.L <= p.n) {
            const uint8_t *sp = s_in + kHalo + qq;
            bool ok = true;
            for (int k = 0; k < p.L; ++k) {
                if (sp[k] != s_pat[k]) {
                    ok = false;
                    break;
                }
            }
            if (ok)
                atomicMax(halo_last, qq);
        }
    }
}

// What a group's count phase can hand its emit (the held-group launch): each thread's
// candidate mask and the wave-inclusive scans of kept counts and last kept starts, and
// in LDS the per-wave totals and the carry-in occurrence of the group's left halo.
struct BfPre {
    uint64_t m;
    uint32_t kinc;
    int linc;
    const uint32_t *wsum;
    const int *wmax;
    const int *halo_last;
};

// The kept starts before a group, for an emit that has them at its start (KnownK)
// or waits for them only where the output address is first needed (a persistent
// launch's prefix word, polled by thread 0; see LateK beside the sweep kernels).
// poll() is called by every thread before a barrier, get() after it.
struct KnownK {
    int64_t k;
    __device__ __forceinline__ void poll() {}
    __device__ __forceinline__ bool get(int64_t &out) const
    {
        out = k;
        return true;
    }
};

// Emit of the staged group g (kept starts before it from kp): per-tile records and
// the output.  The output image in LDS does not depend on the kept starts before
// the group, only its address does, so a late prefix is waited for after the image
// is built.  Ends after the output stores are issued (the caller synchronizes
// before the LDS is reused); false if the prefix wait gave up.
template <bool FULL4, bool IMAGE, int CAP = kBfOutCap, class KP>
__device__ __forceinline__ bool bf_emit_staged(const ReplaceParams &p, int64_t g, KP &kp, const BfLds &s,
                                               uint4 rep_v, const BfRep &rp, const BfPre *pre = nullptr)
{
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wave = tid / kWave;
    const int L = p.L;
    const int r = p.r;
    const int64_t n = p.n;
    const int64_t diff = (int64_t)r - L;
    const int64_t t0 = g * kGroupTiles;
    const int64_t G0 = t0 * kTile;
    const int valid = (int)((n - G0) < kBfGroup ? (n - G0) : kBfGroup);
    const bool last_group = G0 + kBfGroup >= n;

    const int seg = tid * kBfSeg;
    const int64_t gpos = G0 + seg;
    uint64_t m;
    uint32_t kinc;
    int linc;
    const uint32_t *wsum = s.wsum;
    const int *wmax = s.wmax;
    const int *halo_last = s.halo_last;
    if (pre) {
        m = pre->m;
        kinc = pre->kinc;
        linc = pre->linc;
        wsum = pre->wsum;
        wmax = pre->wmax;
        halo_last = pre->halo_last;
    } else {
        bf_halo_carry(p, s.in, s.pat, G0, s.halo_last);
        m = gpos < n ? bf_cand64<FULL4>(p, s.in + kHalo + seg, s.pat, gpos) : 0ull;
        kinc = wave_incl_add_u32((uint32_t)__builtin_popcountll(m));
        linc = wave_incl_max_i32(m ? seg + 63 - (int)__builtin_clzll(m) : INT_MIN);
        if (lane == kWave - 1) {
            s.wsum[wave] = kinc;
            s.wmax[wave] = linc;
        }
        __syncthreads();
    }
    const int k_t = __builtin_popcountll(m);
    const int lq = m ? seg + 63 - (int)__builtin_clzll(m) : INT_MIN;   // last kept start (group-relative)
    const int c0 = *halo_last == INT_MIN ? 0 : *halo_last + L;   // leading group bytes covered
    uint32_t kbefore = kinc - (uint32_t)k_t;
    int blk_prev = wave_shr1_i32(linc, INT_MIN);   // last kept start before my segment
    uint32_t ktot = 0;
    int gl = INT_MIN;   // last kept start of the group
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
        const uint32_t ws = wsum[w];
        const int wm = wmax[w];
        if (w < wave) {
            kbefore += ws;
            blk_prev = max(blk_prev, wm);
        }
        ktot += ws;
        gl = max(gl, wm);
    }
    const int prev_last = max(blk_prev, c0 - L);   // with the virtual kept start covering [0, c0)
    int c_t = prev_last + L - seg;                 // leading bytes of my segment covered
    c_t = c_t < 0 ? 0 : c_t;
    const int c_next = (last_group || gl == INT_MIN) ? 0 : max(0, gl + L - kBfGroup);

    // per-tile records: tile 0 is threads 0-127, tile 1 threads 128-255
    if (tid == 128) {
        s.tile[0] = (int)kbefore;
        s.tile[1] = c_t;
    }
    if (tid == 127 || tid == 255)
        s.tile[2 + (tid >> 7)] = max(blk_prev, lq);   // last kept start up to the tile's end
    // output of the group: [out_base, out_base + out_len) in the destination, out_base =
    // G0 + c0 + k_excl * diff
    const int64_t out_len = (int64_t)valid - c0 + c_next + (int64_t)ktot * diff;
    const int jend = (valid - seg) < kBfSeg ? ((valid - seg) > 0 ? valid - seg : 0) : kBfSeg;
    int j = c_t < jend ? c_t : jend;
    const int64_t o = (int64_t)(seg + j - c0) + (int64_t)kbefore * diff;   // my first output byte
    uint64_t M = m;
    const bool image = IMAGE && p.write_output && out_len + 16 <= CAP && r <= kRepLds;

    auto records = [&](int64_t k_excl) {
        if (tid == 0 && p.info) {
            TileInfo a;
            a.k_excl = k_excl;
            a.last_abs = s.tile[2] >= 0 ? G0 + s.tile[2] : -1;
            a.carry_in = c0;
            a.mode = kModeBorderFree;
            p.info[t0] = a;
            if (t0 + 1 < p.num_tiles) {
                TileInfo b;
                b.k_excl = k_excl + s.tile[0];
                b.last_abs = s.tile[3] >= kTile ? G0 + s.tile[3] : -1;
                b.carry_in = s.tile[1] < kTile ? s.tile[1] : kTile;
                b.mode = kModeBorderFree;
                p.info[t0 + 1] = b;
            }
        }
    };
    const uint8_t *const s_src = s.in + kHalo + seg;
    int64_t k_excl = 0;

    if (image) {
        // my pieces into the LDS image, then the workgroup stores the image
        int oo = (int)o;
        for (;;) {
            const int qk = M ? (int)__builtin_ctzll(M) : jend;
            bf_copy_run(s.out + oo, s_src + j, qk - j);
            oo += qk - j;
            if (qk >= jend)
                break;
            bf_put_rep_lds(s.out + oo, rp, rep_v, s.rep, r);
            oo += r;
            j = qk + L;
            M &= M - 1;
            if (j >= jend)
                break;
        }
        kp.poll();
        __syncthreads();   // the image and s.tile; the prefix
        if (!kp.get(k_excl))
            return false;
        records(k_excl);
        uint8_t *const out_first = p.dst + (G0 + c0 + k_excl * diff);
        const int len = (int)out_len;
        for (int c = tid * 16; c < len; c += kThreads * 16) {
            const uint4 v = *reinterpret_cast<const uint4 *>(s.out + c);
            if (c + 16 <= len)
                __builtin_memcpy(out_first + c, &v, 16);
            else
                store_tail(out_first + c, v, len - c);
        }
        return true;
    }

    kp.poll();
    __syncthreads();   // s.tile; the prefix
    if (!kp.get(k_excl))
        return false;
    records(k_excl);
    if (!p.write_output)
        return true;
    const int64_t out_base = G0 + c0 + k_excl * diff;
    if (!IMAGE) {
        // my pieces straight to HBM: runs as 16-byte stores with the last one ending at
        // the run's end (or one exact tail), replacements from registers
        uint8_t *out = p.dst + out_base + o;
        for (;;) {
            const int qk = M ? (int)__builtin_ctzll(M) : jend;
            bf_store_run(out, s_src + j, qk - j);
            out += qk - j;
            if (qk >= jend)
                break;
            bf_put_rep_global(out, rp, rep_v, s.rep, p.rep, r);
            out += r;
            j = qk + L;
            M &= M - 1;
            if (j >= jend)
                break;
        }
    } else {
        // a group whose output does not fit the image (or a long replacement): every
        // thread writes its pieces straight to HBM
        uint8_t *out = p.dst + out_base + o;
        for (;;) {
            const int qk = M ? (int)__builtin_ctzll(M) : jend;
            if (qk > j)
                store_from_lds(out, s_src, j, qk - j);
            out += qk - j;
            if (qk >= jend)
                break;
            if (r <= kRepLds)
                store_from_lds(out, s.rep, 0, r);
            else
                store_from_global(out, p.rep, r);
            out += r;
            j = qk + L;
            M &= M - 1;
            if (j >= jend)
                break;
        }
    }
    return true;
}

// The LDS arrays of a border-free emit workgroup, declared in the kernel.
#define BF_EMIT_LDS                                                      \
    __shared__ __align__(16) uint8_t s_in[GroupShape<kGroupTiles>::kIn]; \
    __shared__ __align__(16) uint8_t s_out[kBfOutCap];                   \
    __shared__ uint8_t s_pat[kMaxSinglePassPat];                         \
    __shared__ __align__(16) uint8_t s_rep[kRepLds + 64];                \
    __shared__ uint32_t s_wsum[kWaves];                                  \
    __shared__ int s_wmax[kWaves];                                       \
    __shared__ int s_halo_last;                                          \
    __shared__ int s_tile[4];                                            \
    const BfLds lds{s_in, s_out, s_pat, s_rep, s_wsum, s_wmax, &s_halo_last, s_tile};

// Persistent emit: workgroup w takes groups first_group + w, + gridDim.x, ... and
// prefetches each next group's source into registers while it works on the current
// one, so the loads are in flight during the computation instead of before it.
// Workgroups never wait on each other.
template <bool FULL4>
__global__ void __launch_bounds__(kThreads)
bf_emit_kernel(ReplaceParams p, int64_t first_group, int64_t groups)
{
    BF_EMIT_LDS
    const int tid = threadIdx.x;
    const int64_t g_end = first_group + groups;
    for (int i = tid; i < p.L; i += kThreads)
        s_pat[i] = p.pat[i];
    for (int i = tid; i < p.r && i < kRepLds; i += kThreads)
        s_rep[i] = p.rep[i];
    int64_t g = first_group + blockIdx.x;
    BfStage stage;
    bf_prefetch<true>(p, g < g_end ? g : g_end - 1, tid, stage);
    __syncthreads();
    const uint4 rep_v = lds_load16(s_rep);
    const BfRep rp = bf_rep_parts(rep_v, p.r);

    for (; g < g_end; g += gridDim.x) {
        // the group's kept-before count came with the prefetch; the empty asm keeps it in
        // vector registers until here (held in scalar registers, it was waited for
        // right after the prefetch was issued)
        int64_t k_excl = stage.k_base;
        asm volatile("" : "+v"(k_excl));
        __syncthreads();                       // the previous group is done with s_in / s_out
        bf_stage_to_lds(stage, tid, s_in);
        if (tid == 0)
            s_halo_last = INT_MIN;
        // the next group's source (the last iteration loads its own group again)
        bf_prefetch<true>(p, g + gridDim.x < g_end ? g + gridDim.x : g, tid, stage);
        __syncthreads();
        KnownK kk{k_excl};
        bf_emit_staged<FULL4, true>(p, g, kk, lds, rep_v, rp);
    }
}

// The LDS of a border-free workgroup without an output image.
#define BF_DIRECT_LDS                                                    \
    __shared__ __align__(16) uint8_t s_in[GroupShape<kGroupTiles>::kIn]; \
    __shared__ uint8_t s_pat[kMaxSinglePassPat];                         \
    __shared__ __align__(16) uint8_t s_rep[kRepLds + 64];                \
    __shared__ uint32_t s_wsum[kWaves];                                  \
    __shared__ int s_wmax[kWaves];                                       \
    __shared__ int s_halo_last;                                          \
    __shared__ int s_tile[4];                                            \
    const BfLds lds{s_in, nullptr, s_pat, s_rep, s_wsum, s_wmax, &s_halo_last, s_tile};

// Emit with one group per workgroup and no output image: the group comes in by LDS
// DMA and every thread stores its pieces itself.  The small LDS footprint lets twice
// as many workgroups share a compute unit as with an image.
template <bool FULL4>
__global__ void __launch_bounds__(kThreads)
bf_emit_direct_kernel(ReplaceParams p, int64_t first_group)
{
    BF_DIRECT_LDS
    const int tid = threadIdx.x;
    const int64_t g = first_group + blockIdx.x;
    const int64_t t0 = g * kGroupTiles;
    if (t0 >= p.num_tiles)
        return;
    for (int i = tid; i < p.L; i += kThreads)
        s_pat[i] = p.pat[i];
    for (int i = tid; i < p.r && i < kRepLds; i += kThreads)
        s_rep[i] = p.rep[i];
    if (tid == 0)
        s_halo_last = INT_MIN;
    load_group<kGroupTiles>(p, t0, tid, true, s_in);
    const int64_t k_excl = p.chunks_pre[g].k_base;
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    const uint4 rep_v = lds_load16(s_rep);
    const BfRep rp = bf_rep_parts(rep_v, p.r);
    KnownK kk{k_excl};
    bf_emit_staged<FULL4, false>(p, g, kk, lds, rep_v, rp);
}

// ---- single-pass border-free sweep -------------------------------------------------
// One persistent launch does the whole pass.  Workgroup 0 is the scanner: its first
// wave turns the groups' kept counts into exclusive prefixes, polling 16 blocks of 64
// groups at a time.  Every other workgroup is a worker w < W = grid - 1 and
// owns group kW + w of round k.  Its phases alternate: the count phase of round k
// (stage, count, publish the count) runs `ahead` rounds before the emit phase of the
// same group (stage it again -- it was read `ahead` rounds earlier, so the read
// normally hits the memory-side cache -- wait for its prefix, emit).  A phase's data
// is prefetched into registers while the previous phase computes.  Counts and
// prefixes are 64-bit words tagged with the pass's epoch (relaxed agent-scope
// atomics).  Every wait is bounded: one that runs out sets the abort word, every
// waiter sees it and returns, and the host redoes the pass with the three launches.
struct SweepArgs {
    uint64_t *agg;        // per group: epoch << 48 | kept count
    uint64_t *pref;       // per group: epoch << 48 | kept starts before the group
    uint64_t *flag;       // epoch << 48 | 1 once a wait has given up
    uint64_t *arrive;     // epoch << 48 | 1 once the grid's last workgroup has started
    int32_t ahead;        // rounds the count phase runs ahead of the emit phase
    int32_t spin_limit;   // polls before a wait gives up
};

// Workgroups start in order of their index, so once the last one has started every
// workgroup of the grid is (or was) resident.  A wait still unanswered after this
// many polls while the last workgroup has not started gives up at once: another
// kernel holds part of the device, and the grid cannot finish until this one ends.
constexpr int kResidentPolls = 256;

__device__ __forceinline__ bool sweep_should_give_up(const SweepArgs &a, uint64_t tag, int spins)
{
    if (spins > a.spin_limit)
        return true;
    if (spins % 64 != 0)
        return false;
    return __hip_atomic_load(a.flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == (tag | 1) ||
           (spins >= kResidentPolls &&
            __hip_atomic_load(a.arrive, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != (tag | 1));
}

__device__ __forceinline__ uint64_t ld_agent(const uint64_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(uint64_t *p, uint64_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

constexpr uint64_t kTagMask = ~((1ull << kEpochShift) - 1);

// The prefix of a persistent launch's group, waited for inside its emit (KnownK's
// counterpart): thread 0 polls the scanner's word, starting from one it loaded
// earlier, and hands the value (or a give-up) to the workgroup through LDS.
struct LateK {
    const SweepArgs *a;
    uint64_t tag;
    int64_t g;
    uint64_t pv;          // thread 0: the group's prefix word as loaded earlier
    int64_t *s_k;
    int *s_abort;
    __device__ __forceinline__ void poll()
    {
        if (threadIdx.x != 0)
            return;
        uint64_t v = pv;
        int spins = 0;
        while ((v & kTagMask) != tag) {
            if (sweep_should_give_up(*a, tag, ++spins)) {
                st_agent(a->flag, tag | 1);
                *s_abort = 1;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
            v = ld_agent(a->pref + g);
        }
        *s_k = (int64_t)(v & ~kTagMask);
    }
    __device__ __forceinline__ bool get(int64_t &out) const
    {
        if (*s_abort)
            return false;
        out = *s_k;
        return true;
    }
};

__device__ __forceinline__ void bf_scanner(const ReplaceParams &p, const SweepArgs &a, int lane, uint64_t tag,
                                           int64_t G)
{
    // one poll loads the next KB blocks of 64 counts together and scans every leading
    // block whose counts are all published, so a poll at the front of the published
    // counts advances as far as they reach (not one block per round trip)
    constexpr int KB = 16;
    uint64_t running = 0;
    const int64_t nb = (G + kWave - 1) / kWave;
    int spins = 0;
    for (int64_t b = 0; b < nb;) {
        uint64_t v[KB];
#pragma unroll
        for (int i = 0; i < KB; ++i) {
            const int64_t idx = (b + i) * kWave + lane;
            v[i] = ld_agent(a.agg + (idx < G ? idx : G - 1));
        }
        int done = 0;
        bool go = true;   // wave-uniform: every block so far was published
#pragma unroll
        for (int i = 0; i < KB; ++i) {
            const int64_t idx = (b + i) * kWave + lane;
            const bool ok = idx >= G || (v[i] & kTagMask) == tag;
            go = go && b + i < nb && __builtin_amdgcn_ballot_w64(!ok) == 0;
            if (go) {
                const uint32_t c = idx < G ? (uint32_t)(v[i] & 0xffffffffu) : 0u;
                const uint32_t inc = wave_incl_add_u32(c);
                if (idx < G)
                    st_agent(a.pref + idx, tag | (running + inc - c));
                running += (uint32_t)__builtin_amdgcn_readlane((int)inc, kWave - 1);
                done = i + 1;
            }
        }
        b += done;
        if (done == 0) {
            if (sweep_should_give_up(a, tag, ++spins)) {
                if (lane == 0)
                    st_agent(a.flag, tag | 1);
                return;
            }
            __builtin_amdgcn_s_sleep(2);
        } else {
            spins = 0;
        }
    }
    if (lane == 0) {
        st_agent(&p.stats->k_total, tag | running);
    }
}

// Kept count of the staged group (all four waves), published by thread 0.
__device__ __forceinline__ void bf_count_staged(const ReplaceParams &p, int64_t g, const BfLds &s,
                                                const SweepArgs &a, uint64_t tag)
{
    const int tid = threadIdx.x;
    const int64_t gpos = g * (int64_t)kBfGroup + (int64_t)tid * kBfSeg;
    uint32_t cnt = 0;
    if (gpos < p.n) {
        const uint8_t *seg = s.in + kHalo + tid * kBfSeg;
        if (p.L <= 4)
            cnt = bf_count_seg<1>(p, seg, s.pat, gpos);
        else if (p.L <= 8)
            cnt = bf_count_seg<2>(p, seg, s.pat, gpos);
        else if (p.L <= 12)
            cnt = bf_count_seg<3>(p, seg, s.pat, gpos);
        else if (p.L <= 16)
            cnt = bf_count_seg<4>(p, seg, s.pat, gpos);
        else
            cnt = (uint32_t)__builtin_popcountll(bf_cand64<true>(p, seg, s.pat, gpos));
    }
    cnt = wave_incl_add_u32(cnt);
    if ((tid & (kWave - 1)) == kWave - 1)
        s.wsum[tid / kWave] = cnt;
    __syncthreads();
    if (tid == 0)
        st_agent(a.agg + g, tag | (uint64_t)(s.wsum[0] + s.wsum[1] + s.wsum[2] + s.wsum[3]));
}

// Count phase of the held-group launch over the group staged at s_in: everything
// its emit needs before the group's prefix (candidate masks, the wave scans of kept
// counts and last kept starts, per-wave totals in wsum/wmax, the halo carry in
// *halo_last, which thread 0 set to INT_MIN before the barrier ahead of this call),
// and the kept count, published by thread 0.
__device__ __forceinline__ BfPre bf_pre_count_staged(const ReplaceParams &p, int64_t g, const uint8_t *s_in,
                                                    const uint8_t *s_pat, const SweepArgs &a, uint64_t tag,
                                                    uint32_t *wsum, int *wmax, int *halo_last)
{
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wave = tid / kWave;
    const int64_t G0 = g * (int64_t)kBfGroup;
    const int seg = tid * kBfSeg;
    const int64_t gpos = G0 + seg;
    bf_halo_carry(p, s_in, s_pat, G0, halo_last);
    uint64_t m = 0;
    if (gpos < p.n) {
        const uint8_t *sg = s_in + kHalo + seg;
        if (p.L <= 4)
            m = bf_mask_seg<1>(p, sg, s_pat, gpos);
        else if (p.L <= 8)
            m = bf_mask_seg<2>(p, sg, s_pat, gpos);
        else if (p.L <= 12)
            m = bf_mask_seg<3>(p, sg, s_pat, gpos);
        else if (p.L <= 16)
            m = bf_mask_seg<4>(p, sg, s_pat, gpos);
        else
            m = bf_cand64<true>(p, sg, s_pat, gpos);
    }
    BfPre r;
    r.m = m;
    r.kinc = wave_incl_add_u32((uint32_t)__builtin_popcountll(m));
    r.linc = wave_incl_max_i32(m ? seg + 63 - (int)__builtin_clzll(m) : INT_MIN);
    r.wsum = wsum;
    r.wmax = wmax;
    r.halo_last = halo_last;
    if (lane == kWave - 1) {
        wsum[wave] = r.kinc;
        wmax[wave] = r.linc;
    }
    __syncthreads();
    if (tid == 0)
        st_agent(a.agg + g, tag | (uint64_t)(wsum[0] + wsum[1] + wsum[2] + wsum[3]));
    return r;
}

template <bool FULL4, bool IMAGE>
__global__ void __launch_bounds__(kThreads)
bf_sweep_kernel(ReplaceParams p, SweepArgs a)
{
    __shared__ __align__(16) uint8_t s_in[GroupShape<kGroupTiles>::kIn];
    __shared__ __align__(16) uint8_t s_out_mem[IMAGE ? kBfOutCap : 16];
    __shared__ uint8_t s_pat[kMaxSinglePassPat];
    __shared__ __align__(16) uint8_t s_rep[kRepLds + 64];
    __shared__ uint32_t s_wsum[kWaves];
    __shared__ int s_wmax[kWaves];
    __shared__ int s_halo_last;
    __shared__ int s_tile[4];
    const BfLds lds{s_in, IMAGE ? s_out_mem : nullptr, s_pat, s_rep, s_wsum, s_wmax, &s_halo_last, s_tile};
    __shared__ int64_t s_kexcl;
    __shared__ int s_abort;
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    for (int i = tid; i < p.L; i += kThreads)
        s_pat[i] = p.pat[i];
    for (int i = tid; i < p.r && i < kRepLds; i += kThreads)
        s_rep[i] = p.rep[i];
    const uint64_t tag = (uint64_t)p.epoch << kEpochShift;
    const int64_t G = p.num_groups;
    if (tid == 0 && blockIdx.x == gridDim.x - 1)
        st_agent(a.arrive, tag | 1);
    if (blockIdx.x == 0) {
        if (tid < kWave)
            bf_scanner(p, a, lane, tag, G);
        return;
    }
    const int64_t W = gridDim.x - 1;
    const int64_t w = blockIdx.x - 1;
    const int64_t D = a.ahead;
    const int64_t R = w < G ? (G - w + W - 1) / W : 0;   // rounds in which this worker has a group
    const int64_t Dp = D < R ? D : R;                   // count phases before the first emit
    const int64_t mid = R > D ? R - D : 0;              // rounds with a count and an emit phase
    const int64_t phases = 2 * R;
    // phase i: count phases of rounds 0..Dp-1, then (count k, emit k-D) for k = D..R-1,
    // then the remaining emits; returns the group, *emit tells the phase's kind
    auto phase = [&](int64_t i, bool *emit) -> int64_t {
        if (i < Dp) {
            *emit = false;
            return i * W + w;
        }
        const int64_t m = i - Dp;
        if (m < 2 * mid) {
            *emit = (m & 1) != 0;
            return (*emit ? m / 2 : D + m / 2) * W + w;
        }
        *emit = true;
        return (mid + (m - 2 * mid)) * W + w;
    };
    if (tid == 0)
        s_abort = 0;
    __syncthreads();
    const uint4 rep_v = lds_load16(s_rep);
    const BfRep rp = bf_rep_parts(rep_v, p.r);
    if (phases == 0)
        return;
    BfStage stage;
    bool emit = false;
    int64_t g = phase(0, &emit);
    uint64_t pv = 0;       // thread 0: a prefix word loaded a phase before its emit
    int64_t pv_g = -1;     // ... and its group
    if (IMAGE)
        bf_prefetch<false>(p, g, tid, stage);
    for (int64_t i = 0; i < phases; ++i) {
        g = phase(i, &emit);
        __syncthreads();                       // the previous phase is done with the LDS
        if (tid == 0 && i + 1 < phases) {
            // the next emit phase's prefix word, loaded a phase early
            bool ne = false;
            const int64_t gn = phase(i + 1, &ne);
            if (ne) {
                pv = ld_agent(a.pref + gn);
                pv_g = gn;
            }
        }
        if (IMAGE) {
            // register stage: this phase's group, then the next phase's prefetch
            bf_stage_to_lds(stage, tid, s_in);
            bool ne = false;
            const int64_t gn = i + 1 < phases ? phase(i + 1, &ne) : g;
            bf_prefetch<false>(p, gn, tid, stage);
        } else {
            // LDS DMA; the other workgroups on the unit hide the latency
            load_group<kGroupTiles>(p, g * kGroupTiles, tid, true, s_in);
        }
        if (tid == 0)
            s_halo_last = INT_MIN;
        if (!IMAGE)
            __builtin_amdgcn_s_waitcnt(0);     // the LDS DMA of the group
        __syncthreads();
        if (emit) {
            LateK lk{&a, tag, g, pv_g == g ? pv : 0ull, &s_kexcl, &s_abort};
            if (!bf_emit_staged<FULL4, IMAGE>(p, g, lk, lds, rep_v, rp))
                return;
        } else {
            bf_count_staged(p, g, lds, a, tag);
        }
    }
}

// The single launch with each group read once: a worker stages group k+1 of its own
// sequence into the second LDS slot and counts it before it emits group k from the
// first, so the emit phase finds its group still in LDS (count-ahead of one round).
// Two source slots and a smaller output image (groups whose output outgrows it store
// their pieces directly) keep three workgroups on a compute unit.  The scanner,
// tags, bounded waits and give-up word are those of bf_sweep_kernel.
constexpr int kKeepOutCap = kBfGroup + 2048;

template <bool FULL4>
__global__ void __launch_bounds__(kThreads)
bf_sweep_keep_kernel(ReplaceParams p, SweepArgs a)
{
    constexpr int kIn = GroupShape<kGroupTiles>::kIn;
    __shared__ __align__(16) uint8_t s_in[2 * kIn];
    __shared__ __align__(16) uint8_t s_out[kKeepOutCap];
    __shared__ uint8_t s_pat[kMaxSinglePassPat];
    __shared__ __align__(16) uint8_t s_rep[kRepLds + 64];
    __shared__ uint32_t s_wsum[2][kWaves];   // per slot: the count phase's totals for the emit
    __shared__ int s_wmax[2][kWaves];
    __shared__ int s_halo_last[2];
    __shared__ int s_tile[4];
    __shared__ int64_t s_kexcl;
    __shared__ int s_abort;
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    for (int i = tid; i < p.L; i += kThreads)
        s_pat[i] = p.pat[i];
    for (int i = tid; i < p.r && i < kRepLds; i += kThreads)
        s_rep[i] = p.rep[i];
    const uint64_t tag = (uint64_t)p.epoch << kEpochShift;
    const int64_t G = p.num_groups;
    if (tid == 0 && blockIdx.x == gridDim.x - 1)
        st_agent(a.arrive, tag | 1);
    if (blockIdx.x == 0) {
        if (tid < kWave)
            bf_scanner(p, a, lane, tag, G);
        return;
    }
    const int64_t W = gridDim.x - 1;
    const int64_t w = blockIdx.x - 1;
    const int64_t R = w < G ? (G - w + W - 1) / W : 0;   // groups of this worker: kW + w, k < R
    if (tid == 0)
        s_abort = 0;
    __syncthreads();
    const uint4 rep_v = lds_load16(s_rep);
    const BfRep rp = bf_rep_parts(rep_v, p.r);
    if (R == 0)
        return;
    BfStage stage;
    bf_prefetch<false>(p, w, tid, stage);
    // round k: stage and count group k+1 (slot (k+1)&1), then emit group k (slot k&1);
    // round -1 stages and counts group 0
    uint64_t pv = 0;   // thread 0: the emitted group's prefix word, loaded before the count
    BfPre pre_cur{}, pre_next{};      // what the count phases handed the emitted and the counted group
    for (int64_t k = -1; k < R; ++k) {
        const int64_t gc = (k + 1) * W + w;            // the group counted this round
        if (tid == 0 && k >= 0)
            pv = ld_agent(a.pref + k * W + w);
        if (k + 1 < R) {
            const int sl = (int)((k + 1) & 1);
            uint8_t *const slot = s_in + sl * kIn;
            __syncthreads();                           // the previous round's emit is done with slot sl
            bf_stage_to_lds(stage, tid, slot);
            if (tid == 0)
                s_halo_last[sl] = INT_MIN;
            const int64_t gn = k + 2 < R ? gc + W : gc;
            bf_prefetch<false>(p, gn, tid, stage);
            __syncthreads();
            pre_next = bf_pre_count_staged(p, gc, slot, s_pat, a, tag, s_wsum[sl], s_wmax[sl], &s_halo_last[sl]);
        }
        if (k < 0) {
            pre_cur = pre_next;
            continue;
        }
        const int64_t g = k * W + w;
        const int se = (int)(k & 1);
        if (k + 1 >= R)
            __syncthreads();                           // no count phase this round: order the last emit first
        const BfLds le{s_in + se * kIn, s_out, s_pat, s_rep, s_wsum[se], s_wmax[se], &s_halo_last[se], s_tile};
        LateK lk{&a, tag, g, pv, &s_kexcl, &s_abort};
        if (!bf_emit_staged<FULL4, true, kKeepOutCap>(p, g, lk, le, rep_v, rp, &pre_cur))
            return;
        pre_cur = pre_next;
    }
}

// ---- held-group single launch for 2-state carry tables -----------------------
// A 2-byte pattern of one repeated byte (aa, "  ", ...) runs with 2-state carry
// tables.  One persistent launch does the whole pass as bf_sweep_keep_kernel does for
// border-free patterns: worker w stages and counts group (k+1)W + w into one LDS slot
// and emits group kW + w from the other, so the source is read once.  Its count phase
// builds each thread's table over its 64 bytes (two 32-position segments), scans
// the tables across the wave with DPP and publishes the group's table packed into
// one tagged word; the scanner composes the tables in group order into each group's
// (kept starts before it, carry into it).  The emit needs that carry before it can
// choose its kept starts, so it waits first, then resolves each thread's carry from
// the wave scans kept from its count phase.
constexpr int kTblOutCap = kBfGroup + 2048;

// group table entry c: carry_out << 24 | kept (< 2^15) -> 16 packed bits
__device__ __forceinline__ uint64_t tbl2_pack(const Table<2> &t)
{
    auto e = [](uint32_t v) { return ((v >> 24) << 15) | (v & 0x7fffu); };
    return (uint64_t)(e(t.e[0]) | (e(t.e[1]) << 16));
}
__device__ __forceinline__ Table<2> tbl2_unpack(uint64_t x)
{
    Table<2> t;
    t.e[0] = (((uint32_t)(x >> 15) & 1u) << 24) | ((uint32_t)x & 0x7fffu);
    t.e[1] = (((uint32_t)(x >> 31) & 1u) << 24) | ((uint32_t)(x >> 16) & 0x7fffu);
    return t;
}
constexpr int kTblCarryShift = 40;   // prefix word: tag | carry << 40 | kept before

__device__ __forceinline__ void tbl_scanner(const ReplaceParams &p, const SweepArgs &a, int lane, uint64_t tag,
                                            int64_t G)
{
    constexpr int KB = 16;
    uint64_t k_run = 0;
    uint32_t c_run = 0;
    const int64_t nb = (G + kWave - 1) / kWave;
    int spins = 0;
    for (int64_t b = 0; b < nb;) {
        uint64_t v[KB];
#pragma unroll
        for (int i = 0; i < KB; ++i) {
            const int64_t idx = (b + i) * kWave + lane;
            v[i] = ld_agent(a.agg + (idx < G ? idx : G - 1));
        }
        int done = 0;
        bool go = true;
#pragma unroll
        for (int i = 0; i < KB; ++i) {
            const int64_t idx = (b + i) * kWave + lane;
            const bool ok = idx >= G || (v[i] & kTagMask) == tag;
            go = go && b + i < nb && __builtin_amdgcn_ballot_w64(!ok) == 0;
            if (go) {
                const Table<2> t = idx < G ? tbl2_unpack(v[i]) : tidentity<2>();
                const Table<2> inc = wave_incl_table<2>(t);
                const Table<2> ex = wave_shr1_table<2>(inc);
                const uint32_t e = tsel<2>(ex, c_run);
                if (idx < G)
                    st_agent(a.pref + idx, tag | ((uint64_t)(e >> 24) << kTblCarryShift) | (k_run + (e & 0xffffffu)));
                const uint32_t t0 = (uint32_t)__builtin_amdgcn_readlane((int)inc.e[0], kWave - 1);
                const uint32_t t1 = (uint32_t)__builtin_amdgcn_readlane((int)inc.e[1], kWave - 1);
                const uint32_t tc = c_run ? t1 : t0;
                k_run += tc & 0xffffffu;
                c_run = tc >> 24;
                done = i + 1;
            }
        }
        b += done;
        if (done == 0) {
            if (sweep_should_give_up(a, tag, ++spins)) {
                if (lane == 0)
                    st_agent(a.flag, tag | 1);
                return;
            }
            __builtin_amdgcn_s_sleep(2);
        } else {
            spins = 0;
        }
    }
    if (lane == 0)
        st_agent(&p.stats->k_total, tag | k_run);
}

// What a group's count phase hands its emit: the thread's two segment masks and its
// wave-exclusive table (the waves' inclusive tables are in LDS).
struct TblPre {
    uint32_t m0, m1;
    Table<2> ex;
};

__global__ void __launch_bounds__(kThreads)
tbl_sweep_keep_kernel(ReplaceParams p, SweepArgs a)
{
    constexpr int kIn = GroupShape<kGroupTiles>::kIn;
    __shared__ __align__(16) uint8_t s_in[2 * kIn];
    __shared__ __align__(16) uint8_t s_out[kTblOutCap];
    __shared__ uint8_t s_pat[kMaxSinglePassPat];
    __shared__ __align__(16) uint8_t s_rep[kRepLds + 64];
    __shared__ uint32_t s_wtab[2][kWaves][2];   // per slot: each wave's inclusive table
    __shared__ int s_wmax[kWaves];
    __shared__ int s_tile[4];
    __shared__ int64_t s_kexcl;
    __shared__ int s_carry;
    __shared__ int s_abort;
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wave = tid / kWave;
    for (int i = tid; i < p.L; i += kThreads)
        s_pat[i] = p.pat[i];
    for (int i = tid; i < p.r && i < kRepLds; i += kThreads)
        s_rep[i] = p.rep[i];
    const uint64_t tag = (uint64_t)p.epoch << kEpochShift;
    const int64_t G = p.num_groups;
    if (tid == 0 && blockIdx.x == gridDim.x - 1)
        st_agent(a.arrive, tag | 1);
    if (blockIdx.x == 0) {
        if (tid < kWave)
            tbl_scanner(p, a, lane, tag, G);
        return;
    }
    const int64_t W = gridDim.x - 1;
    const int64_t w = blockIdx.x - 1;
    const int64_t R = w < G ? (G - w + W - 1) / W : 0;
    if (tid == 0)
        s_abort = 0;
    __syncthreads();
    const uint4 rep_v = lds_load16(s_rep);
    const BfRep rp = bf_rep_parts(rep_v, p.r);
    if (R == 0)
        return;
    const int L = p.L;
    const int r = p.r;
    const int64_t n = p.n;
    const int64_t diff = (int64_t)r - L;
    const int b = tid * kBfSeg;   // my 64 bytes of a group
    BfStage stage;
    bf_prefetch<false>(p, w, tid, stage);
    TblPre cur{}, nxt{};
    for (int64_t k = -1; k < R; ++k) {
        const int64_t gc = (k + 1) * W + w;
        uint64_t pv = 0;
        if (tid == 0 && k >= 0)
            pv = ld_agent(a.pref + k * W + w);
        if (k + 1 < R) {
            // count phase of group gc in slot sl
            const int sl = (int)((k + 1) & 1);
            uint8_t *const slot = s_in + sl * kIn;
            __syncthreads();                           // the previous round's emit is done with slot sl
            bf_stage_to_lds(stage, tid, slot);
            const int64_t gn = k + 2 < R ? gc + W : gc;
            bf_prefetch<false>(p, gn, tid, stage);
            __syncthreads();
            // the candidates of my 64 positions from the masked quad SADs (per-position
            // compares at the input's edges)
            const int64_t gb = gc * (int64_t)kBfGroup + b;
            const uint64_t m64 = gb < n ? bf_mask_seg<1>(p, slot + kHalo + b, s_pat, gb) : 0ull;
            nxt.m0 = (uint32_t)m64;
            nxt.m1 = (uint32_t)(m64 >> 32);
            const Table<2> t = tcompose<2>(segment_table<2>(nxt.m0, L), segment_table<2>(nxt.m1, L));
            const Table<2> inc = wave_incl_table<2>(t);
            nxt.ex = wave_shr1_table<2>(inc);
            if (lane == kWave - 1) {
                s_wtab[sl][wave][0] = inc.e[0];
                s_wtab[sl][wave][1] = inc.e[1];
            }
            __syncthreads();
            if (tid == 0) {
                Table<2> grp = tidentity<2>();
#pragma unroll
                for (int v = 0; v < kWaves; ++v) {
                    Table<2> wt;
                    wt.e[0] = s_wtab[sl][v][0];
                    wt.e[1] = s_wtab[sl][v][1];
                    grp = tcompose<2>(grp, wt);
                }
                st_agent(a.agg + gc, tag | tbl2_pack(grp));
            }
        }
        if (k < 0) {
            cur = nxt;
            continue;
        }
        // emit phase of group g from slot se: wait for (kept before, carry in)
        const int64_t g = k * W + w;
        const int se = (int)(k & 1);
        if (tid == 0) {
            uint64_t v = pv;
            int spins = 0;
            while ((v & kTagMask) != tag) {
                if (sweep_should_give_up(a, tag, ++spins)) {
                    st_agent(a.flag, tag | 1);
                    s_abort = 1;
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
                v = ld_agent(a.pref + g);
            }
            s_kexcl = (int64_t)(v & ((1ull << kTblCarryShift) - 1));
            s_carry = (int)((v >> kTblCarryShift) & 0xffu);
        }
        __syncthreads();
        if (s_abort)
            return;
        const int64_t k_before = s_kexcl;
        const int c_in = s_carry;
        const int64_t t0 = g * kGroupTiles;
        const int64_t G0 = g * (int64_t)kBfGroup;
        const int valid = (int)((n - G0) < kBfGroup ? (n - G0) : kBfGroup);
        const bool last_group = G0 + kBfGroup >= n;
        Table<2> wpre = tidentity<2>(), agg = tidentity<2>();
#pragma unroll
        for (int v = 0; v < kWaves; ++v) {
            Table<2> wt;
            wt.e[0] = s_wtab[se][v][0];
            wt.e[1] = s_wtab[se][v][1];
            if (v < wave)
                wpre = tcompose<2>(wpre, wt);
            agg = tcompose<2>(agg, wt);
        }
        const uint32_t me = tsel<2>(tcompose<2>(wpre, cur.ex), (uint32_t)c_in);
        const int c_t = (int)(me >> 24);              // leading bytes of my 64 covered
        const int k_loc = (int)(me & 0xffffffu);      // kept starts of the group before my bytes
        int lst0, lst1;
        const uint32_t kept0 = greedy32_pair(cur.m0, c_t, &lst0);
        int c_mid = lst0 + L - 32;
        c_mid = c_mid < 0 ? 0 : c_mid;
        const uint32_t kept1 = greedy32_pair(cur.m1, c_mid, &lst1);
        const uint32_t ge = tsel<2>(agg, (uint32_t)c_in);
        const int c_next = last_group ? 0 : (int)(ge >> 24);
        const int64_t K = (int64_t)(ge & 0xffffffu);
        const int64_t out_len = (int64_t)valid - c_in + c_next + K * diff;
        const int64_t out_base = G0 + c_in + k_before * diff;
        const int jend = (valid - b) < kBfSeg ? ((valid - b) > 0 ? valid - b : 0) : kBfSeg;
        int j = c_t < jend ? c_t : jend;
        const int64_t o = (int64_t)(b + j - c_in) + (int64_t)k_loc * diff;   // my first output byte
        uint64_t M = ((uint64_t)kept1 << 32) | kept0;
        M = j >= 64 ? 0ull : (M & (~0ull << j));

        if (p.info) {
            // tile records: tile 0 is threads 0-127, tile 1 threads 128-255
            const int lastq = kept1 ? b + 32 + lst1 : kept0 ? b + lst0 : kNegInf;
            const int lq = __builtin_amdgcn_readlane(wave_incl_max_i32(lastq), kWave - 1);
            if (lane == 0)
                s_wmax[wave] = lq;
            if (tid == 128) {
                s_tile[0] = k_loc;
                s_tile[1] = c_t;
            }
            __syncthreads();
            if (tid == 0) {
                const int l0 = max(s_wmax[0], s_wmax[1]), l1 = max(s_wmax[2], s_wmax[3]);
                TileInfo ra;
                ra.k_excl = k_before;
                ra.last_abs = l0 >= 0 ? G0 + l0 : -1;
                ra.carry_in = c_in;
                ra.mode = kModeTable;
                p.info[t0] = ra;
                if (t0 + 1 < p.num_tiles) {
                    TileInfo rb;
                    rb.k_excl = k_before + s_tile[0];
                    rb.last_abs = l1 >= 0 ? G0 + l1 : -1;
                    rb.carry_in = s_tile[1];
                    rb.mode = kModeTable;
                    p.info[t0 + 1] = rb;
                }
            }
        }
        const uint8_t *const s_src = s_in + se * kIn + kHalo + b;
        if (!p.write_output) {
            // counts and records only
        } else if (out_len + 16 <= kTblOutCap && r <= kRepLds) {
            int oo = (int)o;
            for (;;) {
                const int q = M ? (int)__builtin_ctzll(M) : jend;
                bf_copy_run(s_out + oo, s_src + j, q - j);
                oo += q - j;
                if (q >= jend)
                    break;
                if (r == 1) {
                    // this kept start and the ones every L bytes after it: one run of
                    // replacement bytes
                    int c = 1;
                    while (q + c * L < 64 && ((M >> (q + c * L)) & 1ull))
                        ++c;
                    const uint32_t b4 = (rep_v.x & 0xffu) * 0x01010101u;
                    const uint4 fill = {b4, b4, b4, b4};
                    int f = 0;
                    for (; f + 16 <= c; f += 16)
                        lds_store16(s_out + oo + f, fill);
                    if (f < c)
                        store_tail(s_out + oo + f, fill, c - f);
                    oo += c;
                    j = q + c * L;
                } else {
                    bf_put_rep_lds(s_out + oo, rp, rep_v, s_rep, r);
                    oo += r;
                    j = q + L;
                }
                M = j >= 64 ? 0ull : (M & (~0ull << j));
                if (j >= jend)
                    break;
            }
            __syncthreads();
            uint8_t *const out_first = p.dst + out_base;
            const int len = (int)out_len;
            for (int c = tid * 16; c < len; c += kThreads * 16) {
                const uint4 v = *reinterpret_cast<const uint4 *>(s_out + c);
                if (c + 16 <= len)
                    __builtin_memcpy(out_first + c, &v, 16);
                else
                    store_tail(out_first + c, v, len - c);
            }
        } else {
            // output larger than the image (or a long replacement): every thread stores
            // its pieces straight to HBM
            uint8_t *out = p.dst + out_base + o;
            for (;;) {
                const int q = M ? (int)__builtin_ctzll(M) : jend;
                if (q > j)
                    store_from_lds(out, s_src, j, q - j);
                out += q - j;
                if (q >= jend)
                    break;
                if (r <= kRepLds)
                    store_from_lds(out, s_rep, 0, r);
                else
                    store_from_global(out, p.rep, r);
                out += r;
                j = q + L;
                M = j >= 64 ? 0ull : (M & (~0ull << j));
                if (j >= jend)
                    break;
            }
            __syncthreads();   // the pieces were read from the slot: done before it is restaged
        }
        cur = nxt;
    }
}

// ---- batches of independent documents ---------------------------------------
// The count and emit launches of a batch round run the same per-group code as a
// single buffer, on a per-document view built from the round's BatchDoc.
__device__ __forceinline__ ReplaceParams batch_view(const BatchArgs &a, int doc)
{
    const BatchDoc d = a.docs[doc];
    ReplaceParams p{};
    p.src = d.src;
    p.dst = d.dst;
    p.n = d.n;
    p.floor = 0;
    p.out_len = -1;
    p.pat = a.pool + d.pat_off;
    p.rep = a.pool + d.rep_off;
    p.L = d.L;
    p.r = d.r;
    p.pat4 = d.pat4;
    p.mask4 = d.mask4;
    p.num_tiles = (d.n + kTile - 1) / kTile;
    p.num_groups = d.ngroups;
    p.write_output = 1;
    return p;
}

template <int MODE, int LT, int GT>
__global__ void __launch_bounds__(kThreads)
batch_count_kernel(BatchArgs a)
{
    constexpr int W = MODE == kModeTable ? LT : 1;
    const BatchGroup bg = a.groups[blockIdx.x];
    const ReplaceParams p = batch_view(a, bg.doc);
    count_group<MODE, LT, GT>(p, bg.g, a.gsum + (int64_t)blockIdx.x * W);
}

template <int MODE, int LT, int GT>
__global__ void __launch_bounds__(kThreads)
batch_emit_kernel(BatchArgs a)
{
    const BatchGroup bg = a.groups[blockIdx.x];
    const ReplaceParams p = batch_view(a, bg.doc);
    emit_group<MODE, LT, GT>(p, bg.g, a.pre + blockIdx.x);
}

// One wave per document: an ordered scan of the document's group summaries, 64
// groups per step, starting from no kept start and carry 0.  Writes every group's
// exclusive state and the document's kept total.
template <int MODE, int LT>
__device__ __forceinline__ void batch_spine_doc(const BatchDoc *docs, const uint32_t *gsum, ChunkState *pre,
                                                int64_t *doc_kept, int doc)
{
    const int lane = threadIdx.x;
    const BatchDoc d = docs[doc];
    int64_t k = 0;
    uint32_t carry = 0;
    for (int64_t b = 0; b < d.ngroups; b += kWave) {
        const int64_t gi = b + lane;
        const int64_t slot = d.group0 + gi;
        if constexpr (MODE != kModeTable) {
            const uint32_t v = gi < d.ngroups ? gsum[slot] : 0u;
            uint32_t inc = v;
#pragma unroll
            for (int dd = 1; dd < kWave; dd <<= 1) {
                const uint32_t u = __shfl_up(inc, dd, kWave);
                if (lane >= dd)
                    inc += u;
            }
            if (gi < d.ngroups) {
                ChunkState cs;
                cs.k_base = k + inc - v;
                cs.carry = 0;
                cs.pad = 0;
                pre[slot] = cs;
            }
            k += __shfl(inc, kWave - 1, kWave);
        } else {
            Table<LT> t = tidentity<LT>();
            if (gi < d.ngroups) {
#pragma unroll
                for (int c = 0; c < LT; ++c)
                    t.e[c] = gsum[slot * LT + c];
            }
            Table64<LT> inc = widen<LT>(t);
#pragma unroll
            for (int dd = 1; dd < kWave; dd <<= 1) {
                Table64<LT> nb;
#pragma unroll
                for (int c = 0; c < LT; ++c)
                    nb.e[c] = __shfl_up(inc.e[c], dd, kWave);
                if (lane >= dd)
                    inc = tcompose64<LT>(nb, inc);
            }
            uint64_t e = (uint64_t)carry << kCarryShift;   // lane 0: nothing before it
#pragma unroll
            for (int c = 0; c < LT; ++c) {
                const uint64_t u = __shfl_up(inc.e[c], 1, kWave);
                if (lane > 0 && (uint32_t)c == carry)
                    e = u;
            }
            if (gi < d.ngroups) {
                ChunkState cs;
                cs.k_base = k + (int64_t)(e & kCountMask);
                cs.carry = (int)(e >> kCarryShift);
                cs.pad = 0;
                pre[slot] = cs;
            }
            const uint64_t el = tsel64<LT>(wave_broadcast_last<LT>(inc), carry);
            k += (int64_t)(el & kCountMask);
            carry = (uint32_t)(el >> kCarryShift);
        }
    }
    if (lane == 0)
        doc_kept[doc] = k;
}

template <int MODE, int LT>
__global__ void __launch_bounds__(kWave)
batch_spine_kernel(BatchArgs a)
{
    batch_spine_doc<MODE, LT>(a.docs, a.gsum, a.pre, a.doc_kept, a.class_docs[blockIdx.x]);
}

// ---- device-driven batch rounds ----------------------------------------------
constexpr int kPlanThreads = 1024;
constexpr int kPlanWaves = kPlanThreads / kWave;

__device__ __forceinline__ ReplaceParams doc_view(const BatchDoc &d, const uint8_t *pool)
{
    ReplaceParams p{};
    p.src = d.src;
    p.dst = d.dst;
    p.n = d.n;
    p.floor = 0;
    p.out_len = -1;
    p.pat = pool + d.pat_off;
    p.rep = pool + d.rep_off;
    p.L = d.L;
    p.r = d.r;
    p.pat4 = d.pat4;
    p.mask4 = d.mask4;
    p.num_tiles = (d.n + kTile - 1) / kTile;
    p.num_groups = d.ngroups;
    p.write_output = 1;
    return p;
}

__device__ __forceinline__ int64_t wave_incl_scan_i64(int64_t v, int lane)
{
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const int64_t u = __shfl_up(v, d, kWave);
        if (lane >= d)
            v += u;
    }
    return v;
}

// Exclusive scan over the workgroup (kPlanThreads) of v; *total gets the sum.
// s_w: kPlanWaves scratch slots.  Contains barriers: call from every thread.
__device__ __forceinline__ int64_t block_excl_scan_i64(int64_t v, int64_t *s_w, int64_t *total)
{
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
    const int64_t inc = wave_incl_scan_i64(v, lane);
    if (lane == kWave - 1)
        s_w[wave] = inc;
    __syncthreads();
    int64_t before = 0, sum = 0;
    for (int w = 0; w < kPlanWaves; ++w) {
        const int64_t x = s_w[w];
        before += w < wave ? x : 0;
        sum += x;
    }
    __syncthreads();   // s_w is reused by the next scan
    *total = sum;
    return before + inc - v;
}

// Round plan: every active document's descriptor and its group range in its
// class's work list (one scan over the round's list; class segments restart).
__global__ void __launch_bounds__(kPlanThreads) batch_plan_kernel(BatchRound a)
{
    __shared__ int64_t s_w[kPlanWaves];
    __shared__ int64_t s_seg[kBatchClasses + 2];   // group prefix at each segment start
    const int tid = threadIdx.x;
    if (a.ctl->overflow)
        return;
    int64_t run = 0;
    for (int base = 0; base < a.rcount; base += kPlanThreads) {
        const int k = base + tid;
        int i = 0, cls = kBatchEmpty;
        int64_t ng = 0;
        BatchDocState st{};
        BatchPair pr{};
        if (k < a.rcount) {
            i = a.rlist[k];
            st = a.states[i];
            pr = a.pairs[st.pair0 + a.round];
            cls = pr.cls;
            if (cls < kBatchClasses)
                ng = ((st.len + kTile - 1) / kTile + a.gt - 1) / a.gt;
        }
        int64_t tot;
        const int64_t ex = run + block_excl_scan_i64(ng, s_w, &tot);
        if (k < a.rcount) {
            for (int c = 0; c < kBatchClasses + 2; ++c)
                if (a.cseg[c] == k)
                    s_seg[c] = ex;
        }
        __syncthreads();
        if (k < a.rcount) {
            const int64_t g0 = ex - s_seg[cls];
            BatchDoc bd{};
            bd.src = a.src_arena + st.off;
            bd.dst = nullptr;
            bd.n = st.len;
            bd.group0 = g0;
            bd.ngroups = (int32_t)ng;
            bd.L = pr.L;
            bd.r = pr.r;
            bd.pat4 = pr.pat4;
            bd.mask4 = pr.mask4;
            bd.pat_off = pr.pat_off;
            bd.rep_off = pr.rep_off;
            a.docs[i] = bd;
        }
        run += tot;
        __syncthreads();   // s_seg entries of this chunk were read
    }
    if (tid < kBatchClasses + 2 && a.cseg[tid] >= a.rcount)
        s_seg[tid] = run;
    __syncthreads();
    if (tid < kBatchClasses)
        a.ctl->class_groups[tid] = s_seg[tid + 1] - s_seg[tid];
    if (tid == 0) {
        int64_t most = 0;
        for (int c = 0; c < kBatchClasses; ++c)
            most = s_seg[c + 1] - s_seg[c] > most ? s_seg[c + 1] - s_seg[c] : most;
        if (most > a.group_cap) {
            a.ctl->need_groups = most;
            a.ctl->overflow = a.round + 1;
        }
    }
}

// Work lists: (document, group) entries of every active counted document.
__global__ void __launch_bounds__(kThreads) batch_expand_kernel(BatchRound a)
{
    if (a.ctl->overflow)
        return;
    const int counted = a.cseg[kBatchClasses];
    for (int k = blockIdx.x; k < counted; k += gridDim.x) {
        const int i = a.rlist[k];
        const int cls = a.pairs[a.states[i].pair0 + a.round].cls;
        const BatchDoc bd = a.docs[i];
        BatchGroup *list = a.groups[cls];
        for (int g = threadIdx.x; g < bd.ngroups; g += kThreads)
            list[bd.group0 + g] = BatchGroup{i, g};
    }
}

// Count or emit over class c's work list: item = workgroup, for the first
// gridDim.x items (the host's estimate of the round's items, sized from the
// documents' lengths; one item per workgroup keeps the register use, and so the
// occupancy, of the single-buffer kernels).
template <int MODE, int LT, int GT, bool EMIT>
__global__ void __launch_bounds__(kThreads) batch_item_kernel(BatchRound a, int c)
{
    constexpr int W = MODE == kModeTable ? LT : 1;
    if (a.ctl->overflow)
        return;
    const int64_t item = blockIdx.x;
    if (item >= a.ctl->class_groups[c])
        return;
    const BatchGroup bg = a.groups[c][item];
    const ReplaceParams p = doc_view(a.docs[bg.doc], a.pool);
    if constexpr (EMIT)
        emit_group<MODE, LT, GT>(p, bg.g, a.pre[c] + item);
    else
        count_group<MODE, LT, GT>(p, bg.g, a.gsum[c] + item * W);
}

// The items past the estimate, if the documents grew beyond it: a persistent
// grid, each workgroup a contiguous share of [first, total).
template <int MODE, int LT, int GT, bool EMIT>
__global__ void __launch_bounds__(kThreads) batch_rest_kernel(BatchRound a, int c, int64_t first)
{
    constexpr int W = MODE == kModeTable ? LT : 1;
    if (a.ctl->overflow)
        return;
    const int64_t total = a.ctl->class_groups[c];
    if (total <= first)
        return;
    const int64_t per = (total - first + gridDim.x - 1) / gridDim.x;
    const int64_t lo = first + (int64_t)blockIdx.x * per;
    const int64_t hi = lo + per < total ? lo + per : total;
    for (int64_t item = lo; item < hi; ++item) {
        const BatchGroup bg = a.groups[c][item];
        const ReplaceParams p = doc_view(a.docs[bg.doc], a.pool);
        if constexpr (EMIT)
            emit_group<MODE, LT, GT>(p, bg.g, a.pre[c] + item);
        else
            count_group<MODE, LT, GT>(p, bg.g, a.gsum[c] + item * W);
        __syncthreads();   // the group functions' LDS is reused by the next item
    }
}

// Persistent spine: one wave per document of class c.
template <int MODE, int LT>
__global__ void __launch_bounds__(kWave) batch_round_spine_kernel(BatchRound a, int c)
{
    if (a.ctl->overflow)
        return;
    const int k0 = a.cseg[c], nd = a.cseg[c + 1] - a.cseg[c];
    for (int k = blockIdx.x; k < nd; k += gridDim.x)
        batch_spine_doc<MODE, LT>(a.docs, a.gsum[c], a.pre[c], a.doc_kept, a.rlist[k0 + k]);
}

// Placement: each active document's new length and destination, packed at
// 256-byte boundaries into the destination arena or, after its last pair, the
// final arena; updates the document states for the next round.  A first sweep
// sums the sizes; if they do not fit, nothing is written, the overflow flag gets
// the round and the sizes needed, and the host grows the buffers and resumes here.
__global__ void __launch_bounds__(kPlanThreads) batch_place_kernel(BatchRound a)
{
    __shared__ int64_t s_w[kPlanWaves];
    const int tid = threadIdx.x;
    BatchCtl *ctl = a.ctl;
    if (ctl->overflow)
        return;
    auto sizes = [&](int k, int *i, BatchDocState *st, bool *last, int64_t *nl) {
        *i = a.rlist[k];
        *st = a.states[*i];
        *last = st->npairs == a.round + 1;
        const BatchPair pr = a.pairs[st->pair0 + a.round];
        *nl = pr.cls == kBatchEmpty ? st->len + (st->len + 1) * (int64_t)pr.r
                                    : st->len + a.doc_kept[*i] * ((int64_t)pr.r - pr.L);
    };
    int64_t fin_total = 0, arena_total = 0;
    for (int base = 0; base < a.rcount; base += kPlanThreads) {
        const int k = base + tid;
        int64_t f = 0, ar = 0;
        if (k < a.rcount) {
            int i;
            BatchDocState st;
            bool last;
            int64_t nl;
            sizes(k, &i, &st, &last, &nl);
            const int64_t sz = (nl + 255) / 256 * 256;
            f = last ? sz : 0;
            ar = last ? 0 : sz;
        }
        int64_t tf, ta;
        (void)block_excl_scan_i64(f, s_w, &tf);
        (void)block_excl_scan_i64(ar, s_w, &ta);
        fin_total += tf;
        arena_total += ta;
    }
    const int64_t fin0 = ctl->fin_used;
    if (fin0 + fin_total > a.fin_cap || arena_total > a.dst_cap) {
        if (tid == 0) {
            ctl->need_fin = fin0 + fin_total;
            ctl->need_arena = arena_total;
            ctl->overflow = a.round + 1;
        }
        return;
    }
    int64_t fin_base = fin0, arena_base = 0, algo = 0;
    for (int base = 0; base < a.rcount; base += kPlanThreads) {
        const int k = base + tid;
        int i = 0;
        BatchDocState st{};
        bool last = false;
        int64_t nl = 0, sz = 0;
        if (k < a.rcount) {
            sizes(k, &i, &st, &last, &nl);
            sz = (nl + 255) / 256 * 256;
        }
        int64_t tf, ta;
        const int64_t exf = block_excl_scan_i64(k < a.rcount && last ? sz : 0, s_w, &tf);
        const int64_t exa = block_excl_scan_i64(k < a.rcount && !last ? sz : 0, s_w, &ta);
        if (k < a.rcount) {
            uint8_t *dst;
            if (last) {
                st.fin_off = fin_base + exf;
                dst = a.fin + st.fin_off;
            } else {
                st.off = arena_base + exa;
                dst = a.dst_arena + st.off;
            }
            algo += st.len + nl;
            st.len = nl;
            a.states[i] = st;
            a.docs[i].dst = dst;
        }
        fin_base += tf;
        arena_base += ta;
    }
    int64_t algo_total;
    (void)block_excl_scan_i64(algo, s_w, &algo_total);
    if (tid == 0) {
        ctl->fin_used = fin_base;
        ctl->algo_bytes += algo_total;
    }
}

// Empty-pattern documents of the round: the replacement at every gap.
__global__ void __launch_bounds__(kThreads) batch_empty_kernel(BatchRound a)
{
    if (a.ctl->overflow)
        return;
    const int k0 = a.cseg[kBatchEmpty], nd = a.cseg[kBatchEmpty + 1] - a.cseg[kBatchEmpty];
    for (int k = blockIdx.x; k < nd; k += gridDim.x) {
        const BatchDoc d = a.docs[a.rlist[k0 + k]];
        const uint8_t *rep = a.pool + d.rep_off;
        const int64_t r = d.r;
        for (int64_t i = threadIdx.x; i <= d.n; i += kThreads) {
            uint8_t *o = d.dst + i * (r + 1);
            for (int64_t q = 0; q < r; ++q)
                o[q] = rep[q];
            if (i < d.n)
                o[r] = d.src[i];
        }
    }
}

// ---- fused multi-pattern pass ------------------------------------------------
// Bytes [0, len) of a and b equal (len 0..16).
__device__ __forceinline__ bool head_equal(uint4 a, uint4 b, int len)
{
    uint32_t d = 0;
    const uint32_t x[4] = {a.x ^ b.x, a.y ^ b.y, a.z ^ b.z, a.w ^ b.w};
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        const int l = len - 4 * w;
        const uint32_t m = l >= 4 ? 0xffffffffu : (l <= 0 ? 0u : (1u << (8 * l)) - 1u);
        d |= x[w] & m;
    }
    return d == 0;
}

// Entry (meta | rep4 << 32) of the pattern occurring at sp (LDS; avail source bytes
// from there), or 0.  The caller has matched the shared first byte; the hash is of
// the next key_len bytes.  One table load per probe step verifies bytes 1..8 (all of
// a pattern of up to 9 bytes); longer patterns compare the rest from their first 16
// bytes and the pool, 16 bytes at a time.
__device__ __forceinline__ uint64_t fused_probe(const FusedTable *t, const uint8_t *sp, int64_t avail)
{
    if (avail < t->min_len)
        return 0;
    uint32_t k1, k2;   // source bytes 1..8 after the candidate
    __builtin_memcpy(&k1, sp + 1, 4);
    __builtin_memcpy(&k2, sp + 5, 4);
    uint32_t h = fused_hash(k1 & t->key_mask);
    for (int probe = 0; probe < kFusedSlots; ++probe) {
        const uint4 sl = t->slot[h];
        if (sl.z == 0)
            return 0;
        const int L = (int)((sl.z >> 8) & 0xff);
        const int nb = L - 1;   // pattern bytes after the first
        const uint32_t mlo = nb >= 4 ? 0xffffffffu : (1u << (8 * nb)) - 1u;
        const uint32_t mhi = nb >= 8 ? 0xffffffffu : (nb <= 4 ? 0u : (1u << (8 * (nb - 4))) - 1u);
        if ((((k1 ^ sl.x) & mlo) | ((k2 ^ sl.y) & mhi)) == 0 && L <= avail) {
            bool ok = true;
            if (L > 1 + kFusedSlotBytes) {
                const int k = (int)(sl.z & 0xff);
                ok = head_equal(lds_load16(sp), t->head[k], L < 16 ? L : 16);
                if (ok && L > 16) {
                    const uint8_t *pp = t->pool + t->pat_off[k];
                    for (int i = 16; ok && i < L; i += 16) {
                        uint4 pv;
                        __builtin_memcpy(&pv, pp + i, 16);
                        ok = head_equal(lds_load16(sp + i), pv, L - i < 16 ? L - i : 16);
                    }
                }
            }
            if (ok)
                return (uint64_t)sl.z | (uint64_t)sl.w << 32;
        }
        h = (h + 1) & (kFusedSlots - 1);
    }
    return 0;
}

// Group summary: the signed length change sum(rep_len - pat_len) over every
// occurrence starting in the group (all occurrences are kept).
__device__ __forceinline__ void fused_count_group(const ReplaceParams &p, const FusedTable *t, int64_t g,
                                                  int32_t *summary)
{
    __shared__ __align__(16) uint8_t s_in[GroupShape<kGroupTiles>::kIn];
    __shared__ int32_t s_red[kWaves];
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wave = tid / kWave;
    const int64_t t0 = g * kGroupTiles;
    if (t0 >= p.num_tiles)
        return;
    load_group<kGroupTiles>(p, t0, tid, false, s_in);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    const int seg = tid * kBytesPerThread;
    int32_t delta = 0;
    for (int i = 0; i < kGroupTiles; ++i) {
        const int64_t ti = t0 + i;
        if (ti >= p.num_tiles)
            break;
        const int64_t gseg = ti * kTile + seg;
        const uint8_t *sp = s_in + kHalo + i * kTile + seg;
        uint32_t m = first_byte32(p, s_in + kHalo + i * kTile + seg, gseg);   // the shared first byte
        uint32_t occ = 0;   // occurrences (all kept: they never overlap)
        while (m) {
            const int q = __builtin_ctz(m);
            m &= m - 1;
            const uint64_t e = fused_probe(t, sp + q, p.n - (gseg + q));
            if (e) {
                delta += fused_entry_rep_len(e) - fused_entry_pat_len(e);
                occ |= 1u << q;
            }
        }
        if (p.masks)
            p.masks[ti * (kTile / 32) + tid] = occ;
    }
    delta = (int)wave_incl_add_u32((uint32_t)delta);   // two's complement: signed sums
    if (lane == kWave - 1)
        s_red[wave] = delta;
    __syncthreads();
    if (tid == 0)
        summary[0] = s_red[0] + s_red[1] + s_red[2] + s_red[3];
}

// Copies len bytes from HBM to LDS (any alignment): 16-byte loads, exact tail.
__device__ __forceinline__ void global_to_lds(uint8_t *dst, const uint8_t *src, int len)
{
    for (; len >= 16; len -= 16, src += 16, dst += 16) {
        uint4 v;
        __builtin_memcpy(&v, src, 16);
        lds_store16(dst, v);
    }
    for (int i = 0; i < len; ++i)
        dst[i] = src[i];
}

// Emits one group of a fused pass: every occurrence of every pattern is replaced.
// pre->k_base is the signed length change of everything before the group.
__device__ __forceinline__ void fused_emit_group(const ReplaceParams &p, const FusedTable *t, int64_t g,
                                                 const ChunkState *pre)
{
    __shared__ __align__(16) uint8_t s_in[GroupShape<kGroupTiles>::kIn];
    __shared__ __align__(16) uint8_t s_out[kStage];
    __shared__ int s_wsum[kWaves];
    __shared__ int s_wmax[kWaves];
    __shared__ int s_halo_end;
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wave = tid / kWave;
    const int64_t n = p.n;
    const int64_t nt = p.num_tiles;
    const int64_t t0 = g * kGroupTiles;
    if (t0 >= nt)
        return;
    if (tid == 0)
        s_halo_end = 0;
    load_group<kGroupTiles>(p, t0, tid, true, s_in);
    int64_t k_run = pre->k_base;   // length change before the current tile
    int c_run = 0;                 // leading bytes of the current tile covered by an earlier match
    // the count's occurrence masks of my segments, loaded with the group
    uint32_t cm[kGroupTiles];
    const bool have_masks = p.masks != nullptr;
#pragma unroll
    for (int i = 0; i < kGroupTiles; ++i)
        cm[i] = (have_masks && t0 + i < nt) ? p.masks[(t0 + i) * (kTile / 32) + tid] : 0u;
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    const int seg = tid * kBytesPerThread;
    const uint8_t first = (uint8_t)p.pat4;
    for (int i = 0; i < kGroupTiles; ++i) {
        const int64_t ti = t0 + i;
        if (ti >= nt)
            break;
        const int64_t start = ti * kTile;
        const int valid = (int)((n - start) < kTile ? (n - start) : kTile);
        const bool last_tile = ti == nt - 1;
        const int seg_valid = valid - seg;
        const int jend = seg_valid < 32 ? (seg_valid > 0 ? seg_valid : 0) : 32;
        const int in_off = kHalo + i * kTile;
        const uint8_t *sp = s_in + in_off + seg;
        // my occurrences: kept mask, summed length change, end of the last one (the
        // count's mask holds only occurrences; without it, every candidate is probed)
        uint32_t m = 0;
        if (have_masks) {
#pragma unroll
            for (int k = 0; k < kGroupTiles; ++k)
                m = k == i ? cm[k] : m;
        } else {
            m = first_byte32(p, s_in + in_off + seg, start + seg);   // the shared first byte
        }
        uint32_t kept = 0;
        int d = 0;
        int e = kNegInf;
        uint64_t e0 = 0, e1 = 0;   // entries of my first two occurrences (the emit reuses them)
        int nk = 0;
        while (m) {
            const int q = __builtin_ctz(m);
            m &= m - 1;
            const uint64_t en = fused_probe(t, sp + q, n - (start + seg + q));
            if (en) {
                kept |= 1u << q;
                d += fused_entry_rep_len(en) - fused_entry_pat_len(en);
                e = seg + q + fused_entry_pat_len(en);
                e1 = nk == 1 ? en : e1;
                e0 = nk == 0 ? en : e0;
                ++nk;
            }
        }
        if (i == 0) {
            // an occurrence starting in the max_len-1 bytes left of the group covers
            // the group's first bytes
            if (ti > 0 && tid < t->max_len - 1) {
                const int q = -1 - tid;
                if (s_in[in_off + q] == first) {
                    const uint64_t en = fused_probe(t, s_in + in_off + q, n - (start + q));
                    if (en && q + fused_entry_pat_len(en) > 0)
                        atomicMax(&s_halo_end, q + fused_entry_pat_len(en));
                }
            }
            __syncthreads();
            c_run = s_halo_end;
        }
        const int c_t = c_run;
        // in-tile scans: length change (sum) and coverage end (max)
        const int ds = (int)wave_incl_add_u32((uint32_t)d);   // two's complement: signed sums
        const int le = wave_incl_max_i32(e);
        const int le_ex = wave_shr1_i32(le, kNegInf);
        if (lane == kWave - 1) {
            s_wsum[wave] = ds;
            s_wmax[wave] = le;
        }
        __syncthreads();
        int wsum = 0, wmax = c_t, tile_d = 0, tile_end = c_t;
        for (int w = 0; w < kWaves; ++w) {
            if (w < wave) {
                wsum += s_wsum[w];
                wmax = s_wmax[w] > wmax ? s_wmax[w] : wmax;
            }
            tile_d += s_wsum[w];
            tile_end = s_wmax[w] > tile_end ? s_wmax[w] : tile_end;
        }
        const int d_loc = wsum + ds - d;
        const int prev_end = le_ex > wmax ? le_ex : wmax;
        int c_i = prev_end - seg;
        if (c_i < 0)
            c_i = 0;
        int c_next = tile_end - kTile;
        if (c_next < 0 || last_tile)
            c_next = 0;
        if (p.write_output) {
            const int64_t out_len = (int64_t)valid + c_next - c_t + tile_d;
            uint8_t *const out_first = p.dst + start + c_t + k_run;
            const bool staged = out_len + 32 <= kStage;
            int64_t o = (seg + c_i - c_t) + d_loc;
            int j = c_i;
            int wi = 0;
            uint32_t M = j >= 32 ? 0u : (kept & (0xffffffffu << j));
            while (j < jend) {
                const int q = M ? __builtin_ctz(M) : jend;
                if (q > j) {
                    if (staged)
                        lds_move(s_out + o, sp + j, q - j);
                    else
                        store_from_lds(out_first + o, s_in, in_off + seg + j, q - j);
                    o += q - j;
                }
                if (q >= jend)
                    break;
                // occurrences are disjoint, so the loop meets my occurrences in order
                const uint64_t en = wi == 0 ? e0 : (wi == 1 ? e1 : fused_probe(t, sp + q, n - (start + seg + q)));
                ++wi;
                const int r = fused_entry_rep_len(en);
                const int k = fused_entry_index(en);
                if (r <= 4) {
                    const uint4 rv = {fused_entry_rep4(en), 0u, 0u, 0u};
                    if (staged)
                        store_tail(s_out + o, rv, r);
                    else
                        store_tail(out_first + o, rv, r);
                } else if (r <= 16) {
                    const uint4 rv = t->rep_head[k];
                    if (r == 16) {
                        if (staged)
                            lds_store16(s_out + o, rv);
                        else
                            __builtin_memcpy(out_first + o, &rv, 16);
                    } else {
                        if (staged)
                            store_tail(s_out + o, rv, r);
                        else
                            store_tail(out_first + o, rv, r);
                    }
                } else {
                    const uint8_t *rp = t->pool + t->rep_off[k];
                    if (staged)
                        global_to_lds(s_out + o, rp, r);
                    else
                        store_from_global(out_first + o, rp, r);
                }
                o += r;
                j = q + fused_entry_pat_len(en);
                M = j >= 32 ? 0u : (M & (0xffffffffu << j));
            }
            if (staged) {
                __syncthreads();
                const int len = (int)out_len;
                for (int c = tid * 16; c < len; c += kThreads * 16) {
                    const uint4 v = *reinterpret_cast<const uint4 *>(s_out + c);
                    if (c + 16 <= len)
                        __builtin_memcpy(out_first + c, &v, 16);
                    else
                        store_tail(out_first + c, v, len - c);
                }
            }
        }
        k_run += tile_d;
        c_run = c_next;
        __syncthreads();   // s_out, s_wsum and s_wmax are reused by the next tile
    }
}

// ---- held-group single launch of a fused multi-pattern pass -------------------
// The structure of bf_sweep_keep_kernel for the fused table: worker w counts group
// (k+1)W + w into one LDS slot and emits group kW + w from the other.  The count
// phase probes every candidate of a thread's 64 bytes (the table's shared first
// byte), keeps the occurrence mask, the first two entries, the signed length change
// and the wave scans of length changes and coverage ends, and the halo probes give
// the coverage of the group's first bytes; the scanner sums the groups' length
// changes in order (signed, 48 bits in the tagged word).  All occurrences are kept and
// disjoint, so the emit builds its LDS image before it waits for the group's prefix.
constexpr int kFusOutCap = kBfGroup + 2048;
constexpr uint64_t kMask48 = (1ull << kEpochShift) - 1;

__device__ __forceinline__ int64_t sext48(uint64_t v)
{
    return (int64_t)(v << (64 - kEpochShift)) >> (64 - kEpochShift);
}

__device__ __forceinline__ void fused_scanner(const ReplaceParams &p, const SweepArgs &a, int lane, uint64_t tag,
                                              int64_t G)
{
    constexpr int KB = 16;
    int64_t running = 0;
    const int64_t nb = (G + kWave - 1) / kWave;
    int spins = 0;
    for (int64_t b = 0; b < nb;) {
        uint64_t v[KB];
#pragma unroll
        for (int i = 0; i < KB; ++i) {
            const int64_t idx = (b + i) * kWave + lane;
            v[i] = ld_agent(a.agg + (idx < G ? idx : G - 1));
        }
        int done = 0;
        bool go = true;
#pragma unroll
        for (int i = 0; i < KB; ++i) {
            const int64_t idx = (b + i) * kWave + lane;
            const bool ok = idx >= G || (v[i] & kTagMask) == tag;
            go = go && b + i < nb && __builtin_amdgcn_ballot_w64(!ok) == 0;
            if (go) {
                const int32_t c = idx < G ? (int32_t)(uint32_t)(v[i] & 0xffffffffu) : 0;
                const int32_t inc = (int32_t)wave_incl_add_u32((uint32_t)c);   // block sums fit 32 bits
                if (idx < G)
                    st_agent(a.pref + idx, tag | ((uint64_t)(running + inc - c) & kMask48));
                running += (int32_t)__builtin_amdgcn_readlane(inc, kWave - 1);
                done = i + 1;
            }
        }
        b += done;
        if (done == 0) {
            if (sweep_should_give_up(a, tag, ++spins)) {
                if (lane == 0)
                    st_agent(a.flag, tag | 1);
                return;
            }
            __builtin_amdgcn_s_sleep(2);
        } else {
            spins = 0;
        }
    }
    if (lane == 0)
        st_agent(&p.stats->k_total, tag | ((uint64_t)running & kMask48));
}

// Positions among the 64 whose bytes start at s_seg (16-byte aligned LDS) that hold
// the byte p.pat4 & 0xff, within [floor, n - L]: per dword, the exact zero-byte test
// of (d xor f * 0x01010101) leaves bit 7 of each matching byte, and a multiply by
// 0x00204081 gathers those four bits into the top nibble (each lands on its own bit
// 28 + i, and no other product term reaches bits 28..31).
__device__ __forceinline__ uint64_t first_byte64(const ReplaceParams &p, const uint8_t *s_seg, int64_t gpos)
{
    const uint32_t f4 = (p.pat4 & 0xffu) * 0x01010101u;
    const uint4 *q = reinterpret_cast<const uint4 *>(s_seg);
    uint32_t half[2] = {0u, 0u};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint4 v = q[k];
        const uint32_t d[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const uint32_t t = d[i] ^ f4;
            const uint32_t z = ~(((t & 0x7f7f7f7fu) + 0x7f7f7f7fu) | t) & 0x80808080u;
            const int idx = 4 * k + i;
            half[idx / 8] |= ((z * 0x00204081u) >> 28) << (4 * (idx % 8));
        }
    }
    uint64_t m = ((uint64_t)half[1] << 32) | half[0];
    const int64_t lo_lim = p.floor - gpos;
    const int64_t hi_lim = p.n - (int64_t)p.L - gpos;
    if (lo_lim > 0 || hi_lim < 63) {
        if (hi_lim < 0 || lo_lim > 63 || lo_lim > hi_lim) {
            m = 0;
        } else {
            const int a = lo_lim < 0 ? 0 : (int)lo_lim;
            const int b = hi_lim > 63 ? 63 : (int)hi_lim;
            m &= (b == 63 ? ~0ull : ((2ull << b) - 1ull)) & (~0ull << a);
        }
    }
    return m;
}

// What a group's count phase hands its emit (per thread).
struct FusPre {
    uint64_t occ;        // occurrences starting in my 64 bytes
    uint64_t e0, e1;     // entries of my first two occurrences
    int d;               // my summed length change
    int ds;              // wave-inclusive sum of d
    int le;              // wave-inclusive max of my coverage end
};

__global__ void __launch_bounds__(kThreads)
fused_sweep_keep_kernel(ReplaceParams p, const FusedTable *t, SweepArgs a)
{
    constexpr int kIn = GroupShape<kGroupTiles>::kIn;
    __shared__ __align__(16) uint8_t s_in[2 * kIn];
    __shared__ __align__(16) uint8_t s_out[kFusOutCap];
    __shared__ int s_wsum[2][kWaves];
    __shared__ int s_wmax[2][kWaves];
    __shared__ int s_halo[2];
    __shared__ int64_t s_kexcl;
    __shared__ int s_abort;
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wave = tid / kWave;
    const uint64_t tag = (uint64_t)p.epoch << kEpochShift;
    const int64_t G = p.num_groups;
    if (tid == 0 && blockIdx.x == gridDim.x - 1)
        st_agent(a.arrive, tag | 1);
    if (blockIdx.x == 0) {
        if (tid < kWave)
            fused_scanner(p, a, lane, tag, G);
        return;
    }
    const int64_t W = gridDim.x - 1;
    const int64_t w = blockIdx.x - 1;
    const int64_t R = w < G ? (G - w + W - 1) / W : 0;
    if (tid == 0)
        s_abort = 0;
    __syncthreads();
    if (R == 0)
        return;
    const int64_t n = p.n;
    const uint8_t first = (uint8_t)p.pat4;
    const int b = tid * kBfSeg;   // my 64 bytes of a group
    BfStage stage;
    bf_prefetch<false>(p, w, tid, stage);
    FusPre cur{}, nxt{};
    for (int64_t k = -1; k < R; ++k) {
        const int64_t gc = (k + 1) * W + w;
        uint64_t pv = 0;
        if (tid == 0 && k >= 0)
            pv = ld_agent(a.pref + k * W + w);
        if (k + 1 < R) {
            // count phase of group gc in slot sl
            const int sl = (int)((k + 1) & 1);
            uint8_t *const slot = s_in + sl * kIn;
            __syncthreads();                           // the previous round's emit is done with slot sl
            bf_stage_to_lds(stage, tid, slot);
            if (tid == 0)
                s_halo[sl] = 0;
            const int64_t gn = k + 2 < R ? gc + W : gc;
            bf_prefetch<false>(p, gn, tid, stage);
            __syncthreads();
            const int64_t G0c = gc * (int64_t)kBfGroup;
            const int64_t gb = G0c + b;
            const uint8_t *sp = slot + kHalo + b;
            uint64_t m = gb < n ? first_byte64(p, sp, gb) : 0ull;   // the shared first byte
            nxt.occ = 0;
            nxt.e0 = nxt.e1 = 0;
            int d = 0, e = kNegInf, nk = 0;
            while (m) {
                const int q = (int)__builtin_ctzll(m);
                m &= m - 1;
                const uint64_t en = fused_probe(t, sp + q, n - (gb + q));
                if (en) {
                    nxt.occ |= 1ull << q;
                    d += fused_entry_rep_len(en) - fused_entry_pat_len(en);
                    e = b + q + fused_entry_pat_len(en);
                    nxt.e1 = nk == 1 ? en : nxt.e1;
                    nxt.e0 = nk == 0 ? en : nxt.e0;
                    ++nk;
                }
            }
            // an occurrence starting in the max_len-1 bytes left of the group covers
            // the group's first bytes
            if (G0c > 0 && tid < t->max_len - 1) {
                const int q = -1 - tid;
                if (slot[kHalo + q] == first) {
                    const uint64_t en = fused_probe(t, slot + kHalo + q, n - (G0c + q));
                    if (en && q + fused_entry_pat_len(en) > 0)
                        atomicMax(&s_halo[sl], q + fused_entry_pat_len(en));
                }
            }
            nxt.d = d;
            nxt.ds = (int)wave_incl_add_u32((uint32_t)d);   // two's complement: signed sums
            nxt.le = wave_incl_max_i32(e);
            if (lane == kWave - 1) {
                s_wsum[sl][wave] = nxt.ds;
                s_wmax[sl][wave] = nxt.le;
            }
            __syncthreads();
            if (tid == 0)
                st_agent(a.agg + gc,
                         tag | (uint64_t)(uint32_t)(s_wsum[sl][0] + s_wsum[sl][1] + s_wsum[sl][2] + s_wsum[sl][3]));
        }
        if (k < 0) {
            cur = nxt;
            continue;
        }
        // emit phase of group g from slot se
        const int64_t g = k * W + w;
        const int se = (int)(k & 1);
        if (k + 1 >= R)
            __syncthreads();                           // no count phase this round: order the last emit first
        const int64_t G0 = g * (int64_t)kBfGroup;
        const int valid = (int)((n - G0) < kBfGroup ? (n - G0) : kBfGroup);
        const bool last_group = G0 + kBfGroup >= n;
        const int c_in = s_halo[se];
        int wsum = 0, wmax = c_in, grp_d = 0, grp_end = c_in;
#pragma unroll
        for (int v = 0; v < kWaves; ++v) {
            const int ws = s_wsum[se][v], wm = s_wmax[se][v];
            if (v < wave) {
                wsum += ws;
                wmax = wm > wmax ? wm : wmax;
            }
            grp_d += ws;
            grp_end = wm > grp_end ? wm : grp_end;
        }
        const int le_ex = wave_shr1_i32(cur.le, kNegInf);
        const int d_loc = wsum + cur.ds - cur.d;
        const int prev_end = le_ex > wmax ? le_ex : wmax;
        int c_i = prev_end - b;
        c_i = c_i < 0 ? 0 : c_i;
        int c_next = grp_end - kBfGroup;
        c_next = (c_next < 0 || last_group) ? 0 : c_next;
        const int64_t out_len = (int64_t)valid + c_next - c_in + grp_d;
        const int jend = (valid - b) < kBfSeg ? ((valid - b) > 0 ? valid - b : 0) : kBfSeg;
        int j = c_i < jend ? c_i : jend;
        int64_t o = (int64_t)(b + j - c_in) + d_loc;   // my first output byte
        uint64_t M = j >= 64 ? 0ull : (cur.occ & (~0ull << j));
        const uint8_t *const sp = s_in + se * kIn + kHalo + b;
        const bool staged = p.write_output && out_len + 32 <= kFusOutCap;
        // the group's prefix (its output's offset from its input position); with the
        // image it is waited for after the image is built
        auto wait_prefix = [&]() -> bool {
            if (tid == 0) {
                uint64_t v = pv;
                int spins = 0;
                while ((v & kTagMask) != tag) {
                    if (sweep_should_give_up(a, tag, ++spins)) {
                        st_agent(a.flag, tag | 1);
                        s_abort = 1;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(2);
                    v = ld_agent(a.pref + g);
                }
                s_kexcl = sext48(v & kMask48);
            }
            __syncthreads();
            return !s_abort;
        };
        auto put_rep = [&](uint8_t *dst, bool lds, uint64_t en) {
            const int r = fused_entry_rep_len(en);
            const int kk = fused_entry_index(en);
            if (r <= 4) {
                const uint4 rv = {fused_entry_rep4(en), 0u, 0u, 0u};
                store_tail(dst, rv, r);
            } else if (r <= 16) {
                const uint4 rv = t->rep_head[kk];
                if (r == 16)
                    __builtin_memcpy(dst, &rv, 16);
                else
                    store_tail(dst, rv, r);
            } else if (lds) {
                global_to_lds(dst, t->pool + t->rep_off[kk], r);
            } else {
                store_from_global(dst, t->pool + t->rep_off[kk], r);
            }
        };
        if (!p.write_output) {
            if (!wait_prefix())
                return;
        } else if (staged) {
            int wi = 0;
            for (;;) {
                const int q = M ? (int)__builtin_ctzll(M) : jend;
                bf_copy_run(s_out + o, sp + j, q - j);
                o += q - j;
                if (q >= jend)
                    break;
                const uint64_t en = wi == 0 ? cur.e0 : (wi == 1 ? cur.e1 : fused_probe(t, sp + q, n - (G0 + b + q)));
                ++wi;
                put_rep(s_out + o, true, en);
                o += fused_entry_rep_len(en);
                j = q + fused_entry_pat_len(en);
                M = j >= 64 ? 0ull : (M & (~0ull << j));
                if (j >= jend)
                    break;
            }
            if (!wait_prefix())
                return;
            uint8_t *const out_first = p.dst + (G0 + c_in + s_kexcl);
            const int len = (int)out_len;
            for (int c = tid * 16; c < len; c += kThreads * 16) {
                const uint4 v = *reinterpret_cast<const uint4 *>(s_out + c);
                if (c + 16 <= len)
                    __builtin_memcpy(out_first + c, &v, 16);
                else
                    store_tail(out_first + c, v, len - c);
            }
        } else {
            // output larger than the image: every thread stores its pieces to HBM
            if (!wait_prefix())
                return;
            uint8_t *out = p.dst + (G0 + c_in + s_kexcl) + o;
            int wi = 0;
            for (;;) {
                const int q = M ? (int)__builtin_ctzll(M) : jend;
                if (q > j)
                    store_from_lds(out, sp, j, q - j);
                out += q - j;
                if (q >= jend)
                    break;
                const uint64_t en = wi == 0 ? cur.e0 : (wi == 1 ? cur.e1 : fused_probe(t, sp + q, n - (G0 + b + q)));
                ++wi;
                put_rep(out, false, en);
                out += fused_entry_rep_len(en);
                j = q + fused_entry_pat_len(en);
                M = j >= 64 ? 0ull : (M & (~0ull << j));
                if (j >= jend)
                    break;
            }
            __syncthreads();   // the pieces were read from the slot: done before it is restaged
        }
        cur = nxt;
    }
}

__global__ void __launch_bounds__(kThreads)
fused_count_kernel(ReplaceParams p, const FusedTable *t, int64_t first_group)
{
    const int64_t g = first_group + blockIdx.x;
    fused_count_group(p, t, g, reinterpret_cast<int32_t *>(p.gsum) + g);
}

// Six waves per SIMD, as many as its LDS allows: left alone the compiler takes 91
// VGPRs, which leaves room for five.
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(6)))
fused_emit_kernel(ReplaceParams p, const FusedTable *t, int64_t first_group)
{
    const int64_t g = first_group + blockIdx.x;
    fused_emit_group(p, t, g, p.chunks_pre + g);
}

// ---- bitmap mode: candidate bitmap -------------------------------------------
__global__ void __launch_bounds__(kThreads)
mark_bitmap_kernel(ReplaceParams p, int64_t words)
{
    const int64_t wi = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (wi >= words)
        return;
    const int64_t base = wi * 32;
    const int L = p.L;
    // the 32 positions' first four bytes from nine dword loads; the dword that holds
    // the end of the buffer is read byte by byte, so nothing past n is touched
    uint32_t d[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const int64_t o = base + 4 * k;
        if (o + 4 <= p.n) {
            __builtin_memcpy(&d[k], p.src + o, 4);
        } else {
            uint32_t v = 0;
            for (int b = 0; b < 4 && o + b < p.n; ++b)
                v |= (uint32_t)p.src[o + b] << (8 * b);
            d[k] = v;
        }
    }
    uint32_t m = 0;
#pragma unroll
    for (int j = 31; j >= 0; --j) {
        const uint32_t w = (j % 4 == 0) ? d[j / 4] : __builtin_amdgcn_alignbyte(d[j / 4 + 1], d[j / 4], j % 4);
        m = m + m + (uint32_t)(((w ^ p.pat4) & p.mask4) == 0u);
    }
    const int64_t lo = p.floor - base;
    const int64_t hi = p.n - (int64_t)L - base;
    uint32_t rm = 0;
    if (hi >= 0 && lo <= 31 && lo <= hi) {
        const int a = lo < 0 ? 0 : (int)lo;
        const int b = hi > 31 ? 31 : (int)hi;
        rm = (b == 31 ? 0xffffffffu : ((1u << (b + 1)) - 1u)) & (0xffffffffu << a);
    }
    m &= rm;
    uint32_t mm = m;
    m = 0;
    while (mm) {
        const int j = __builtin_ctz(mm);
        mm &= mm - 1;
        const int64_t pos = base + j;
        bool ok = true;
        for (int k = 4; k < L; ++k) {
            if (p.src[pos + k] != p.pat[k]) {
                ok = false;
                break;
            }
        }
        m |= (uint32_t)ok << j;
    }
    p.bitmap[wi] = m;
}

// One wave walks the candidate bitmap in order, one tile (256 words) per step,
// keeps the leftmost non-overlapping candidates, rewrites the bitmap with the kept
// starts and records each tile's carry-in and kept prefix.
__global__ void __launch_bounds__(kWave)
greedy_bitmap_kernel(ReplaceParams p, int64_t words)
{
    const int lane = threadIdx.x;
    const int L = p.L;
    constexpr int kWordsPerTile = kTile / 32;   // 256
    int64_t next = 0;        // first position a candidate may be kept at
    int64_t k = 0;
    int64_t last = -1;
    for (int64_t t = 0; t < p.num_tiles; ++t) {
        const int64_t w0 = t * kWordsPerTile;
        uint32_t w[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int64_t wi = w0 + q * kWave + lane;
            w[q] = wi < words ? p.bitmap[wi] : 0u;
        }
        if (lane == 0) {
            TileInfo ti;
            ti.k_excl = k;
            const int64_t tstart = t * (int64_t)kTile;
            ti.carry_in = (int32_t)(next > tstart ? next - tstart : 0);
            ti.last_abs = -1;
            ti.mode = kModeBitmap;
            p.info[t] = ti;
        }
        int64_t tile_last = -1;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            uint64_t nz = __ballot(w[q] != 0u);
            uint32_t keep_mine = 0;
            while (nz) {
                const int j = __builtin_ctzll(nz);
                nz &= nz - 1;
                const uint32_t wj = __shfl(w[q], j, kWave);
                const int64_t pos0 = (w0 + q * kWave + j) * 32;
                uint32_t mm = wj;
                if (next > pos0)
                    mm = (next - pos0 >= 32) ? 0u : (mm & (0xffffffffu << (int)(next - pos0)));
                uint32_t kw = 0;
                while (mm) {
                    const int b = __builtin_ctz(mm);
                    kw |= 1u << b;
                    ++k;
                    last = pos0 + b;
                    tile_last = last;
                    next = last + L;
                    const int64_t s = (int64_t)b + L;
                    mm = s >= 32 ? 0u : (mm & (0xffffffffu << (int)s));
                }
                if (lane == j)
                    keep_mine = kw;
            }
            w[q] = keep_mine;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int64_t wi = w0 + q * kWave + lane;
            if (wi < words)
                p.bitmap[wi] = w[q];
        }
        if (lane == 0)
            p.info[t].last_abs = tile_last;
    }
    if (lane == 0) {
        p.stats->k_total = ((uint64_t)p.epoch << kEpochShift) | (uint64_t)k;
        p.stats->last_kept = ((uint64_t)p.epoch << kEpochShift) | (uint64_t)(last + 1);
    }
}

// ---- bitmap mode: parallel leftmost non-overlapping selection ----------------
// The serial greedy above walks every candidate in one wave.  The parallel form
// (cuda_replace.cu:905-968 is the serial rule it reproduces) works on tiles:
//   tables   for every entry state s (the tile's first s bytes are covered by an
//            earlier kept match, s < L) the greedy over the tile's candidates gives
//            (carry out, kept count); L states, one walk each;
//   compose  64 consecutive tables at a time into their composition, level by
//            level, until at most 64 remain;
//   top      one thread walks the last level from state 0: each entry's (kept
//            before, entry state);
//   expand   each parent's entry walked over its 64 children, level by level
//            down to the tiles;
//   resolve  every tile runs its greedy from its entry state, rewrites its bitmap
//            words to the kept starts and writes its tile record.
// A border-free pattern keeps every candidate, so its tile tables have one state
// (the tile's candidate count), and a tile's carry in comes from the last
// candidate of the tiles before it.
constexpr int kBmFan = 64;

// A tile's candidate words in LDS, and per word the first nonzero word and the
// first word that is not all ones after it (kTile / 32 when none) packed beside it,
// so a walk step reads one 8-byte entry.
struct BmTile {
    uint32_t w[kTile / 32];
    uint64_t wn[kTile / 32];   // w[i] | (first nonzero > i) << 32 | (first not all ones > i) << 48
    uint64_t bnz[kTile / 32 / kWave], bnf[kTile / 32 / kWave];   // per 64 words: nonzero / not all ones
};

// First index >= i whose bit is set in the per-64-word masks b, or kTile / 32.
__device__ __forceinline__ int bm_next_set(const uint64_t *b, int i)
{
    constexpr int W = kTile / 32;
    if (i >= W)
        return W;
    int c = i >> 6;
    const uint64_t m = b[c] >> (i & 63);
    if (m)
        return i + (int)__builtin_ctzll(m);
    for (++c; c < W / kWave; ++c)
        if (b[c])
            return (c << 6) + (int)__builtin_ctzll(b[c]);
    return W;
}

// Word i of the tile (v) into s, with the wave ballots of the nonzero and the
// not-all-ones words; every lane of the wave calls it for the same 64-word span.
__device__ __forceinline__ void bm_put_word(BmTile &s, int i, uint32_t v)
{
    s.w[i] = v;
    const uint64_t bz = __builtin_amdgcn_ballot_w64(v != 0u);
    const uint64_t bf = __builtin_amdgcn_ballot_w64(v != 0xffffffffu);
    if ((i & (kWave - 1)) == 0) {
        s.bnz[i / kWave] = bz;
        s.bnf[i / kWave] = bf;
    }
}

// The packed entries from the words and ballots (after every bm_put_word).
__device__ __forceinline__ void bm_finish_tile(BmTile &s)
{
    constexpr int W = kTile / 32;
    __syncthreads();
    for (int i = threadIdx.x; i < W; i += blockDim.x)
        s.wn[i] = (uint64_t)s.w[i] | (uint64_t)bm_next_set(s.bnz, i + 1) << 32 |
                  (uint64_t)bm_next_set(s.bnf, i + 1) << 48;
    __syncthreads();
}

// Loads tile t's candidate words (zero past the buffer) and builds the packed
// entries; every thread of the block takes part (blockDim.x is a multiple of 64
// dividing 256, so each wave covers whole 64-word spans).
__device__ __forceinline__ void bm_load_tile(const ReplaceParams &p, int64_t t, BmTile &s)
{
    constexpr int W = kTile / 32;
    const int64_t words = p.num_tiles * W;
    for (int i = threadIdx.x; i < W; i += blockDim.x) {
        const int64_t wi = t * W + i;
        bm_put_word(s, i, wi < words ? p.bitmap[wi] : 0u);
    }
    bm_finish_tile(s);
}

// The greedy over tile offsets [x, end) from entry offset x (end <= kTile): kept
// count, carry past end, last kept; with MARK each kept start's bit is set in kept[].
// A step takes the first candidate c at or after x and the run of consecutive
// candidates from c (to the first clear bit, through any all-ones words): the
// greedy keeps c, c + L, ... while inside the run, ceil(len / L) starts at once,
// and goes on from the first position the last of them does not cover.  A word
// with no candidate left moves x to the next nonempty word.  All-'a' input is then
// one step per walk; a sparse or periodic input one step per kept start.
// inv = floor((2^32 - 1) / L) + 1 makes the high half of (len + L - 1) * inv equal
// ceil(len / L) for len <= kTile.
template <bool MARK>
__device__ __forceinline__ void bm_walk(const BmTile &s, int x, int end, int L, uint32_t inv, int *count, int *carry,
                                        int *last, uint32_t *kept)
{
    constexpr int W = kTile / 32;
    int cnt = 0, reach = x;
    while (x < end) {
        // a volatile LDS read: the whole entry every step rather than a branch
        // around the read of the packed indices
        const uint64_t e = *(const volatile __attribute__((address_space(3))) uint64_t *)(&s.wn[x >> 5]);
        const uint32_t w = (uint32_t)e;
        const uint32_t m = w >> (x & 31);
        const int nz = (int)((e >> 32) & 0xffffu);
        const int nf = (int)(e >> 48);
        const int c = x + (int)__builtin_ctz(m | 0x80000000u);
        const int s5 = c & 31;
        const int rl = (int)__builtin_ctzll(~(uint64_t)(w >> s5));   // <= 32 - s5
        int run_end = c + rl;
        if (s5 + rl == 32) {
            // the run reaches the word's end: on through the all-ones words
            const uint32_t wf = s.w[nf < W ? nf : W - 1];
            run_end = nf < W ? (nf << 5) + (int)__builtin_ctzll(~(uint64_t)wf) : kTile;
        }
        const int lim = run_end < end ? run_end : end;
        const bool has = m != 0u && c < end;
        const int k = has ? (int)__umulhi((uint32_t)(lim - c + L - 1), inv) : 0;
        if (MARK) {
            for (int j = 0; j < k; ++j) {
                const int q = c + j * L;
                atomicOr(&kept[q >> 5], 1u << (q & 31));
            }
        }
        cnt += k;
        const int nxt = c + k * L;
        reach = has ? nxt : reach;
        x = m != 0u ? (has ? nxt : end) : (nz << 5);
    }
    *count = cnt;
    *carry = reach > end ? reach - end : 0;
    *last = cnt ? reach - L : -1;
}

template <int NQ>
__device__ __forceinline__ uint32_t bm_mask32_qsad(const ReplaceParams &p, const uint8_t *s_seg)
{
    return mask32_qsad<NQ>(p.qref, p.qx, s_seg);
}

__device__ __forceinline__ uint32_t bm_mask32(const ReplaceParams &p, const uint8_t *s_seg, int nq)
{
    switch (nq) {
    case 1: return bm_mask32_qsad<1>(p, s_seg);
    case 2: return bm_mask32_qsad<2>(p, s_seg);
    case 3: return bm_mask32_qsad<3>(p, s_seg);
    case 4: return bm_mask32_qsad<4>(p, s_seg);
    case 5: return bm_mask32_qsad<5>(p, s_seg);
    case 6: return bm_mask32_qsad<6>(p, s_seg);
    case 7: return bm_mask32_qsad<7>(p, s_seg);
    default: return bm_mask32_qsad<8>(p, s_seg);
    }
}

// The leaf tables of tile t (self-overlapping pattern; the tile's words in s): entry
// st of a leaf = carry_out << 16 | kept (kept <= 8192, carry < L <= 256) for the
// greedy over the leaf whose first st bytes are covered.  Threads take (leaf, entry
// state) pairs in turn.
__device__ __forceinline__ void bm_leaf_tables(const ReplaceParams &p, const BmTile &s, int64_t t, uint32_t *tab,
                                               int states, int lpt)
{
    const int leaf_bytes = kTile / lpt;
    const uint32_t inv = 0xffffffffu / (uint32_t)p.L + 1u;
    for (int i = threadIdx.x; i < lpt * states; i += blockDim.x) {
        const int leaf = i / states, st = i - leaf * states;
        int cnt, carry, lst;
        bm_walk<false>(s, leaf * leaf_bytes + st, (leaf + 1) * leaf_bytes, p.L, inv, &cnt, &carry, &lst, nullptr);
        tab[(t * lpt + leaf) * states + st] = ((uint32_t)carry << 16) | (uint32_t)cnt;
    }
}

// Candidates of tile t from LDS (L <= kHalo), one 32-position word per thread.  A
// pattern of up to 32 bytes takes the masked quad SADs (ceil(L / 4) SAD instructions
// per four positions).  A longer one compares the head word at every position as the
// count kernels do, then pattern word k (bytes 4k..4k+3, masked past L) at every
// position, while any lane of the wave still has a candidate.  With TABLES (a
// self-overlapping pattern) the same workgroup then builds the tile's leaf tables
// from the words it holds, instead of a second kernel reading them back.
template <bool TABLES>
__global__ void __launch_bounds__(kThreads) bm_mark_lds_kernel(ReplaceParams p, uint32_t *tab, int states, int lpt)
{
    __shared__ __align__(16) uint8_t s_in[GroupShape<1>::kIn];
    __shared__ __align__(16) uint8_t s_pat[kMaxSinglePassPat + 4];
    const int64_t t = blockIdx.x;
    const int tid = threadIdx.x;
    const int L = p.L;
    for (int i = tid; i < kMaxSinglePassPat + 4; i += kThreads)
        s_pat[i] = i < L ? p.pat[i] : 0;
    load_group<1>(p, t, tid, false, s_in);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    const int seg = tid * 32;
    const int64_t gseg = t * kTile + (int64_t)seg;
    uint32_t m = 0;
    if (gseg < p.n) {
        const int64_t lo = p.floor - gseg;
        const int64_t hi = p.n - (int64_t)L - gseg;
        if (hi >= 0 && lo <= 31 && lo <= hi) {
            const int a = lo < 0 ? 0 : (int)lo;
            const int b = hi > 31 ? 31 : (int)hi;
            m = (b == 31 ? 0xffffffffu : ((1u << (b + 1)) - 1u)) & (0xffffffffu << a);
        }
    }
    const int words = (L + 3) / 4;
    if (L <= 32) {
        if (__builtin_amdgcn_ballot_w64(m != 0u) != 0)
            m &= bm_mask32(p, s_in + kHalo + seg, words);
    } else {
        for (int k = 0; k < words; ++k) {
            if (__builtin_amdgcn_ballot_w64(m != 0u) == 0)
                break;
            const uint32_t pw = *reinterpret_cast<const uint32_t *>(s_pat + 4 * k);
            const int left = L - 4 * k;
            const uint32_t pm = left >= 4 ? 0xffffffffu : (0xffffffffu >> (8 * (4 - left)));
            const uint32_t *w32 = reinterpret_cast<const uint32_t *>(s_in + kHalo + seg + 4 * k);
            uint32_t d[9];
#pragma unroll
            for (int q = 0; q < 9; ++q)
                d[q] = w32[q];
            uint32_t mk = 0;
#pragma unroll
            for (int j = 31; j >= 0; --j) {
                const uint32_t w = (j % 4 == 0) ? d[j / 4] : __builtin_amdgcn_alignbyte(d[j / 4 + 1], d[j / 4], j % 4);
                mk = mk + mk + (uint32_t)(((w ^ pw) & pm) == 0u);
            }
            m &= mk;
        }
    }
    if (gseg >= p.n)
        m = 0u;
    p.bitmap[t * (kTile / 32) + tid] = m;
    if constexpr (TABLES) {
        __shared__ BmTile s;
        bm_put_word(s, tid, m);
        bm_finish_tile(s);
        bm_leaf_tables(p, s, t, tab, states, lpt);
    }
}

// Level-0 tables, one per leaf (a tile holds lpt leaves of kTile / lpt bytes): entry s
// of a leaf = carry_out << 16 | kept (kept <= 8192, carry < L <= 256) for the greedy
// over the leaf whose first s bytes are covered.  Border-free (one state, lpt = 1):
// the tile's candidate count, and its last candidate (tile offset, or -1) in tlast.
__global__ void __launch_bounds__(kThreads) bm_tables_kernel(ReplaceParams p, uint32_t *tab, int states, int lpt,
                                                              int32_t *tlast)
{
    __shared__ BmTile s;
    const int64_t t = blockIdx.x;
    bm_load_tile(p, t, s);
    const int tid = threadIdx.x;
    if (tlast) {
        // border-free: every candidate is kept
        const uint32_t v = s.w[tid];
        uint32_t c = (uint32_t)__builtin_popcount(v);
        int l = v ? tid * 32 + 31 - __builtin_clz(v) : -1;
        c = wave_sum_u32(c);
#pragma unroll
        for (int d = kWave / 2; d > 0; d >>= 1)
            l = max(l, __shfl_xor(l, d, kWave));
        __shared__ uint32_t s_c[kWaves];
        __shared__ int s_l[kWaves];
        if ((tid & (kWave - 1)) == 0) {
            s_c[tid / kWave] = c;
            s_l[tid / kWave] = l;
        }
        __syncthreads();
        if (tid == 0) {
            tab[t] = s_c[0] + s_c[1] + s_c[2] + s_c[3];
            tlast[t] = max(max(s_l[0], s_l[1]), max(s_l[2], s_l[3]));
        }
        return;
    }
    bm_leaf_tables(p, s, t, tab, states, lpt);
}

// Entry e of a table at a level: level 0 packs carry << 16 | count in 32 bits,
// higher levels carry << 32 | count in 64.
__device__ __forceinline__ void bm_entry(const uint32_t *t32, const uint64_t *t64, int64_t idx, uint32_t *carry,
                                         uint64_t *count)
{
    if (t32) {
        const uint32_t e = t32[idx];
        *carry = e >> 16;
        *count = e & 0xffffu;
    } else {
        const uint64_t e = t64[idx];
        *carry = (uint32_t)(e >> 32);
        *count = e & 0xffffffffull;
    }
}

// Tables of at most this many states are staged in LDS by the compose and expand
// kernels (64 children x states x 8 bytes <= 32 KiB of dynamic LDS): all of a
// block's child entries are loaded at once, then the walks run over LDS instead of a
// chain of 64 dependent global loads.
constexpr int kBmStageStates = 64;

static size_t bm_stage_bytes(int states)
{
    return states <= kBmStageStates ? (size_t)kBmFan * (size_t)states * sizeof(uint64_t) : 0;
}

// Block b's children's entries (first .. first + nch - 1, every state) into s_t as
// carry << 32 | count.
__device__ __forceinline__ void bm_stage_children(const uint32_t *in32, const uint64_t *in64, int64_t first,
                                                  int nch, int states, uint64_t *s_t)
{
    for (int i = threadIdx.x; i < nch * states; i += blockDim.x) {
        uint32_t c;
        uint64_t k;
        bm_entry(in32, in64, first * states + i, &c, &k);
        s_t[i] = (uint64_t)c << 32 | k;
    }
    __syncthreads();
}

// out[b] = in[64b] then in[64b+1] ... then in[64b+63] (composition), per state.
__global__ void __launch_bounds__(kThreads) bm_compose_kernel(const uint32_t *in32, const uint64_t *in64,
                                                               int64_t n_in, int states, uint64_t *out)
{
    extern __shared__ uint64_t s_t[];
    const int64_t b = blockIdx.x;
    const int64_t first = b * kBmFan;
    const int nch = (int)(n_in - first < kBmFan ? n_in - first : kBmFan);
    const bool staged = states <= kBmStageStates;
    if (staged)
        bm_stage_children(in32, in64, first, nch, states, s_t);
    for (int s0 = threadIdx.x; s0 < states; s0 += blockDim.x) {
        uint32_t st = (uint32_t)s0;
        uint64_t cnt = 0;
        for (int j = 0; j < nch; ++j) {
            uint32_t c;
            uint64_t k;
            if (staged) {
                const uint64_t e = s_t[j * states + (int)st];
                c = (uint32_t)(e >> 32);
                k = e & 0xffffffffull;
            } else {
                bm_entry(in32, in64, (first + j) * states + st, &c, &k);
            }
            cnt += k;
            st = c;
        }
        out[b * states + s0] = ((uint64_t)st << 32) | cnt;
    }
}

// One thread walks the last level from state 0: ent[i] = (kept before, entry state).
__global__ void bm_top_kernel(const uint32_t *in32, const uint64_t *in64, int64_t n, int states, ChunkState *ent,
                              LaunchStats *stats, uint32_t epoch)
{
    if (threadIdx.x != 0)
        return;
    int64_t k = 0;
    uint32_t st = 0;
    for (int64_t i = 0; i < n; ++i) {
        ChunkState cs;
        cs.k_base = k;
        cs.carry = (int)st;
        cs.pad = 0;
        ent[i] = cs;
        uint32_t c;
        uint64_t kk;
        bm_entry(in32, in64, i * states + st, &c, &kk);
        k += (int64_t)kk;
        st = c;
    }
    stats->k_total = ((uint64_t)epoch << kEpochShift) | ((uint64_t)k & ((1ull << kEpochShift) - 1));
}

// Children of parent b get their entries by walking its 64 tables from b's entry
// (thread 0, over the staged entries when they fit LDS).
__global__ void bm_expand_kernel(const uint32_t *in32, const uint64_t *in64, int64_t n_in, int states,
                                 const ChunkState *parent, ChunkState *child)
{
    extern __shared__ uint64_t s_t[];
    const int64_t b = blockIdx.x;
    const int64_t first = b * kBmFan;
    const int nch = (int)(n_in - first < kBmFan ? n_in - first : kBmFan);
    const bool staged = states <= kBmStageStates;
    if (staged)
        bm_stage_children(in32, in64, first, nch, states, s_t);
    if (threadIdx.x != 0)
        return;
    int64_t k = parent[b].k_base;
    uint32_t st = (uint32_t)parent[b].carry;
    for (int j = 0; j < nch; ++j) {
        ChunkState cs;
        cs.k_base = k;
        cs.carry = (int)st;
        cs.pad = 0;
        child[first + j] = cs;
        uint32_t c;
        uint64_t kk;
        if (staged) {
            const uint64_t e = s_t[j * states + (int)st];
            c = (uint32_t)(e >> 32);
            kk = e & 0xffffffffull;
        } else {
            bm_entry(in32, in64, (first + j) * states + st, &c, &kk);
        }
        k += (int64_t)kk;
        st = c;
    }
}

// The kept count of the whole pass, written by bm_top_kernel earlier in the stream.
__device__ __forceinline__ int64_t bm_total(const ReplaceParams &p)
{
    return (int64_t)(p.stats->k_total & ((1ull << kEpochShift) - 1));
}

// Self-overlapping: each leaf of tile t runs its greedy from its entry state (lane
// = leaf), the tile's bitmap words are rewritten to the kept starts and its record
// is written.  Only the last tile holding kept starts (the kept before it plus its
// own equal the pass's total) reports the last kept start: one atomic per pass
// instead of one per tile on a single address.
__global__ void __launch_bounds__(kWave) bm_resolve_kernel(ReplaceParams p, const ChunkState *ent, int lpt)
{
    __shared__ BmTile s;
    __shared__ uint32_t s_kept[kTile / 32];
    __shared__ int s_last;
    const int64_t t = blockIdx.x;
    const int lane = threadIdx.x;
    bm_load_tile(p, t, s);
    for (int i = lane; i < kTile / 32; i += kWave)
        s_kept[i] = 0u;
    if (lane == 0)
        s_last = -1;
    __syncthreads();
    int cnt = 0;
    if (lane < lpt) {
        const int leaf_bytes = kTile / lpt;
        const ChunkState e = ent[t * lpt + lane];
        const uint32_t inv = 0xffffffffu / (uint32_t)p.L + 1u;
        int carry, lst;
        bm_walk<true>(s, lane * leaf_bytes + e.carry, (lane + 1) * leaf_bytes, p.L, inv, &cnt, &carry, &lst,
                      s_kept);
        if (lst >= 0)
            atomicMax(&s_last, lst);
    }
    const int64_t tile_kept = (int64_t)wave_sum_u32((uint32_t)cnt);
    __syncthreads();
    const int64_t words = p.num_tiles * (kTile / 32);
    for (int i = lane; i < kTile / 32; i += kWave) {
        const int64_t wi = t * (kTile / 32) + i;
        if (wi < words)
            p.bitmap[wi] = s_kept[i];
    }
    if (lane == 0) {
        const ChunkState e = ent[t * lpt];
        const int64_t tstart = t * (int64_t)kTile;
        TileInfo ti;
        ti.k_excl = e.k_base;
        ti.last_abs = s_last >= 0 ? tstart + s_last : -1;
        ti.carry_in = e.carry;
        ti.mode = kModeBitmap;
        p.info[t] = ti;
        if (s_last >= 0 && e.k_base + tile_kept == bm_total(p))
            atomicMax((unsigned long long *)&p.stats->last_kept,
                      (unsigned long long)(((uint64_t)p.epoch << kEpochShift) | (uint64_t)(tstart + s_last + 1)));
    }
}

// Border-free: tile records from the entries (kept before), the tile's last
// candidate and the carry from the latest candidate of the tiles before it (a
// candidate ends at most L - 1 bytes later, so only tiles within L bytes matter).
__global__ void __launch_bounds__(kThreads) bm_bf_records_kernel(ReplaceParams p, const ChunkState *ent,
                                                                  const int32_t *tlast)
{
    const int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (t >= p.num_tiles)
        return;
    const int64_t tstart = t * (int64_t)kTile;
    int64_t carry = 0;
    for (int64_t u = t - 1; u >= 0 && (u + 1) * (int64_t)kTile + p.L > tstart; --u) {
        if (tlast[u] >= 0) {
            const int64_t end = u * (int64_t)kTile + tlast[u] + p.L;
            carry = end > tstart ? end - tstart : 0;
            break;
        }
    }
    TileInfo ti;
    ti.k_excl = ent[t].k_base;
    ti.last_abs = tlast[t] >= 0 ? tstart + tlast[t] : -1;
    ti.carry_in = (int32_t)(carry < INT_MAX ? carry : INT_MAX);
    ti.mode = kModeBitmap;
    p.info[t] = ti;
    // only the last tile with candidates reports (every candidate is kept)
    if (tlast[t] >= 0 && (t + 1 == p.num_tiles || ent[t + 1].k_base == bm_total(p)))
        atomicMax((unsigned long long *)&p.stats->last_kept,
                  (unsigned long long)(((uint64_t)p.epoch << kEpochShift) | (uint64_t)(tstart + tlast[t] + 1)));
}

// ---- empty pattern: bytes.replace(b"", rep) ---------------------------------
__global__ void empty_pattern_kernel(ReplaceParams p)
{
    const int64_t r = p.r;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= p.n; i += stride) {
        uint8_t *o = p.dst + i * (r + 1);
        for (int64_t q = 0; q < r; ++q)
            o[q] = p.rep[q];
        if (i < p.n)
            o[r] = p.src[i];
    }
}

__global__ void fill_u64_kernel(uint64_t *ptr, uint64_t value, size_t count)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride)
        ptr[i] = value;
}

// ---- prefix query -----------------------------------------------------------
__device__ bool match_at(const ReplaceParams &p, int64_t pos)
{
    if (pos < p.floor || pos + p.L > p.n)
        return false;
    for (int k = 0; k < p.L; ++k)
        if (p.src[pos + k] != p.pat[k])
            return false;
    return true;
}

// out2[0] = kept starts < limit, out2[1] = last kept start < limit (or -1)
__global__ void __launch_bounds__(kThreads)
query_kernel(ReplaceParams p, int mode, int64_t limit, int64_t *out2)
{
    __shared__ uint32_t s_mask[kThreads];
    __shared__ unsigned long long s_last_before;
    __shared__ int s_part_last;
    __shared__ int s_part_count;
    const int tid = threadIdx.x;
    const int64_t t = (limit - 1) / kTile;
    const int64_t start = t * kTile;
    if (tid == 0) {
        s_last_before = 0;
        s_part_last = kNegInf;
        s_part_count = 0;
    }
    __syncthreads();
    // last kept start before tile t: walk back kThreads tiles at a time and stop at
    // the first block of tiles that holds one (records are in position order)
    for (int64_t b = t - 1; b >= 0; b -= kThreads) {
        const int64_t i = b - tid;
        if (i >= 0) {
            const int64_t la = p.info[i].last_abs;
            if (la >= 0)
                atomicMax(&s_last_before, (unsigned long long)(la + 1));
        }
        __syncthreads();
        const bool found = s_last_before != 0;
        __syncthreads();
        if (found)
            break;
    }
    const int64_t gseg = start + (int64_t)tid * 32;
    uint32_t m = 0;
    if (mode == kModeBitmap) {
        m = p.bitmap[t * (kTile / 32) + tid];
    } else {
        for (int j = 0; j < 32; ++j)
            if (match_at(p, gseg + j))
                m |= 1u << j;
    }
    s_mask[tid] = m;
    __syncthreads();
    if (mode == kModeTable) {
        if (tid == 0) {
            int64_t pos = start + p.info[t].carry_in;    // first position a candidate may be kept
            int cnt = 0, lst = kNegInf;
            for (int w = 0; w < kThreads; ++w) {
                const int64_t b = start + (int64_t)w * 32;
                uint32_t mm = s_mask[w];
                if (pos > b)
                    mm = pos - b >= 32 ? 0u : (mm & (0xffffffffu << (int)(pos - b)));
                while (mm) {
                    const int q = __builtin_ctz(mm);
                    const int64_t g = b + q;
                    if (g >= limit) {
                        mm = 0;
                        break;
                    }
                    ++cnt;
                    lst = (int)(g - start);
                    pos = g + p.L;
                    mm = pos - b >= 32 ? 0u : (mm & (0xffffffffu << (int)(pos - b)));
                }
            }
            s_part_count = cnt;
            s_part_last = lst;
        }
    } else {
        // kept == mask for border-free and bitmap modes
        uint32_t mm = m;
        const int64_t lim_rel = limit - gseg;
        if (lim_rel <= 0)
            mm = 0;
        else if (lim_rel < 32)
            mm &= (1u << lim_rel) - 1u;
        if (mm) {
            atomicAdd(&s_part_count, __builtin_popcount(mm));
            atomicMax(&s_part_last, (int)(gseg - start) + 31 - __builtin_clz(mm));
        }
    }
    __syncthreads();
    if (tid == 0) {
        out2[0] = p.info[t].k_excl + s_part_count;
        out2[1] = s_part_last != kNegInf ? start + s_part_last : (int64_t)s_last_before - 1;
    }
}

// Largest kept start over all tiles (from the per-tile records) + 1, max-reduced
// into *out1 (cleared beforehand); 0 when nothing was kept.
__global__ void __launch_bounds__(kThreads)
last_kept_kernel(const TileInfo *info, int64_t num_tiles, unsigned long long *out1)
{
    unsigned long long best = 0;
    const int64_t stride = (int64_t)gridDim.x * kThreads;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < num_tiles; i += stride) {
        const int64_t la = info[i].last_abs;
        if ((unsigned long long)(la + 1) > best)
            best = (unsigned long long)(la + 1);
    }
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) {
        const unsigned long long o = __shfl_xor(best, d, kWave);
        best = o > best ? o : best;
    }
    if ((threadIdx.x & (kWave - 1)) == 0 && best)
        atomicMax(out1, best);
}

// ---- presence index: 256-bit byte-value mask per chunk -----------------------
// Grid-stride over the chunks (64-bit indices), so any chunk count fits the grid.
__global__ void __launch_bounds__(kThreads)
presence_kernel(const uint8_t *src, int64_t n, int64_t chunk, uint64_t *masks, int64_t num_chunks)
{
    __shared__ unsigned int s_bits[8];
    for (int64_t c = blockIdx.x; c < num_chunks; c += gridDim.x) {
        if (threadIdx.x < 8)
            s_bits[threadIdx.x] = 0;
        __syncthreads();
        const int64_t b = c * chunk;
        int64_t e = b + chunk;
        if (e > n)
            e = n;
        unsigned int local[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int64_t i = b + threadIdx.x; i < e; i += kThreads) {
            const uint8_t v = src[i];
            local[v >> 5] |= 1u << (v & 31);
        }
        for (int q = 0; q < 8; ++q)
            if (local[q])
                atomicOr(&s_bits[q], local[q]);
        __syncthreads();
        if (threadIdx.x < 4)
            masks[c * 4 + threadIdx.x] =
                (uint64_t)s_bits[2 * threadIdx.x] | ((uint64_t)s_bits[2 * threadIdx.x + 1] << 32);
        __syncthreads();   // s_bits is cleared for the next chunk
    }
}

// ---- launchers --------------------------------------------------------------
template <int MODE, int LT, int GT>
hipError_t count_t(const ReplaceParams &p, int64_t first_group, int64_t groups, hipStream_t s)
{
    hipLaunchKernelGGL((count_kernel<MODE, LT, GT>), dim3((unsigned)groups), dim3(kThreads), 0, s, p, first_group);
    return hipGetLastError();
}

template <int MODE, int LT, int GT>
hipError_t emit_t(const ReplaceParams &p, int64_t first_group, int64_t groups, hipStream_t s)
{
    hipLaunchKernelGGL((emit_kernel<MODE, LT, GT>), dim3((unsigned)groups), dim3(kThreads), 0, s, p, first_group);
    return hipGetLastError();
}

template <int MODE, int LT>
hipError_t spine_t(const ReplaceParams &p, int64_t first_group, int64_t groups, int64_t chunk, int last_chunk,
                   hipStream_t s)
{
    hipLaunchKernelGGL((spine_kernel<MODE, LT>), dim3(1), dim3(kSpineThreads), 0, s, p, first_group, groups, chunk,
                       last_chunk);
    return hipGetLastError();
}

// Calls F::template run<MODE, LT, GT>() for a runtime (mode, table size, tiles per
// group); bitmap mode exists for kGroupTiles only.
template <class F>
hipError_t dispatch_mode(int mode, int table_size, int gt, F &&f)
{
    auto by_gt = [&](auto mode_c, auto lt_c) -> hipError_t {
        constexpr int M = decltype(mode_c)::value, LT = decltype(lt_c)::value;
        switch (gt) {
        case 1:
            return f.template operator()<M, LT, 1>();
        case 2:
            return f.template operator()<M, LT, 2>();
        case 4:
            return f.template operator()<M, LT, 4>();
        default:
            return hipErrorInvalidValue;
        }
    };
    using BF = std::integral_constant<int, kModeBorderFree>;
    using TB = std::integral_constant<int, kModeTable>;
    if (mode == kModeBorderFree)
        return by_gt(BF{}, std::integral_constant<int, 1>{});
    if (mode != kModeTable)
        return hipErrorInvalidValue;
    switch (table_size) {
    case 2:
        return by_gt(TB{}, std::integral_constant<int, 2>{});
    case 4:
        return by_gt(TB{}, std::integral_constant<int, 4>{});
    case 8:
        return by_gt(TB{}, std::integral_constant<int, 8>{});
    case 16:
        return by_gt(TB{}, std::integral_constant<int, 16>{});
    default:
        return hipErrorInvalidValue;
    }
}

hipError_t launch_count(const ReplaceParams &p, int mode, int table_size, int gt, int64_t first_group, int64_t groups,
                        hipStream_t s)
{
    if (groups <= 0)
        return hipSuccess;
    return dispatch_mode(mode, table_size, gt, [&]<int M, int LT, int GT>() {
        return count_t<M, LT, GT>(p, first_group, groups, s);
    });
}

template <int MODE, int LT>
hipError_t spine_blocks_t(const ReplaceParams &p, int64_t groups, hipStream_t s)
{
    constexpr int64_t kPass = spine_block_groups<MODE, LT>();
    const int64_t nb = (groups + kPass - 1) / kPass;
    if constexpr (MODE == kModeTable) {
        // one launch while every lane composes at most kSmallSpineV groups (fewer for
        // 8- and 16-state tables, whose compositions cost O(states^2))
        constexpr int kSmallSpineV = LT <= 4 ? 16 : 4;
        if (groups <= (int64_t)kSpineThreads * kSmallSpineV) {
            const int v = (int)((groups + kSpineThreads - 1) / kSpineThreads);
            const int64_t lanes = (groups + v - 1) / v;
            const unsigned threads = (unsigned)((lanes + kWave - 1) / kWave * kWave);
            hipLaunchKernelGGL((spine_small_kernel<LT>), dim3(1), dim3(threads), 0, s, p, groups, v);
            return hipGetLastError();
        }
    }
    if (nb <= 1)   // one block: the single-workgroup spine is the whole job
        return spine_t<MODE, LT>(p, 0, groups, 0, 1, s);
    hipLaunchKernelGGL((spine_reduce_kernel<MODE, LT>), dim3((unsigned)nb), dim3(kSpineThreads), 0, s, p, groups);
    hipLaunchKernelGGL((spine_top_kernel<MODE, LT>), dim3(1), dim3(kWave), 0, s, p, nb);
    hipLaunchKernelGGL((spine_kernel<MODE, LT, true>), dim3((unsigned)nb), dim3(kSpineThreads), 0, s, p, (int64_t)0,
                       groups, (int64_t)0, 0);
    return hipGetLastError();
}

int64_t spine_blocks_needed(int64_t groups)
{
    // the smallest block has 64 x 1 x 16 groups
    return groups / ((int64_t)kWave * kSpineWaves) + 2;
}

hipError_t launch_spine_blocks(const ReplaceParams &p, int mode, int table_size, int64_t groups, hipStream_t s)
{
    if (groups <= 0)
        return hipSuccess;
    if (mode == kModeBorderFree)
        return spine_blocks_t<kModeBorderFree, 1>(p, groups, s);
    if (mode == kModeFused)
        return spine_blocks_t<kModeFused, 1>(p, groups, s);
    if (mode != kModeTable)
        return hipErrorInvalidValue;
    switch (table_size) {
    case 2:
        return spine_blocks_t<kModeTable, 2>(p, groups, s);
    case 4:
        return spine_blocks_t<kModeTable, 4>(p, groups, s);
    case 8:
        return spine_blocks_t<kModeTable, 8>(p, groups, s);
    case 16:
        return spine_blocks_t<kModeTable, 16>(p, groups, s);
    default:
        return hipErrorInvalidValue;
    }
}

hipError_t launch_spine(const ReplaceParams &p, int mode, int table_size, int64_t first_group, int64_t groups,
                        int64_t chunk, int last_chunk, hipStream_t s)
{
    if (groups <= 0)
        return hipSuccess;
    if (mode == kModeBorderFree)
        return spine_t<kModeBorderFree, 1>(p, first_group, groups, chunk, last_chunk, s);
    if (mode != kModeTable)
        return hipErrorInvalidValue;
    switch (table_size) {
    case 2:
        return spine_t<kModeTable, 2>(p, first_group, groups, chunk, last_chunk, s);
    case 4:
        return spine_t<kModeTable, 4>(p, first_group, groups, chunk, last_chunk, s);
    case 8:
        return spine_t<kModeTable, 8>(p, first_group, groups, chunk, last_chunk, s);
    case 16:
        return spine_t<kModeTable, 16>(p, first_group, groups, chunk, last_chunk, s);
    default:
        return hipErrorInvalidValue;
    }
}

hipError_t launch_emit(const ReplaceParams &p, int mode, int table_size, int gt, int64_t first_group, int64_t groups,
                       hipStream_t s)
{
    if (groups <= 0)
        return hipSuccess;
    if (mode == kModeBitmap)
        return gt == kGroupTiles ? emit_t<kModeBitmap, 1, kGroupTiles>(p, first_group, groups, s)
                                 : hipErrorInvalidValue;
    return dispatch_mode(mode, table_size, gt, [&]<int M, int LT, int GT>() {
        return emit_t<M, LT, GT>(p, first_group, groups, s);
    });
}

hipError_t launch_bf_count(const ReplaceParams &p, int64_t first_group, int64_t groups, hipStream_t s)
{
    if (groups <= 0)
        return hipSuccess;
    if (p.L >= 4)
        hipLaunchKernelGGL((bf_count_kernel<true>), dim3((unsigned)groups), dim3(kThreads), 0, s, p, first_group);
    else
        hipLaunchKernelGGL((bf_count_kernel<false>), dim3((unsigned)groups), dim3(kThreads), 0, s, p, first_group);
    return hipGetLastError();
}

// Workgroups of a persistent launch: the resident capacity of the device for the
// kernel (compute units x workgroups per unit), looked up once per device.
template <class K>
int64_t resident_grid(K kernel)
{
    static std::mutex mtx;
    static std::map<std::pair<int, const void *>, int64_t> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess)
        return 1024;
    std::lock_guard<std::mutex> lk(mtx);
    const auto key = std::make_pair(dev, (const void *)kernel);
    auto it = cache.find(key);
    if (it != cache.end())
        return it->second;
    int cus = 0, per = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, kThreads, 0) != hipSuccess || per <= 0)
        per = 1;
    (void)hipGetLastError();
    const int64_t grid = (int64_t)cus * per;
    cache[key] = grid;
    return grid;
}

// Output staging of the border-free kernels: 1 = an LDS image of each group's output
// stored by the whole workgroup (with register prefetch), 0 = every thread stores
// its own pieces (CR_BF_IMAGE, read per launch).
bool bf_image()
{
    const char *e = getenv("CR_BF_IMAGE");
    return e ? atoi(e) != 0 : true;
}

template <bool FULL4, bool IMAGE>
hipError_t sweep_t(const ReplaceParams &p, const SweepArgs &a, hipStream_t s)
{
    const int64_t grid = std::min(p.num_groups + 1, resident_grid(bf_sweep_kernel<FULL4, IMAGE>));
    hipLaunchKernelGGL((bf_sweep_kernel<FULL4, IMAGE>), dim3((unsigned)std::max<int64_t>(grid, 2)), dim3(kThreads),
                       0, s, p, a);
    return hipGetLastError();
}

template <bool FULL4>
hipError_t sweep_keep_t(const ReplaceParams &p, const SweepArgs &a, hipStream_t s)
{
    const int64_t grid = std::min(p.num_groups + 1, resident_grid(bf_sweep_keep_kernel<FULL4>));
    hipLaunchKernelGGL((bf_sweep_keep_kernel<FULL4>), dim3((unsigned)std::max<int64_t>(grid, 2)), dim3(kThreads),
                       0, s, p, a);
    return hipGetLastError();
}

// keep: the form that holds each group in LDS from its count to its emit (it needs
// the output image; without one the re-staging form runs)
hipError_t launch_bf_sweep(const ReplaceParams &p, uint64_t *sweep, int ahead, int spin_limit, bool keep,
                           hipStream_t s)
{
    const int64_t G = p.num_groups;
    SweepArgs a{sweep, sweep + G, sweep + 2 * G, sweep + 2 * G + 1, ahead, spin_limit};
    const bool img = bf_image();
    if (img && keep)
        return p.L >= 4 ? sweep_keep_t<true>(p, a, s) : sweep_keep_t<false>(p, a, s);
    if (p.L >= 4)
        return img ? sweep_t<true, true>(p, a, s) : sweep_t<true, false>(p, a, s);
    return img ? sweep_t<false, true>(p, a, s) : sweep_t<false, false>(p, a, s);
}

hipError_t launch_tbl_sweep(const ReplaceParams &p, uint64_t *sweep, int spin_limit, hipStream_t s)
{
    const int64_t G = p.num_groups;
    SweepArgs a{sweep, sweep + G, sweep + 2 * G, sweep + 2 * G + 1, 1, spin_limit};
    const int64_t grid = std::min(G + 1, resident_grid(tbl_sweep_keep_kernel));
    hipLaunchKernelGGL(tbl_sweep_keep_kernel, dim3((unsigned)std::max<int64_t>(grid, 2)), dim3(kThreads), 0, s, p, a);
    return hipGetLastError();
}

int64_t sweep_words(int64_t groups) { return 2 * groups + 2; }

hipError_t launch_bf_emit(const ReplaceParams &p, int64_t first_group, int64_t groups, hipStream_t s)
{
    if (groups <= 0)
        return hipSuccess;
    if (!bf_image()) {
        if (p.L >= 4)
            hipLaunchKernelGGL((bf_emit_direct_kernel<true>), dim3((unsigned)groups), dim3(kThreads), 0, s, p,
                               first_group);
        else
            hipLaunchKernelGGL((bf_emit_direct_kernel<false>), dim3((unsigned)groups), dim3(kThreads), 0, s, p,
                               first_group);
        return hipGetLastError();
    }
    if (p.L >= 4) {
        const int64_t grid = std::min(groups, resident_grid(bf_emit_kernel<true>));
        hipLaunchKernelGGL((bf_emit_kernel<true>), dim3((unsigned)grid), dim3(kThreads), 0, s, p, first_group,
                           groups);
    } else {
        const int64_t grid = std::min(groups, resident_grid(bf_emit_kernel<false>));
        hipLaunchKernelGGL((bf_emit_kernel<false>), dim3((unsigned)grid), dim3(kThreads), 0, s, p, first_group,
                           groups);
    }
    return hipGetLastError();
}

hipError_t launch_bitmap_mark(const ReplaceParams &p, hipStream_t s)
{
    const int64_t words = p.num_tiles * (kTile / 32);
    const int64_t blocks = (words + kThreads - 1) / kThreads;
    hipLaunchKernelGGL(mark_bitmap_kernel, dim3((unsigned)blocks), dim3(kThreads), 0, s, p, words);
    return hipGetLastError();
}

hipError_t launch_bitmap_greedy(const ReplaceParams &p, hipStream_t s)
{
    const int64_t words = p.num_tiles * (kTile / 32);
    hipLaunchKernelGGL(greedy_bitmap_kernel, dim3(1), dim3(kWave), 0, s, p, words);
    return hipGetLastError();
}

namespace {
// Scratch layout of a parallel bitmap selection: level sizes (level 0 = tiles, each
// next level 1/64 of the previous, until at most kBmFan tables remain), the tables
// of every level, the entries of every level, and the border-free last candidates.
struct BmPlan {
    int states = 0;
    int levels = 0;              // tables levels 0..levels-1
    int64_t n[24] = {};
    size_t tab_off[24] = {};
    size_t ent_off[24] = {};
    size_t tlast_off = 0;
    size_t bytes = 0;
};

BmPlan bm_plan(int64_t tiles, int states)
{
    BmPlan pl;
    pl.states = states;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return o;
    };
    int64_t cnt = tiles;
    for (;;) {
        pl.n[pl.levels] = cnt;
        pl.tab_off[pl.levels] = take((size_t)cnt * states * (pl.levels == 0 ? 4 : 8));
        ++pl.levels;
        if (cnt <= kBmFan)
            break;
        cnt = (cnt + kBmFan - 1) / kBmFan;
    }
    for (int l = 0; l < pl.levels; ++l)
        pl.ent_off[l] = take((size_t)pl.n[l] * sizeof(ChunkState));
    pl.tlast_off = take((size_t)tiles * 4);
    pl.bytes = off;
    return pl;
}
}   // namespace

// Leaves per tile: 1 KiB leaves for patterns of up to 32 bytes (a leaf's walks are
// then 1/8 as long and a tile resolves eight leaves at once), else the tile.
int bm_leaves(int L, bool border_free)
{
    return !border_free && L <= 32 ? 8 : 1;
}

size_t bitmap_select_bytes(int64_t num_tiles, int L, bool border_free)
{
    if (!border_free && L > kMaxSinglePassPat)
        return 0;   // a self-overlapping pattern over 256 bytes keeps the serial greedy
    return bm_plan(num_tiles * bm_leaves(L, border_free), border_free ? 1 : L).bytes;
}

hipError_t launch_bitmap_select(const ReplaceParams &p, bool border_free, uint8_t *scratch, hipStream_t s)
{
    const int64_t tiles = p.num_tiles;
    const int lpt = bm_leaves(p.L, border_free);
    const BmPlan pl = bm_plan(tiles * lpt, border_free ? 1 : p.L);
    const int st = pl.states;
    auto t32 = [&](int l) { return l == 0 ? reinterpret_cast<uint32_t *>(scratch + pl.tab_off[0]) : nullptr; };
    auto t64 = [&](int l) { return l == 0 ? nullptr : reinterpret_cast<uint64_t *>(scratch + pl.tab_off[l]); };
    auto ent = [&](int l) { return reinterpret_cast<ChunkState *>(scratch + pl.ent_off[l]); };
    int32_t *tlast = border_free ? reinterpret_cast<int32_t *>(scratch + pl.tlast_off) : nullptr;
    // the mark of a pattern of up to 32 bytes takes the masked-SAD words (pattern_qsad),
    // whose first word is never zero once set
    if (p.L <= 32 && p.qref[0] == 0u)
        return hipErrorInvalidValue;
    if (p.L <= kHalo && !border_free) {
        // candidates and leaf tables in one pass over each tile
        hipLaunchKernelGGL(bm_mark_lds_kernel<true>, dim3((unsigned)tiles), dim3(kThreads), 0, s, p, t32(0), st, lpt);
    } else {
        if (p.L <= kHalo) {
            hipLaunchKernelGGL(bm_mark_lds_kernel<false>, dim3((unsigned)tiles), dim3(kThreads), 0, s, p, nullptr, 0,
                               0);
        } else {
            const hipError_t e = launch_bitmap_mark(p, s);
            if (e != hipSuccess)
                return e;
        }
        // one wave per tile when the states fit one (more tiles in flight per unit:
        // the walks are latency-bound); border-free tiles reduce with the full block
        const unsigned tab_threads = (!border_free && st * lpt <= kWave) ? (unsigned)kWave : (unsigned)kThreads;
        hipLaunchKernelGGL(bm_tables_kernel, dim3((unsigned)tiles), dim3(tab_threads), 0, s, p, t32(0), st, lpt,
                           tlast);
    }
    const unsigned comp_threads = (unsigned)(st >= kThreads ? kThreads : (st + kWave - 1) / kWave * kWave);
    const size_t stage = bm_stage_bytes(st);
    for (int l = 1; l < pl.levels; ++l)
        hipLaunchKernelGGL(bm_compose_kernel, dim3((unsigned)pl.n[l]), dim3(comp_threads), stage, s, t32(l - 1),
                           t64(l - 1), pl.n[l - 1], st, t64(l));
    const int top = pl.levels - 1;
    hipLaunchKernelGGL(bm_top_kernel, dim3(1), dim3(1), 0, s, t32(top), t64(top), pl.n[top], st, ent(top), p.stats,
                       p.epoch);
    for (int l = top; l >= 1; --l)
        hipLaunchKernelGGL(bm_expand_kernel, dim3((unsigned)pl.n[l]), dim3(stage ? kWave : 1), stage, s, t32(l - 1),
                           t64(l - 1), pl.n[l - 1], st, ent(l), ent(l - 1));
    if (border_free)
        hipLaunchKernelGGL(bm_bf_records_kernel, dim3((unsigned)((tiles + kThreads - 1) / kThreads)), dim3(kThreads),
                           0, s, p, ent(0), tlast);
    else
        hipLaunchKernelGGL(bm_resolve_kernel, dim3((unsigned)tiles), dim3(kWave), 0, s, p, ent(0), lpt);
    return hipGetLastError();
}

hipError_t launch_empty_pattern(const ReplaceParams &p, hipStream_t s)
{
    int64_t blocks = (p.n + 1 + kThreads - 1) / kThreads;
    if (blocks > 65536)
        blocks = 65536;
    hipLaunchKernelGGL(empty_pattern_kernel, dim3((unsigned)blocks), dim3(kThreads), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_fill_u64(uint64_t *ptr, uint64_t value, size_t count, hipStream_t s)
{
    if (count == 0)
        return hipSuccess;
    size_t blocks = (count + kThreads - 1) / kThreads;
    if (blocks > 4096)
        blocks = 4096;
    hipLaunchKernelGGL(fill_u64_kernel, dim3((unsigned)blocks), dim3(kThreads), 0, s, ptr, value, count);
    return hipGetLastError();
}

hipError_t launch_query(const ReplaceParams &p, int mode, int64_t limit, int64_t *out2, hipStream_t s)
{
    hipLaunchKernelGGL(query_kernel, dim3(1), dim3(kThreads), 0, s, p, mode, limit, out2);
    return hipGetLastError();
}

// *out1 = last kept start + 1 over the whole pass (0: none)
hipError_t launch_last_kept(const ReplaceParams &p, int64_t *out1, hipStream_t s)
{
    hipError_t e = launch_fill_u64(reinterpret_cast<uint64_t *>(out1), 0, 1, s);
    if (e != hipSuccess)
        return e;
    int64_t blocks = (p.num_tiles + kThreads - 1) / kThreads;
    if (blocks > 1024)
        blocks = 1024;
    hipLaunchKernelGGL(last_kept_kernel, dim3((unsigned)blocks), dim3(kThreads), 0, s, p.info, p.num_tiles,
                       reinterpret_cast<unsigned long long *>(out1));
    return hipGetLastError();
}

hipError_t launch_presence_index(const uint8_t *src, int64_t n, int64_t chunk, uint64_t *masks,
                                 int64_t num_chunks, hipStream_t s)
{
    if (num_chunks <= 0)
        return hipSuccess;
    const int64_t grid = num_chunks < (1 << 20) ? num_chunks : (1 << 20);
    hipLaunchKernelGGL(presence_kernel, dim3((unsigned)grid), dim3(kThreads), 0, s, src, n, chunk, masks,
                       num_chunks);
    return hipGetLastError();
}

// ---- batch launchers ----------------------------------------------------------
hipError_t launch_batch_count(const BatchArgs &a, int mode, int table_size, int gt, int64_t groups, hipStream_t s)
{
    if (groups <= 0)
        return hipSuccess;
    return dispatch_mode(mode, table_size, gt, [&]<int M, int LT, int GT>() {
        hipLaunchKernelGGL((batch_count_kernel<M, LT, GT>), dim3((unsigned)groups), dim3(kThreads), 0, s, a);
        return hipGetLastError();
    });
}

hipError_t launch_batch_spine(const BatchArgs &a, int mode, int table_size, int64_t docs, hipStream_t s)
{
    if (docs <= 0)
        return hipSuccess;
    // the spine reads group summaries only, so one instantiation serves every group size
    return dispatch_mode(mode, table_size, kGroupTiles, [&]<int M, int LT, int GT>() {
        hipLaunchKernelGGL((batch_spine_kernel<M, LT>), dim3((unsigned)docs), dim3(kWave), 0, s, a);
        return hipGetLastError();
    });
}

hipError_t launch_batch_emit(const BatchArgs &a, int mode, int table_size, int gt, int64_t groups, hipStream_t s)
{
    if (groups <= 0)
        return hipSuccess;
    return dispatch_mode(mode, table_size, gt, [&]<int M, int LT, int GT>() {
        hipLaunchKernelGGL((batch_emit_kernel<M, LT, GT>), dim3((unsigned)groups), dim3(kThreads), 0, s, a);
        return hipGetLastError();
    });
}

template <int M, int LT, int GT, bool EMIT>
hipError_t launch_work(const BatchRound &a, int c, const BatchRoundShape &shape, hipStream_t s)
{
    const int64_t est = shape.work_grid[c] > 0 ? shape.work_grid[c] : 1;
    hipLaunchKernelGGL((batch_item_kernel<M, LT, GT, EMIT>), dim3((unsigned)est), dim3(kThreads), 0, s, a, c);
    const unsigned rest = (unsigned)(shape.cus > 0 ? shape.cus : 1);
    hipLaunchKernelGGL((batch_rest_kernel<M, LT, GT, EMIT>), dim3(rest), dim3(kThreads), 0, s, a, c, est);
    return hipGetLastError();
}

hipError_t launch_batch_round(const BatchRound &a, const BatchRoundShape &shape, hipStream_t s)
{
    static const int kClassMode[kBatchClasses] = {kModeBorderFree, kModeTable, kModeTable, kModeTable, kModeTable};
    static const int kClassLt[kBatchClasses] = {1, 2, 4, 8, 16};
    hipLaunchKernelGGL(batch_plan_kernel, dim3(1), dim3(kPlanThreads), 0, s, a);
    int counted = 0;
    for (int c = 0; c < kBatchClasses; ++c)
        counted += shape.present[c] ? shape.class_docs[c] : 0;
    if (counted > 0) {
        const int eg = counted < 1024 ? counted : 1024;
        hipLaunchKernelGGL(batch_expand_kernel, dim3((unsigned)eg), dim3(kThreads), 0, s, a);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return e;
    for (int c = 0; c < kBatchClasses && e == hipSuccess; ++c) {
        if (!shape.present[c])
            continue;
        e = dispatch_mode(kClassMode[c], kClassLt[c], a.gt, [&]<int M, int LT, int GT>() {
            return launch_work<M, LT, GT, false>(a, c, shape, s);
        });
        if (e != hipSuccess)
            return e;
        const int nd = shape.class_docs[c];
        const unsigned sg = (unsigned)(nd < 2048 ? nd : 2048);
        e = dispatch_mode(kClassMode[c], kClassLt[c], kGroupTiles, [&]<int M, int LT, int GT>() {
            hipLaunchKernelGGL((batch_round_spine_kernel<M, LT>), dim3(sg), dim3(kWave), 0, s, a, c);
            return hipGetLastError();
        });
    }
    if (e != hipSuccess)
        return e;
    hipLaunchKernelGGL(batch_place_kernel, dim3(1), dim3(kPlanThreads), 0, s, a);
    for (int c = 0; c < kBatchClasses && e == hipSuccess; ++c) {
        if (!shape.present[c])
            continue;
        e = dispatch_mode(kClassMode[c], kClassLt[c], a.gt, [&]<int M, int LT, int GT>() {
            return launch_work<M, LT, GT, true>(a, c, shape, s);
        });
    }
    if (e != hipSuccess)
        return e;
    if (shape.present[kBatchEmpty]) {
        const int nd = shape.class_docs[kBatchEmpty];
        hipLaunchKernelGGL(batch_empty_kernel, dim3((unsigned)(nd < 1024 ? nd : 1024)), dim3(kThreads), 0, s, a);
    }
    return hipGetLastError();
}

hipError_t launch_fused_sweep(const ReplaceParams &p, const FusedTable *t, uint64_t *sweep, int spin_limit,
                              hipStream_t s)
{
    const int64_t G = p.num_groups;
    SweepArgs a{sweep, sweep + G, sweep + 2 * G, sweep + 2 * G + 1, 1, spin_limit};
    const int64_t grid = std::min(G + 1, resident_grid(fused_sweep_keep_kernel));
    hipLaunchKernelGGL(fused_sweep_keep_kernel, dim3((unsigned)std::max<int64_t>(grid, 2)), dim3(kThreads), 0, s, p,
                       t, a);
    return hipGetLastError();
}

hipError_t launch_fused(const ReplaceParams &p, const FusedTable *t, int which, int64_t first_group, int64_t groups,
                        hipStream_t s)
{
    if (groups <= 0)
        return hipSuccess;
    if (which == 0)
        hipLaunchKernelGGL(fused_count_kernel, dim3((unsigned)groups), dim3(kThreads), 0, s, p, t, first_group);
    else if (which == 1)
        return launch_spine_blocks(p, kModeFused, 1, groups, s);
    else
        hipLaunchKernelGGL(fused_emit_kernel, dim3((unsigned)groups), dim3(kThreads), 0, s, p, t, first_group);
    return hipGetLastError();
}

} // namespace cr
