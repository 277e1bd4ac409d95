// cr_kernels.hip -- HIP kernels for leftmost non-overlapping byte replace on gfx950.
//
// One workgroup (256 threads, 4 wave64) processes one 8 KiB tile of the source in a
// single pass: stage tile + halo in LDS, build 32-bit candidate masks per thread,
// resolve the leftmost-greedy suppression, obtain the tile's output offset with a
// decoupled look-back over earlier tiles, then assemble the tile's contiguous output
// in LDS and write it with aligned 16-byte stores.
//
// Replaces the reference's mark / suppress / scan / scatter chain
// (cuda_replace.cu:737-866 fused mark+suppress, 237-334 three-stage scan,
// 1391-1608 shrink-pad / expansion scatter, 2097-2159 apply_once) with one kernel.
//
// Suppression modes:
//   border-free pattern (no proper prefix equals a suffix): occurrences can never
//     overlap, so every candidate at or above the seed floor is kept; the carry into a
//     tile is recomputed from the L-1 bytes left of it; look-back sums kept counts.
//   self-overlapping pattern, L <= 16: each thread summarises its 32 bytes as a table
//     carry_in -> (carry_out, kept) over the L carry states; tables compose
//     associatively (wave64 shuffle scan, then look-back over tile tables).
//   otherwise: candidate bitmap, one-wave serial greedy, then the same scatter.
#include "cr_internal.h"

#include <limits.h>
#include <stdio.h>
#include <stdlib.h>

namespace cr {

constexpr int kWave = 64;
constexpr int kWaves = kThreads / kWave;
constexpr int kNegInf = INT_MIN / 4;
// Bound on look-back polls of one tile (legitimate waits are a few microseconds).
constexpr int kSpinLimit = 1 << 18;
// Replacements up to this length are staged in LDS; longer ones are read from HBM.
constexpr int kRepLds = 256;

// A look-back that waits too long raises the launch's abort flag; every other
// waiter sees it and leaves, and the host re-runs the pass on the bitmap path.
__device__ __forceinline__ bool aborted(const ReplaceParams &p)
{
    const uint64_t w = __hip_atomic_load(&p.stats->pad[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return (uint32_t)(w >> kEpochShift) == p.epoch;
}
__device__ __forceinline__ void raise_abort(const ReplaceParams &p)
{
    __hip_atomic_store(&p.stats->pad[1], ((uint64_t)p.epoch << kEpochShift) | 1u, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}

// ---- look-back words --------------------------------------------------------
__device__ __forceinline__ uint64_t make_word(uint32_t epoch, uint64_t status, uint64_t payload)
{
    return ((uint64_t)epoch << kEpochShift) | (status << kStatusShift) | (payload & kPayloadMask);
}
__device__ __forceinline__ uint64_t load_word(const uint64_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_word(uint64_t *p, uint64_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t word_status(uint64_t w, uint32_t epoch)
{
    return ((uint32_t)(w >> kEpochShift) == epoch) ? (uint32_t)((w >> kStatusShift) & 3u) : 0u;
}

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

// ---- carry tables (self-overlapping patterns) -------------------------------
// entry = carry_out << 24 | kept_count ; carry < 16, count <= 8192 inside a tile
template <int LT>
struct Table {
    uint32_t e[LT];
};

template <int LT>
__device__ __forceinline__ uint32_t tsel(const Table<LT> &t, uint32_t idx)
{
    uint32_t v = t.e[0];
#pragma unroll
    for (int i = 1; i < LT; ++i)
        v = (idx == (uint32_t)i) ? t.e[i] : v;
    return v;
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

// ---- tile kernel ------------------------------------------------------------
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

// Copies len bytes inside LDS without loop-carried LDS round trips: up to 3
// predicated head bytes until the destination is 4-byte aligned, then blocks of up
// to 8 aligned dword writes whose source dwords (re-aligned with alignbyte) are
// all read first, then up to 3 predicated tail bytes.  Reads at most 4 bytes past
// the source range.
__device__ __forceinline__ void lds_copy(uint8_t *D, int dst, const uint8_t *S, int src, int len)
{
    const int h0 = (-dst) & 3;
    const int h = h0 < len ? h0 : len;
    if (h > 0)
        D[dst] = S[src];
    if (h > 1)
        D[dst + 1] = S[src + 1];
    if (h > 2)
        D[dst + 2] = S[src + 2];
    dst += h;
    src += h;
    len -= h;
    const uint32_t sh = (uint32_t)(src & 3);
    const uint32_t *sp = reinterpret_cast<const uint32_t *>(S + (src & ~3));
    uint32_t *dp = reinterpret_cast<uint32_t *>(D + dst);
    int nd = len >> 2;
    while (nd > 0) {
        uint32_t v[9];
#pragma unroll
        for (int k = 0; k < 9; ++k)
            v[k] = k <= nd ? sp[k] : 0u;
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (k < nd)
                dp[k] = __builtin_amdgcn_alignbyte(v[k + 1], v[k], sh);
        const int blk = nd < 8 ? nd : 8;
        sp += blk;
        dp += blk;
        nd -= blk;
    }
    const int t = len & 3;
    const int ts = src + (len & ~3), td = dst + (len & ~3);
    if (t > 0)
        D[td] = S[ts];
    if (t > 1)
        D[td + 1] = S[ts + 1];
    if (t > 2)
        D[td + 2] = S[ts + 2];
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

// Sum over the wave of the published words at addr[0..S) of every lane (null
// addresses contribute 0), waiting until every one is published.  All loads of a
// round are in flight together.  Returns false when the wait gave up (abort).
// Second half of a poll whose words w were already loaded from addr.
template <int S>
__device__ __forceinline__ bool finish_poll(const ReplaceParams &p, const uint64_t *const (&addr)[S], uint64_t (&w)[S],
                                            int64_t *sum)
{
    int spins = 0;
    while (true) {
        bool nr = false;
#pragma unroll
        for (int q = 0; q < S; ++q)
            nr |= word_status(w[q], p.epoch) == 0;
        if (!__ballot(nr))
            break;
        if (++spins > kSpinLimit || aborted(p))
            return false;
        __builtin_amdgcn_s_sleep(2);
#pragma unroll
        for (int q = 0; q < S; ++q)
            if (word_status(w[q], p.epoch) == 0)
                w[q] = load_word(addr[q]);
    }
    int64_t v = 0;
#pragma unroll
    for (int q = 0; q < S; ++q)
        v += (int64_t)(w[q] & kPayloadMask);
#pragma unroll
    for (int dd = kWave / 2; dd > 0; dd >>= 1)
        v += __shfl_xor(v, dd, kWave);
    *sum = v;
    return true;
}

template <int S>
__device__ __forceinline__ bool poll_sum(const ReplaceParams &p, const uint64_t *const (&addr)[S], int64_t *sum)
{
    uint64_t w[S];
#pragma unroll
    for (int q = 0; q < S; ++q)
        w[q] = addr[q] ? load_word(addr[q]) : make_word(p.epoch, kStatusAgg, 0);
    return finish_poll<S>(p, addr, w, sum);
}


// 64-bit carry tables for the prefix hierarchy: entry = carry_out << 40 | count.
template <int LT>
struct Table64 {
    uint64_t e[LT];
};

template <int LT>
__device__ __forceinline__ uint64_t tsel64(const Table64<LT> &t, uint32_t idx)
{
    uint64_t v = t.e[0];
#pragma unroll
    for (int i = 1; i < LT; ++i)
        v = (idx == (uint32_t)i) ? t.e[i] : v;
    return v;
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
__device__ __forceinline__ Table64<LT> wave_broadcast0(const Table64<LT> &t)
{
    Table64<LT> r;
#pragma unroll
    for (int c = 0; c < LT; ++c)
        r.e[c] = __shfl(t.e[c], 0, kWave);
    return r;
}

// Loads the published LT-word table at addr (identity when addr is null), waiting
// until every word is published.  Returns false when the wait gave up.
template <int LT>
__device__ __forceinline__ bool poll_table(const ReplaceParams &p, const uint64_t *addr, Table64<LT> &t)
{
    uint64_t w[LT];
#pragma unroll
    for (int c = 0; c < LT; ++c)
        w[c] = addr ? load_word(addr + c) : make_word(p.epoch, kStatusAgg, (uint64_t)c << kCarryShift);
    int spins = 0;
    while (true) {
        bool nr = false;
#pragma unroll
        for (int c = 0; c < LT; ++c)
            nr |= word_status(w[c], p.epoch) == 0;
        if (!__ballot(nr))
            break;
        if (++spins > kSpinLimit || aborted(p))
            return false;
        __builtin_amdgcn_s_sleep(2);
#pragma unroll
        for (int c = 0; c < LT; ++c)
            if (word_status(w[c], p.epoch) == 0)
                w[c] = load_word(addr + c);
    }
#pragma unroll
    for (int c = 0; c < LT; ++c)
        t.e[c] = w[c] & kPayloadMask;
    return true;
}

// ---- persistent tile kernel -------------------------------------------------
// The next tile is fetched straight into LDS with global_load_lds (no registers):
// waves 2-3 load the 8 KiB body (1 KiB per instruction), wave 1 the two 256-byte
// halos (4 bytes per lane).  Wave 0 never issues these loads, so its look-back
// polls never wait behind them.
__device__ __forceinline__ void prefetch_lds(const ReplaceParams &p, int64_t tile, int tid, bool halos,
                                             uint8_t *s_in)
{
    typedef __attribute__((address_space(3))) void lds_void;
    const int64_t start = tile * kTile;
    const int wave = tid / kWave;
    const int lane = tid & (kWave - 1);
    if (wave >= 2) {
        // The source is a 16-byte aligned device allocation (16-byte granular), so
        // the granule holding the last valid byte is readable; bytes past n are
        // never matched or emitted.
        const int64_t avail = p.n - start;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int off = (wave - 2) * 1024 + k * 2048;
            if (off + lane * 16 < avail)
                __builtin_amdgcn_global_load_lds((const void *)(p.src + start + off + lane * 16),
                                                 (lds_void *)(s_in + kHalo + off), 16, 0, 0);
        }
    } else if (wave == 1 && halos) {
        // left halo [start-256, start), right halo [start+kTile, start+kTile+256)
        if (start >= kHalo)
            __builtin_amdgcn_global_load_lds((const void *)(p.src + start - kHalo + lane * 4),
                                             (lds_void *)(s_in), 4, 0, 0);
        if (start + kTile + lane * 4 < p.n)
            __builtin_amdgcn_global_load_lds((const void *)(p.src + start + kTile + lane * 4),
                                             (lds_void *)(s_in + kHalo + kTile), 4, 0, 0);
    }
}

// Persistent: gridDim.x workgroups, all co-resident (cooperative launch), walk the
// tiles blockIdx.x, blockIdx.x + gridDim.x, ...  Every tile a look-back waits on
// belongs to a resident workgroup and an earlier or the same round, so the waits
// always end; the bounded spin only turns an unexpected stall into an abort that
// the host answers by re-running the pass on the bitmap path.
template <int MODE, int LT>
__global__ void __launch_bounds__(kThreads)
tile_kernel(ReplaceParams p)
{
    __shared__ __align__(16) uint8_t s_in[kHalo + kTile + kHalo];
    __shared__ __align__(16) uint8_t s_out[kStage];
    __shared__ uint8_t s_pat[kMaxSinglePassPat];
    __shared__ __align__(16) uint8_t s_rep[kRepLds + 32];
    __shared__ int s_wsum[kWaves];
    __shared__ int s_wmax[kWaves];
    __shared__ int s_halo_last;
    __shared__ int s_abort;
    __shared__ int64_t s_kexcl;
    __shared__ int s_carry_in;
    __shared__ uint32_t s_wtab[kWaves][LT > 1 ? LT : 1];

    const int L = p.L;
    const int r = p.r;
    const int64_t n = p.n;
    const int64_t nt = p.num_tiles;
    const int64_t G = gridDim.x;
    const int64_t diff = (int64_t)r - (int64_t)L;
    const bool need_src = MODE != kModeBitmap || p.write_output;
    const bool halos = MODE != kModeBitmap;
    int tid = threadIdx.x;

    int64_t tile = (int64_t)blockIdx.x;
    if (tile >= nt)
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
        s_abort = 0;
    if (need_src)
        prefetch_lds(p, tile, tid, halos, s_in);

    for (; tile < nt; tile += G) {
        // Re-derive the thread index each round through an opaque register move:
        // keeps lane-dependent addresses from being hoisted out of the loop (they
        // would pin dozens of VGPRs for the whole kernel).
        asm volatile("" : "+v"(tid));
        const int lane = tid & (kWave - 1);
        const int wave = tid / kWave;
        const int seg = tid * kBytesPerThread;     // tile-relative start of my segment
        const int64_t start = tile * kTile;
        const int valid = (int)((n - start) < kTile ? (n - start) : kTile);
        const bool last_tile = tile == nt - 1;
        const int seg_valid = valid - seg;
        const int jend = seg_valid < 32 ? (seg_valid > 0 ? seg_valid : 0) : 32;

        if (wave >= 1)
            __builtin_amdgcn_s_waitcnt(0);         // this wave's prefetch has landed
        if (tid == 0)
            s_halo_last = kNegInf;
        uint32_t m = 0;   // candidate mask, then kept mask, of this thread's 32 positions
        if constexpr (MODE == kModeBitmap)
            m = p.bitmap[tile * (kTile / 32) + tid];
        __syncthreads();

        // ---- candidate mask ----------------------------------------------------
        if constexpr (MODE != kModeBitmap) {
            const int64_t gseg = start + seg;
            const uint32_t *w32 = reinterpret_cast<const uint32_t *>(s_in + kHalo + seg);
            uint32_t d[9];
#pragma unroll
            for (int k = 0; k < 9; ++k)
                d[k] = w32[k];
            const uint32_t pat4 = p.pat4, mask4 = p.mask4;
#pragma unroll
            for (int j = 31; j >= 0; --j) {
                const uint32_t w = __builtin_amdgcn_alignbyte(d[j / 4 + 1], d[j / 4], j % 4);
                m = (m << 1) | (uint32_t)(((w ^ pat4) & mask4) == 0u);
            }
            // bounds: floor <= pos <= n - L
            const int64_t lo = p.floor - gseg;
            const int64_t hi = n - (int64_t)L - gseg;
            uint32_t rm = 0;
            if (hi >= 0 && lo <= 31 && lo <= hi) {
                const int a = lo < 0 ? 0 : (int)lo;
                const int b = hi > 31 ? 31 : (int)hi;
                rm = (b == 31 ? 0xffffffffu : ((1u << (b + 1)) - 1u)) & (0xffffffffu << a);
            }
            m &= rm;
            if (L > 4 && m) {
                uint32_t mm = m;
                m = 0;
                const uint8_t *sp = s_in + kHalo + seg;
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
        }

        // ---- border-free: carry from candidates in the L-1 bytes left of the tile
        if constexpr (MODE == kModeBorderFree) {
            if (tile > 0 && tid < L - 1) {
                const int q = -1 - tid;                      // tile-relative position
                const int64_t g = start + q;
                if (g >= p.floor && g + L <= n) {
                    const uint8_t *sp = s_in + kHalo + q;
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
        }

        int c_i = 0;        // leading bytes of my segment covered by an earlier kept match
        int k_loc = 0;      // kept starts in the tile before my segment
        int tile_k = 0;     // kept starts in the tile
        int tile_last = kNegInf;   // last kept start in the tile (tile-relative)
        int c_t = 0;        // carry into the tile
        int c_next = 0;     // carry into the next tile
        int64_t k_excl = 0; // kept starts before the tile

        if constexpr (MODE == kModeBorderFree || MODE == kModeBitmap) {
            // kept mask is final: scan counts (sum) and last kept (max)
            const int k = __builtin_popcount(m);
            const int lastq = m ? (seg + 31 - __builtin_clz(m)) : kNegInf;
            int ks = k, lq = lastq;
#pragma unroll
            for (int dd = 1; dd < kWave; dd <<= 1) {
                const int uk = __shfl_up(ks, dd, kWave);
                const int ul = __shfl_up(lq, dd, kWave);
                if (lane >= dd) {
                    ks += uk;
                    lq = ul > lq ? ul : lq;
                }
            }
            int lq_ex = __shfl_up(lq, 1, kWave);
            if (lane == 0)
                lq_ex = kNegInf;
            if (lane == kWave - 1) {
                s_wsum[wave] = ks;
                s_wmax[wave] = lq;
            }
            __syncthreads();
            int init_last;
            if constexpr (MODE == kModeBorderFree) {
                init_last = s_halo_last;
                c_t = init_last + L > 0 ? init_last + L : 0;
            } else {
                c_t = p.info[tile].carry_in;
                k_excl = p.info[tile].k_excl;
                init_last = c_t - L;   // a virtual kept start covering [.., c_t)
            }
            int wsum = 0, wmax = init_last;
            for (int w = 0; w < kWaves; ++w) {
                if (w < wave) {
                    wsum += s_wsum[w];
                    wmax = s_wmax[w] > wmax ? s_wmax[w] : wmax;
                }
                tile_k += s_wsum[w];
                tile_last = s_wmax[w] > tile_last ? s_wmax[w] : tile_last;
            }
            k_loc = wsum + ks - k;
            const int prev_last = lq_ex > wmax ? lq_ex : wmax;
            c_i = prev_last + L - seg;
            if (c_i < 0)
                c_i = 0;
            if constexpr (MODE == kModeBorderFree) {
                c_next = tile_last + L - kTile;
                if (c_next < 0 || last_tile)
                    c_next = 0;
            } else {
                c_next = last_tile ? 0 : p.info[tile + 1].carry_in;
            }
        }

        // Writes this thread's output (source runs between kept starts, the
        // replacement at each kept start) into the tile-relative LDS stage.
        auto stage_output = [&](uint32_t kept_mask) {
            int o = (seg + c_i - c_t) + k_loc * (int)diff;
            int j = c_i;
            uint32_t M = j >= 32 ? 0u : (kept_mask & (0xffffffffu << j));
            while (j < jend) {
                const int q = M ? __builtin_ctz(M) : jend;
                if (q > j) {
                    lds_copy(s_out, o, s_in, kHalo + seg + j, q - j);
                    o += q - j;
                }
                if (q >= jend)
                    break;
                lds_copy(s_out, o, s_rep, 0, r);
                o += r;
                j = q + L;
                M = j >= 32 ? 0u : (M & (0xffffffffu << j));
            }
        };
        // tile output length (before the look-back for border-free and bitmap)
        int64_t out_len = (int64_t)valid + c_next - c_t + (int64_t)tile_k * diff;
        bool staged = p.write_output && r <= kRepLds && out_len + 32 <= kStage;

        uint32_t kept = m;
        if constexpr (MODE == kModeBorderFree) {
            // Hierarchical prefix of kept counts, fan-out 64: every tile publishes its
            // count, the last tile of each 64-tile group also publishes the group sum,
            // and likewise for groups of groups.  A tile's exclusive prefix is the sum
            // of at most 63 words per level, all published by tiles of this or an
            // earlier round shortly after their scans: no chain of inclusive prefixes.
            uint64_t *const a0 = p.state;
            uint64_t *const a1 = a0 + nt;
            uint64_t *const a2 = a1 + (nt >> 6) + 1;
            uint64_t *const a3 = a2 + (nt >> 12) + 1;
            const int64_t d0 = tile & 63, d1 = (tile >> 6) & 63, d2 = (tile >> 12) & 63, d3 = tile >> 18;
            const int64_t items = d0 + d1 + d2 + d3;
            // prefix item it -> its published word: a0 words of this group, a1 of this
            // 64-group block, a2 of this 4096-group block, a3 of all earlier blocks
            auto item_addr = [&](int64_t it) -> const uint64_t * {
                if (it < d0)
                    return &a0[tile - d0 + it];
                if (it < d0 + d1)
                    return &a1[(tile >> 6) - d1 + (it - d0)];
                if (it < d0 + d1 + d2)
                    return &a2[(tile >> 12) - d2 + (it - d0 - d1)];
                if (it < items)
                    return &a3[it - d0 - d1 - d2];
                return nullptr;
            };
            bool ok = true;
            uint64_t pw[4];
            if (wave == 0) {
                if (lane == 0)
                    store_word(&a0[tile], make_word(p.epoch, kStatusAgg, (uint64_t)tile_k));
                if ((tile & 63) == 63) {
                    int64_t s1 = 0;
                    const uint64_t *g0[1] = {lane < 63 ? &a0[tile - 63 + lane] : nullptr};
                    ok = poll_sum<1>(p, g0, &s1);
                    s1 += tile_k;
                    if (ok && lane == 0)
                        store_word(&a1[tile >> 6], make_word(p.epoch, kStatusAgg, (uint64_t)s1));
                    if (ok && ((tile >> 6) & 63) == 63) {
                        const int64_t g = tile >> 6;
                        int64_t s2 = 0;
                        const uint64_t *g1[1] = {lane < 63 ? &a1[g - 63 + lane] : nullptr};
                        ok = poll_sum<1>(p, g1, &s2);
                        s2 += s1;
                        if (ok && lane == 0)
                            store_word(&a2[tile >> 12], make_word(p.epoch, kStatusAgg, (uint64_t)s2));
                        if (ok && ((tile >> 12) & 63) == 63) {
                            const int64_t h = tile >> 12;
                            int64_t s3 = 0;
                            const uint64_t *g2[1] = {lane < 63 ? &a2[h - 63 + lane] : nullptr};
                            ok = poll_sum<1>(p, g2, &s3);
                            s3 += s2;
                            if (ok && lane == 0)
                                store_word(&a3[tile >> 18], make_word(p.epoch, kStatusAgg, (uint64_t)s3));
                        }
                    }
                }
                // first poll round in flight while the output is staged
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const uint64_t *a = item_addr(q * kWave + lane);
                    pw[q] = a ? load_word(a) : make_word(p.epoch, kStatusAgg, 0);
                }
            }
            // build the output while the predecessors' words are on their way
            if (staged)
                stage_output(m);
            if (wave == 0) {
                int64_t acc = 0;
                if (ok) {
                    const uint64_t *pa[4];
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        pa[q] = item_addr(q * kWave + lane);
                    int64_t sum = 0;
                    ok = finish_poll<4>(p, pa, pw, &sum);
                    acc += sum;
                }
                for (int64_t base = 4 * kWave; ok && base < items; base += 4 * kWave) {
                    const uint64_t *addr[4];
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        addr[q] = item_addr(base + q * kWave + lane);
                    int64_t sum = 0;
                    ok = poll_sum<4>(p, addr, &sum);
                    acc += sum;
                }
                if (lane == 0) {
                    if (!ok) {
                        raise_abort(p);
                        s_abort = 1;
                    }
                    s_kexcl = acc;
                }
            }
            __syncthreads();
            if (s_abort)
                return;
            k_excl = s_kexcl;
        } else if constexpr (MODE == kModeBitmap) {
            if (staged)
                stage_output(m);
            __syncthreads();
        } else {
            // per-thread table over carry states c in [0, LT)
            Table<LT> t;
#pragma unroll
            for (int c = 0; c < LT; ++c) {
                int lst;
                const uint32_t km = greedy32(m, c, L, &lst);
                int co = lst + L - 32;
                co = co < 0 ? 0 : co;
                t.e[c] = ((uint32_t)co << 24) | (uint32_t)__builtin_popcount(km);
            }
            // wave-inclusive scan of tables
            Table<LT> inc = t;
#pragma unroll
            for (int dd = 1; dd < kWave; dd <<= 1) {
                Table<LT> nb;
#pragma unroll
                for (int c = 0; c < LT; ++c)
                    nb.e[c] = __shfl_up(inc.e[c], dd, kWave);
                if (lane >= dd)
                    inc = tcompose<LT>(nb, inc);
            }
            Table<LT> ex;
#pragma unroll
            for (int c = 0; c < LT; ++c) {
                const uint32_t v = __shfl_up(inc.e[c], 1, kWave);
                ex.e[c] = lane == 0 ? ((uint32_t)c << 24) : v;
            }
            if (lane == kWave - 1) {
#pragma unroll
                for (int c = 0; c < LT; ++c)
                    s_wtab[wave][c] = inc.e[c];
            }
            __syncthreads();
            Table<LT> wpre = tidentity<LT>();
            Table<LT> agg = tidentity<LT>();
            for (int w = 0; w < kWaves; ++w) {
                Table<LT> wt;
#pragma unroll
                for (int c = 0; c < LT; ++c)
                    wt.e[c] = s_wtab[w][c];
                if (w < wave)
                    wpre = tcompose<LT>(wpre, wt);
                agg = tcompose<LT>(agg, wt);
            }
            const Table<LT> thread_ex = tcompose<LT>(wpre, ex);

            // hierarchical prefix of carry tables (same fan-out-64 layout as the
            // border-free counts; entries composed in position order)
            if (wave == 0) {
                uint64_t *const t0 = p.state;
                uint64_t *const t1 = t0 + nt * LT;
                uint64_t *const t2 = t1 + ((nt >> 6) + 1) * LT;
                uint64_t *const t3 = t2 + ((nt >> 12) + 1) * LT;
                const Table64<LT> mine = widen<LT>(agg);
                bool ok = true;
#pragma unroll
                for (int c = 0; c < LT; ++c)
                    if (lane == c)
                        store_word(&t0[tile * LT + c], make_word(p.epoch, kStatusAgg, mine.e[c]));
                if ((tile & 63) == 63) {
                    Table64<LT> gt;
                    ok = poll_table<LT>(p, lane < 63 ? &t0[(tile - 63 + lane) * LT] : nullptr, gt);
                    if (lane == 63)
                        gt = mine;
                    gt = wave_broadcast0<LT>(wave_compose<LT>(gt, lane));
#pragma unroll
                    for (int c = 0; c < LT; ++c)
                        if (ok && lane == c)
                            store_word(&t1[(tile >> 6) * LT + c], make_word(p.epoch, kStatusAgg, gt.e[c]));
                    if (ok && ((tile >> 6) & 63) == 63) {
                        const int64_t g = tile >> 6;
                        Table64<LT> ht;
                        ok = poll_table<LT>(p, lane < 63 ? &t1[(g - 63 + lane) * LT] : nullptr, ht);
                        if (lane == 63)
                            ht = gt;
                        ht = wave_broadcast0<LT>(wave_compose<LT>(ht, lane));
#pragma unroll
                        for (int c = 0; c < LT; ++c)
                            if (ok && lane == c)
                                store_word(&t2[(tile >> 12) * LT + c], make_word(p.epoch, kStatusAgg, ht.e[c]));
                        if (ok && ((tile >> 12) & 63) == 63) {
                            const int64_t h = tile >> 12;
                            Table64<LT> kt;
                            ok = poll_table<LT>(p, lane < 63 ? &t2[(h - 63 + lane) * LT] : nullptr, kt);
                            if (lane == 63)
                                kt = ht;
                            kt = wave_broadcast0<LT>(wave_compose<LT>(kt, lane));
#pragma unroll
                            for (int c = 0; c < LT; ++c)
                                if (ok && lane == c)
                                    store_word(&t3[(tile >> 18) * LT + c],
                                               make_word(p.epoch, kStatusAgg, kt.e[c]));
                        }
                    }
                }
                // prefix items in position order: t3 blocks, t2, t1, t0
                const int64_t d0 = tile & 63, d1 = (tile >> 6) & 63, d2 = (tile >> 12) & 63, d3 = tile >> 18;
                const int64_t items = d0 + d1 + d2 + d3;
                Table64<LT> pre = tidentity64<LT>();
                for (int64_t base = 0; ok && base < items; base += kWave) {
                    const int64_t it = base + lane;
                    const uint64_t *addr = nullptr;
                    if (it < d3)
                        addr = &t3[it * LT];
                    else if (it < d3 + d2)
                        addr = &t2[((tile >> 12) - d2 + (it - d3)) * LT];
                    else if (it < d3 + d2 + d1)
                        addr = &t1[((tile >> 6) - d1 + (it - d3 - d2)) * LT];
                    else if (it < items)
                        addr = &t0[(tile - d0 + (it - d3 - d2 - d1)) * LT];
                    Table64<LT> tt;
                    ok = poll_table<LT>(p, addr, tt);
                    pre = tcompose64<LT>(pre, wave_broadcast0<LT>(wave_compose<LT>(tt, lane)));
                }
                const uint64_t e0 = pre.e[0];   // carry 0 into tile 0
                if (lane == 0) {
                    if (!ok) {
                        raise_abort(p);
                        s_abort = 1;
                    }
                    s_kexcl = (int64_t)(e0 & kCountMask);
                    s_carry_in = (int)(e0 >> kCarryShift);
                }
            }
            __syncthreads();
            if (s_abort)
                return;
            k_excl = s_kexcl;
            c_t = s_carry_in;
            const uint32_t me = tsel<LT>(thread_ex, (uint32_t)c_t);
            c_i = (int)(me >> 24);
            k_loc = (int)(me & 0xffffffu);
            c_next = last_tile ? 0 : (int)(tsel<LT>(agg, (uint32_t)c_t) >> 24);
            int lst;
            kept = greedy32(m, c_i, L, &lst);
            const int lastq = kept ? seg + lst : kNegInf;
            // tile totals: kept count and last kept
            int ks = __builtin_popcount(kept), lq = lastq;
#pragma unroll
            for (int dd = kWave / 2; dd > 0; dd >>= 1) {
                ks += __shfl_xor(ks, dd, kWave);
                const int o = __shfl_xor(lq, dd, kWave);
                lq = o > lq ? o : lq;
            }
            if (lane == 0) {
                s_wsum[wave] = ks;
                s_wmax[wave] = lq;
            }
            __syncthreads();
            for (int w = 0; w < kWaves; ++w) {
                tile_k += s_wsum[w];
                tile_last = s_wmax[w] > tile_last ? s_wmax[w] : tile_last;
            }
            out_len = (int64_t)valid + c_next - c_t + (int64_t)tile_k * diff;
            staged = p.write_output && r <= kRepLds && out_len + 32 <= kStage;
            if (staged)
                stage_output(kept);
            __syncthreads();
        }

        // ---- records for prefix queries and launch stats -----------------------
        if (tid == 0) {
            if constexpr (MODE != kModeBitmap) {
                TileInfo ti;
                ti.k_excl = k_excl;
                ti.last_abs = tile_k > 0 ? start + tile_last : -1;
                ti.carry_in = c_t;
                ti.mode = MODE;
                p.info[tile] = ti;
            }
            if (last_tile)
                p.stats->k_total = ((uint64_t)p.epoch << kEpochShift) | (uint64_t)(k_excl + tile_k);
        }

        // ---- output -------------------------------------------------------------
        uint8_t *const out_first = p.dst + start + c_t + k_excl * diff;   // tile's first output byte
        if (staged) {
            // s_in is no longer read: the next tile streams in while wave 0 copies
            // the staged output out with unaligned 16-byte stores
            if (need_src && tile + G < nt)
                prefetch_lds(p, tile + G, tid, halos, s_in);
            if (wave == 0) {
                const int len = (int)out_len;
                for (int c = lane * 16; c < len; c += kWave * 16) {
                    const uint4 v = *reinterpret_cast<const uint4 *>(s_out + c);
                    if (c + 16 <= len)
                        __builtin_memcpy(out_first + c, &v, 16);
                    else
                        store_tail(out_first + c, v, len - c);
                }
            }
        } else {
            if (p.write_output) {
                uint8_t *out = out_first + (int64_t)(seg + c_i - c_t) + (int64_t)k_loc * diff;
                int j = c_i;
                uint32_t M = j >= 32 ? 0u : (kept & (0xffffffffu << j));
                while (j < jend) {
                    const int q = M ? __builtin_ctz(M) : jend;
                    if (q > j) {
                        store_from_lds(out, s_in, kHalo + seg + j, q - j);
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
            __syncthreads();   // s_in fully read
            if (need_src && tile + G < nt)
                prefetch_lds(p, tile + G, tid, halos, s_in);
        }
    }
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
    const uint8_t p0 = p.pat[0];
    uint32_t m = 0;
    for (int j = 0; j < 32; ++j) {
        const int64_t pos = base + j;
        if (pos < p.floor || pos + L > p.n)
            continue;
        if (p.src[pos] != p0)
            continue;
        bool ok = true;
        for (int k = 1; k < L; ++k) {
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
    unsigned long long lb = 0;   // last + 1
    for (int64_t i = tid; i < t; i += kThreads) {
        const int64_t la = p.info[i].last_abs;
        if (la + 1 > (int64_t)lb)
            lb = (unsigned long long)(la + 1);
    }
    atomicMax(&s_last_before, lb);
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

// Largest kept start over all tiles (from the per-tile records), or -1.
__global__ void __launch_bounds__(kThreads)
last_kept_kernel(const TileInfo *info, int64_t num_tiles, int64_t *out1)
{
    __shared__ unsigned long long s_best;
    if (threadIdx.x == 0)
        s_best = 0;
    __syncthreads();
    unsigned long long best = 0;   // last + 1
    for (int64_t i = threadIdx.x; i < num_tiles; i += kThreads) {
        const int64_t la = info[i].last_abs;
        if ((unsigned long long)(la + 1) > best)
            best = (unsigned long long)(la + 1);
    }
    atomicMax(&s_best, best);
    __syncthreads();
    if (threadIdx.x == 0)
        out1[0] = (int64_t)s_best - 1;
}

// ---- presence index: 256-bit byte-value mask per chunk -----------------------
__global__ void __launch_bounds__(kThreads)
presence_kernel(const uint8_t *src, int64_t n, int64_t chunk, uint64_t *masks)
{
    __shared__ unsigned int s_bits[8];
    if (threadIdx.x < 8)
        s_bits[threadIdx.x] = 0;
    __syncthreads();
    const int64_t b = (int64_t)blockIdx.x * chunk;
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
        masks[blockIdx.x * 4 + threadIdx.x] =
            (uint64_t)s_bits[2 * threadIdx.x] | ((uint64_t)s_bits[2 * threadIdx.x + 1] << 32);
}

// ---- launchers --------------------------------------------------------------
// Persistent launch: one workgroup per resident slot (occupancy x CUs, at most one
// per tile).  Look-back modes launch cooperatively, which guarantees that all
// workgroups are resident at once.
template <int MODE, int LT>
hipError_t launch_tile(const ReplaceParams &p, hipStream_t s)
{
    constexpr int kMaxDev = 64;
    static int per_cu[kMaxDev];
    static int cus[kMaxDev];
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess)
        return e;
    if (dev < 0 || dev >= kMaxDev)
        return hipErrorInvalidDevice;
    if (per_cu[dev] <= 0) {
        int nb = 0, nc = 0;
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, tile_kernel<MODE, LT>, kThreads, 0);
        if (e != hipSuccess)
            return e;
        e = hipDeviceGetAttribute(&nc, hipDeviceAttributeMultiprocessorCount, dev);
        if (e != hipSuccess)
            return e;
        cus[dev] = nc > 0 ? nc : 1;
        // one slot per CU of margin: with the full reported occupancy not every
        // workgroup was resident at once on MI355X (look-backs then stall)
        per_cu[dev] = nb > 1 ? nb - 1 : 1;
        if (getenv("CR_DEBUG"))
            fprintf(stderr, "[cr] tile_kernel<%d,%d>: %d workgroups per CU x %d CUs\n", MODE, LT, per_cu[dev],
                    cus[dev]);
    }
    int64_t grid = (int64_t)per_cu[dev] * cus[dev];
    if (grid > p.num_tiles)
        grid = p.num_tiles;
    ReplaceParams pc = p;
    if (MODE == kModeBitmap) {
        hipLaunchKernelGGL((tile_kernel<MODE, LT>), dim3((unsigned)grid), dim3(kThreads), 0, s, pc);
        return hipGetLastError();
    }
    // CR_COOPERATIVE=0: plain launch of the same (resident-sized) grid; rocprofv3
    // crashes at process exit after cooperative launches and loses its trace
    static const bool coop = !(getenv("CR_COOPERATIVE") && getenv("CR_COOPERATIVE")[0] == '0');
    if (!coop) {
        hipLaunchKernelGGL((tile_kernel<MODE, LT>), dim3((unsigned)grid), dim3(kThreads), 0, s, pc);
        return hipGetLastError();
    }
    void *args[] = {&pc};
    return hipLaunchCooperativeKernel(reinterpret_cast<const void *>(&tile_kernel<MODE, LT>), dim3((unsigned)grid),
                                      dim3(kThreads), args, 0, s);
}

hipError_t launch_border_free(const ReplaceParams &p, hipStream_t s)
{
    return launch_tile<kModeBorderFree, 1>(p, s);
}

hipError_t launch_table(const ReplaceParams &p, int table_size, hipStream_t s)
{
    switch (table_size) {
    case 2:
        return launch_tile<kModeTable, 2>(p, s);
    case 4:
        return launch_tile<kModeTable, 4>(p, s);
    case 8:
        return launch_tile<kModeTable, 8>(p, s);
    case 16:
        return launch_tile<kModeTable, 16>(p, s);
    default:
        return hipErrorInvalidValue;
    }
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

hipError_t launch_bitmap_scatter(const ReplaceParams &p, hipStream_t s)
{
    return launch_tile<kModeBitmap, 1>(p, s);
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
    if (blocks > 65536)
        blocks = 65536;
    hipLaunchKernelGGL(fill_u64_kernel, dim3((unsigned)blocks), dim3(kThreads), 0, s, ptr, value, count);
    return hipGetLastError();
}

hipError_t launch_query(const ReplaceParams &p, int mode, int64_t limit, int64_t *out2, hipStream_t s)
{
    hipLaunchKernelGGL(query_kernel, dim3(1), dim3(kThreads), 0, s, p, mode, limit, out2);
    return hipGetLastError();
}

hipError_t launch_last_kept(const ReplaceParams &p, int64_t *out1, hipStream_t s)
{
    hipLaunchKernelGGL(last_kept_kernel, dim3(1), dim3(kThreads), 0, s, p.info, p.num_tiles, out1);
    return hipGetLastError();
}

hipError_t launch_presence_index(const uint8_t *src, int64_t n, int64_t chunk, uint64_t *masks,
                                 int64_t num_chunks, hipStream_t s)
{
    hipLaunchKernelGGL(presence_kernel, dim3((unsigned)num_chunks), dim3(kThreads), 0, s, src, n, chunk,
                       masks);
    return hipGetLastError();
}

} // namespace cr
