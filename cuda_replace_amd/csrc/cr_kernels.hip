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

namespace cr {

constexpr int kWave = 64;
constexpr int kWaves = kThreads / kWave;
constexpr int kNegInf = INT_MIN / 4;
// A look-back waits on tiles that already hold a ticket, so it ends; the bound only
// turns a logic error into a reported failure instead of a hung device.
constexpr int kSpinLimit = 1 << 20;

__device__ __forceinline__ void report_error(const ReplaceParams &p, uint64_t code)
{
    atomicMax((unsigned long long *)&p.stats->pad[0],
              (unsigned long long)(((uint64_t)p.epoch << kEpochShift) | code));
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
template <int MODE, int LT>
__global__ void __launch_bounds__(kThreads)
tile_kernel(ReplaceParams p)
{
    __shared__ __align__(16) uint8_t s_in[kHalo + kTile + kHalo];
    __shared__ __align__(16) uint8_t s_out[kOutStage];
    __shared__ uint8_t s_pat[kMaxSinglePassPat];
    __shared__ uint8_t s_rep[256];
    __shared__ int s_wsum[kWaves];
    __shared__ int s_wmax[kWaves];
    __shared__ int s_halo_last;
    __shared__ int64_t s_tile;
    __shared__ int64_t s_kexcl;
    __shared__ int s_carry_in;
    __shared__ int s_carry_out;
    __shared__ uint32_t s_wtab[kWaves][LT > 1 ? LT : 1];

    const int tid = threadIdx.x;
    const int lane = lane_id();
    const int wave = tid / kWave;
    const int L = p.L;
    const int r = p.r;
    const int64_t n = p.n;

    if (tid == 0) {
        s_tile = (int64_t)(atomicAdd(p.ticket, 1ull) - p.ticket_base);
        s_halo_last = kNegInf;
    }
    __syncthreads();
    const int64_t tile = s_tile;
    if (tile < 0 || tile >= p.num_tiles) {
        if (tid == 0)
            report_error(p, 1);
        return;
    }
    const int64_t start = tile * kTile;
    const int valid = (int)((n - start) < kTile ? (n - start) : kTile);
    const bool last_tile = tile == p.num_tiles - 1;
    const int64_t diff = (int64_t)r - (int64_t)L;

    // ---- stage source tile + halos, pattern, replacement -----------------------
    uint32_t m = 0;   // kept mask of this thread's 32 positions (after suppression)
    if constexpr (MODE == kModeBitmap) {
        m = p.bitmap[tile * (kTile / 32) + tid];
        if (p.write_output) {
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const int off = c * (kThreads * 16) + tid * 16;
                uint4 v = make_uint4(0, 0, 0, 0);
                if (off + 16 <= valid) {
                    v = *reinterpret_cast<const uint4 *>(p.src + start + off);
                } else if (off < valid) {
                    uint8_t b[16];
                    for (int i = 0; i < 16; ++i)
                        b[i] = (off + i < valid) ? p.src[start + off + i] : 0;
                    __builtin_memcpy(&v, b, 16);
                }
                *reinterpret_cast<uint4 *>(s_in + kHalo + off) = v;
            }
        }
    } else {
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int off = c * (kThreads * 16) + tid * 16;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (off + 16 <= valid) {
                v = *reinterpret_cast<const uint4 *>(p.src + start + off);
            } else if (off < valid) {
                    uint8_t b[16];
                for (int i = 0; i < 16; ++i)
                    b[i] = (off + i < valid) ? p.src[start + off + i] : 0;
                __builtin_memcpy(&v, b, 16);
            }
            *reinterpret_cast<uint4 *>(s_in + kHalo + off) = v;
        }
        for (int i = tid; i < L - 1; i += kThreads) {
            const int64_t g = start + kTile + i;
            s_in[kHalo + kTile + i] = g < n ? p.src[g] : 0;
            const int64_t h = start - (L - 1) + i;
            s_in[kHalo - (L - 1) + i] = h >= 0 ? p.src[h] : 0;
        }
        // bytes after the right halo may be read by a 4-byte window; mask4 drops them
        for (int i = tid; i < L; i += kThreads)
            s_pat[i] = p.pat[i];
    }
    if (p.write_output) {
        for (int i = tid; i < r && i < 256; i += kThreads)
            s_rep[i] = p.rep[i];
    }
    __syncthreads();

    const int seg = tid * kBytesPerThread;     // tile-relative start of my segment
    const int64_t gseg = start + seg;

    // ---- candidate mask --------------------------------------------------------
    if constexpr (MODE != kModeBitmap) {
        const uint32_t *w32 = reinterpret_cast<const uint32_t *>(s_in + kHalo + seg);
        uint32_t d[9];
#pragma unroll
        for (int k = 0; k < 9; ++k)
            d[k] = w32[k];
        const uint32_t pat4 = p.pat4, mask4 = p.mask4;
#pragma unroll
        for (int j = 0; j < 32; ++j) {
            const uint32_t w = __builtin_amdgcn_alignbyte(d[j / 4 + 1], d[j / 4], j % 4);
            m |= (uint32_t)(((w ^ pat4) & mask4) == 0u) << j;
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

    // ---- border-free: carry from candidates in the L-1 bytes left of the tile ----
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

    if constexpr (MODE == kModeBorderFree) {
        // decoupled look-back over kept counts
        uint64_t *state = p.state;
        if (wave == 0) {
            if (tile == 0) {
                if (lane == 0)
                    store_word(&state[0], make_word(p.epoch, kStatusInc, (uint64_t)tile_k));
            } else {
                if (lane == 0)
                    store_word(&state[tile], make_word(p.epoch, kStatusAgg, (uint64_t)tile_k));
                int64_t acc = 0;
                int64_t pred = tile - 1;
                int spins = 0;
                while (true) {
                    const int64_t idx = pred - lane;
                    uint64_t w = idx >= 0 ? load_word(&state[idx]) : make_word(p.epoch, kStatusInc, 0);
                    const uint32_t st = word_status(w, p.epoch);
                    const uint64_t inc_bal = __ballot(st == kStatusInc);
                    const uint64_t nr_bal = __ballot(st == 0);
                    const int first_inc = inc_bal ? __builtin_ctzll(inc_bal) : kWave;
                    const uint64_t need = first_inc >= kWave - 1 ? ~0ull : ((2ull << first_inc) - 1ull);
                    if (nr_bal & need) {
                        if (++spins > kSpinLimit) {
                            if (lane == 0)
                                report_error(p, 2);
                            break;
                        }
                        __builtin_amdgcn_s_sleep(1);
                        continue;
                    }
                    int64_t v = lane <= first_inc ? (int64_t)(w & kPayloadMask) : 0;
#pragma unroll
                    for (int dd = kWave / 2; dd > 0; dd >>= 1)
                        v += __shfl_xor(v, dd, kWave);
                    acc += v;
                    if (first_inc < kWave)
                        break;
                    pred -= kWave;
                }
                if (lane == 0)
                    store_word(&state[tile], make_word(p.epoch, kStatusInc, (uint64_t)(acc + tile_k)));
                if (lane == 0)
                    s_kexcl = acc;
            }
            if (tile == 0 && lane == 0)
                s_kexcl = 0;
        }
        __syncthreads();
        k_excl = s_kexcl;
    }

    uint32_t kept = m;
    if constexpr (MODE == kModeTable) {
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

        // look-back over tile tables: state[tile*(LT+1)] inclusive, [+1+c] aggregate
        uint64_t *state = p.state;
        const int64_t stride = LT + 1;
        if (wave == 0) {
            uint32_t cin = 0;
            int64_t kin = 0;
            if (tile > 0) {
                if (lane < LT) {
                    const uint32_t a = agg.e[lane];
                    store_word(&state[tile * stride + 1 + lane],
                               make_word(p.epoch, kStatusAgg,
                                         ((uint64_t)(a >> 24) << kCarryShift) | (uint64_t)(a & 0xffffffu)));
                }
                // H: composition of visited predecessor aggregates (lane c = entry c)
                uint32_t h_c = (uint32_t)(lane < LT ? lane : 0);
                int64_t h_k = 0;
                int64_t j = tile - 1;
                int spins = 0;
                while (true) {
                    if (j < 0) {
                        cin = __shfl(h_c, 0, kWave);
                        kin = __shfl(h_k, 0, kWave);
                        break;
                    }
                    const uint64_t iw = load_word(&state[j * stride]);
                    if (word_status(iw, p.epoch) == kStatusInc) {
                        const uint32_t c0 = (uint32_t)((iw & kPayloadMask) >> kCarryShift);
                        const int64_t k0 = (int64_t)(iw & kCountMask);
                        cin = __shfl(h_c, (int)c0, kWave);
                        kin = k0 + __shfl(h_k, (int)c0, kWave);
                        break;
                    }
                    uint64_t aw = lane < LT ? load_word(&state[j * stride + 1 + lane])
                                            : make_word(p.epoch, kStatusAgg, 0);
                    const uint64_t nr = __ballot(word_status(aw, p.epoch) == 0);
                    if (nr) {
                        if (++spins > kSpinLimit) {
                            if (lane == 0)
                                report_error(p, 2);
                            break;
                        }
                        __builtin_amdgcn_s_sleep(1);
                        continue;
                    }
                    const uint32_t ac = (uint32_t)((aw & kPayloadMask) >> kCarryShift);
                    const int64_t ak = (int64_t)(aw & kCountMask);
                    const uint32_t nh_c = __shfl(h_c, (int)(lane < LT ? ac : 0), kWave);
                    const int64_t nh_k = __shfl(h_k, (int)(lane < LT ? ac : 0), kWave);
                    h_c = nh_c;
                    h_k = ak + nh_k;
                    --j;
                }
            }
            const uint32_t a = tsel<LT>(agg, cin);
            if (lane == 0) {
                store_word(&state[tile * stride],
                           make_word(p.epoch, kStatusInc,
                                     ((uint64_t)(a >> 24) << kCarryShift) |
                                         (uint64_t)(kin + (int64_t)(a & 0xffffffu))));
                s_kexcl = kin;
                s_carry_in = (int)cin;
                s_carry_out = (int)(a >> 24);
            }
        }
        __syncthreads();
        k_excl = s_kexcl;
        c_t = s_carry_in;
        c_next = last_tile ? 0 : s_carry_out;
        const uint32_t me = tsel<LT>(thread_ex, (uint32_t)c_t);
        c_i = (int)(me >> 24);
        k_loc = (int)(me & 0xffffffu);
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
    }

    // ---- records for prefix queries and launch stats ---------------------------
    if (tid == 0) {
        TileInfo ti;
        ti.k_excl = k_excl;
        ti.last_abs = tile_k > 0 ? start + tile_last : -1;
        ti.carry_in = c_t;
        ti.mode = MODE;
        if constexpr (MODE != kModeBitmap)
            p.info[tile] = ti;
        if (tile_k > 0)
            atomicMax((unsigned long long *)&p.stats->last_kept,
                      (unsigned long long)(((uint64_t)p.epoch << kEpochShift) |
                                           (uint64_t)(start + tile_last + 1)));
        if (last_tile)
            p.stats->k_total = ((uint64_t)p.epoch << kEpochShift) | (uint64_t)(k_excl + tile_k);
    }
    if (!p.write_output)
        return;

    // ---- assemble the tile's output in LDS, write with aligned 16-byte stores ----
    const int64_t out_first = start + c_t + k_excl * diff;   // global output offset
    const int64_t out_len = (int64_t)valid + c_next - c_t + (int64_t)tile_k * diff;
    if (out_len <= 0)
        return;
    const int shift = (int)(out_first & 15);
    const int64_t my_rel = (int64_t)(seg + c_i - c_t) + (int64_t)k_loc * diff + shift;
    const int seg_valid = valid - seg;     // may be <= 0
    const uint8_t *sp = s_in + kHalo + seg;
    const uint8_t *rep_src = r <= 256 ? s_rep : p.rep;
    const int64_t total = shift + out_len;
    const int64_t g_aligned = out_first - shift;

    for (int64_t wbase = 0; wbase < total; wbase += kOutStage) {
        const bool single = total <= kOutStage;
        // emit my bytes that fall in [wbase, wbase + kOutStage)
        {
            int64_t pos = my_rel;
            int j = c_i;
            const int jend = seg_valid < 32 ? seg_valid : 32;
            if (single) {
                uint8_t *o = s_out + pos;
                while (j < jend) {
                    if ((kept >> j) & 1u) {
                        for (int q = 0; q < r; ++q)
                            o[q] = rep_src[q];
                        o += r;
                        j += L;
                    } else {
                        *o++ = sp[j++];
                    }
                }
            } else {
                const int64_t whi = wbase + kOutStage;
                while (j < jend && pos < whi) {
                    if ((kept >> j) & 1u) {
                        for (int q = 0; q < r; ++q) {
                            const int64_t x = pos + q;
                            if (x >= wbase && x < whi)
                                s_out[x - wbase] = rep_src[q];
                        }
                        pos += r;
                        j += L;
                    } else {
                        if (pos >= wbase)
                            s_out[pos - wbase] = sp[j];
                        ++pos;
                        ++j;
                    }
                }
            }
        }
        __syncthreads();
        // copy the window out
        const int64_t wlen = (total - wbase) < kOutStage ? (total - wbase) : kOutStage;
        for (int c = tid; c * 16 < wlen; c += kThreads) {
            const int64_t lo = wbase + (int64_t)c * 16;          // window-relative -> total-relative
            const int64_t g0 = g_aligned + lo;
            const bool full = lo >= shift && lo + 16 <= total;
            if (full) {
                *reinterpret_cast<uint4 *>(p.dst + g0) = *reinterpret_cast<const uint4 *>(s_out + c * 16);
            } else {
                for (int b = 0; b < 16; ++b) {
                    const int64_t x = lo + b;
                    if (x >= shift && x < total)
                        p.dst[g0 + b] = s_out[c * 16 + b];
                }
            }
        }
        __syncthreads();
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
hipError_t launch_border_free(const ReplaceParams &p, hipStream_t s)
{
    hipLaunchKernelGGL((tile_kernel<kModeBorderFree, 1>), dim3((unsigned)p.num_tiles), dim3(kThreads),
                       0, s, p);
    return hipGetLastError();
}

hipError_t launch_table(const ReplaceParams &p, int table_size, hipStream_t s)
{
    const dim3 g((unsigned)p.num_tiles), b(kThreads);
    switch (table_size) {
    case 2:
        hipLaunchKernelGGL((tile_kernel<kModeTable, 2>), g, b, 0, s, p);
        break;
    case 4:
        hipLaunchKernelGGL((tile_kernel<kModeTable, 4>), g, b, 0, s, p);
        break;
    case 8:
        hipLaunchKernelGGL((tile_kernel<kModeTable, 8>), g, b, 0, s, p);
        break;
    case 16:
        hipLaunchKernelGGL((tile_kernel<kModeTable, 16>), g, b, 0, s, p);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// The bitmap covers whole tiles (num_tiles * 256 words): words past n stay zero,
// so the scatter kernel can read a full tile of mask words.
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
    hipLaunchKernelGGL((tile_kernel<kModeBitmap, 1>), dim3((unsigned)p.num_tiles), dim3(kThreads), 0,
                       s, p);
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

hipError_t launch_presence_index(const uint8_t *src, int64_t n, int64_t chunk, uint64_t *masks,
                                 int64_t num_chunks, hipStream_t s)
{
    hipLaunchKernelGGL(presence_kernel, dim3((unsigned)num_chunks), dim3(kThreads), 0, s, src, n, chunk,
                       masks);
    return hipGetLastError();
}

} // namespace cr
