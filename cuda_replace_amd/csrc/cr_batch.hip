// cr_batch.hip -- batch engine: many independent documents of uneven size, each with
// its own ordered list of (pattern, replacement) pairs.
//
// Round j applies pair j of every document that has one.  A round is: upload the
// round's document descriptors and pattern pool; per pattern class (border-free, or
// carry tables of 2/4/8/16 states) one batched count launch over all groups of the
// class's documents and one spine launch (one wave per document); read back the
// documents' kept counts; pack the new lengths into the destination arena on the
// host; per class one batched emit launch.  Documents whose pattern is empty or needs
// the bitmap path run one at a time after the batched emit.  Documents that finish
// in a round are written into the final arena, packed in finishing order, so no
// later round overwrites them; one device-to-host copy returns them all.
//
// The engine keeps its arenas, descriptor buffers and pinned staging resident across
// runs (sized by the largest run so far).  No counterpart in the reference, whose
// callers loop over cuda_replace_unified (cuda_replace.cu:2579-2601) per document
// and pair.
#include "cr_internal.h"
#include "cr_pattern.h"
#include "cr_host.h"
#include "../../include/cuda_replace.h"

#include <algorithm>
#include <array>
#include <chrono>
#include <map>
#include <mutex>
#include <new>
#include <stdlib.h>
#include <string.h>
#include <thread>
#include <vector>

using namespace cr;

namespace {

constexpr size_t kDocAlign = 256;      // documents start at 256-byte boundaries in the arenas
constexpr size_t kGranule = 2u << 20;  // device buffers grow in 2 MiB steps
constexpr int kClasses = 5;            // border-free, tables of 2, 4, 8, 16 states
constexpr int kOther = kClasses;       // empty pattern or bitmap path: one at a time
constexpr int kMaxSummaryWords = kMaxTablePat;

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

#define CK(expr)                                  \
    do {                                          \
        hipError_t e_ = (expr);                   \
        if (e_ != hipSuccess)                     \
            return (int)e_;                       \
    } while (0)

struct DevBuf {
    uint8_t *p = nullptr;
    size_t cap = 0;
};
struct HostBuf {
    uint8_t *p = nullptr;
    size_t cap = 0;
};

// Grows b to at least need bytes; preserve copies the first `used` bytes over.
hipError_t dev_reserve(DevBuf &b, size_t need, bool preserve, size_t used, hipStream_t s)
{
    if (b.p && b.cap >= need)
        return hipSuccess;
    size_t nc = align_up(std::max(need, b.cap + b.cap / 2), kGranule);
    uint8_t *np = nullptr;
    hipError_t e = hipMalloc((void **)&np, nc);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        nc = align_up(need > 0 ? need : 1, kGranule);
        e = hipMalloc((void **)&np, nc);
        if (e != hipSuccess)
            return e;
    }
    if (b.p) {
        if (preserve && used > 0) {
            e = hipMemcpyAsync(np, b.p, used, hipMemcpyDeviceToDevice, s);
            if (e != hipSuccess) {
                (void)hipFree(np);
                return e;
            }
        }
        e = hipStreamSynchronize(s);   // queued work may still read the old buffer
        (void)hipFree(b.p);
        if (e != hipSuccess) {
            b.p = np;
            b.cap = nc;
            return e;
        }
    }
    b.p = np;
    b.cap = nc;
    return hipSuccess;
}

hipError_t host_reserve(HostBuf &b, size_t need)
{
    if (b.p && b.cap >= need)
        return hipSuccess;
    const size_t nc = align_up(std::max(need, b.cap + b.cap / 2), 1u << 20);
    void *np = nullptr;
    hipError_t e = hipHostMalloc(&np, nc, hipHostMallocDefault);
    if (e != hipSuccess)
        return e;
    if (b.p)
        (void)hipHostFree(b.p);
    b.p = (uint8_t *)np;
    b.cap = nc;
    return hipSuccess;
}

struct Piece {
    size_t dst_off;
    const uint8_t *src;
    size_t len;
};

// Host memcpy of many pieces into one buffer, split over threads by bytes.
void copy_pieces(uint8_t *dst, const std::vector<Piece> &pieces, size_t total)
{
    const int nthreads = total < (32u << 20) ? 1 : host_copy_threads();
    auto run = [&](size_t lo, size_t hi) {   // byte range [lo, hi) of the concatenation
        size_t acc = 0;
        for (const Piece &pc : pieces) {
            const size_t a = std::max(lo, acc), b = std::min(hi, acc + pc.len);
            if (a < b)
                memcpy(dst + pc.dst_off + (a - acc), pc.src + (a - acc), b - a);
            acc += pc.len;
            if (acc >= hi)
                break;
        }
    };
    if (nthreads <= 1) {
        run(0, total);
        return;
    }
    std::vector<std::thread> ts;
    for (int t = 0; t < nthreads; ++t)
        ts.emplace_back(run, total * t / nthreads, total * (t + 1) / nthreads);
    for (auto &t : ts)
        t.join();
}

// Packs the documents into the pinned staging h_in (at their arena offsets) and
// uploads them to dst, in slices of about 64 MiB: the host copy of slice k+1 runs
// while slice k's transfer is in flight.
hipError_t stage_and_upload(uint8_t *h_in, uint8_t *dst, const std::vector<Piece> &pieces, size_t arena_in,
                            hipStream_t s)
{
    constexpr size_t kSlice = 64u << 20;
    size_t i0 = 0;
    while (i0 < pieces.size()) {
        size_t i1 = i0, bytes = 0;
        while (i1 < pieces.size() && (bytes < kSlice || i1 == i0))
            bytes += pieces[i1++].len;
        const std::vector<Piece> part(pieces.begin() + (long)i0, pieces.begin() + (long)i1);
        copy_pieces(h_in, part, bytes);
        const size_t lo = pieces[i0].dst_off;
        const size_t hi = i1 < pieces.size() ? pieces[i1].dst_off : arena_in;
        if (hi > lo) {
            hipError_t e = hipMemcpyAsync(dst + lo, h_in + lo, hi - lo, hipMemcpyHostToDevice, s);
            if (e != hipSuccess)
                return e;
        }
        i0 = i1;
    }
    return hipSuccess;
}

struct DocState {
    uint8_t *cur = nullptr;   // device location of the current contents
    int64_t len = 0;
    int64_t pair0 = 0;        // first pair of the document in the flattened lists
    int npairs = 0;
    size_t fin_off = 0;       // final arena offset once finished
};

struct Engine {
    std::mutex mtx;
    int device = 0;
    int num_cus = 256;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    DevBuf arena[2], fin;
    DevBuf d_docs, d_groups, d_cdocs, d_pool, d_gsum, d_pre, d_kept, d_other;
    // descriptor staging alternates by round parity: round j+1 fills one copy while
    // round j's uploads may still read the other (round j+1's sync orders them)
    HostBuf h_docs[2], h_groups[2], h_cdocs[2], h_pool[2];
    HostBuf h_in, h_kept, h_out, h_misc;
    // device-driven path
    int mode = 0;   // 0 automatic, 1 host-driven rounds only, 2 device-driven when the pairs allow
    DevBuf g_pairs, g_pool, g_states, g_states0, g_ctl, g_ctl0, g_docs, g_kept;
    DevBuf g_groups[kBatchClasses], g_gsum[kBatchClasses], g_pre[kBatchClasses], g_rlist;
    HostBuf h_pairs, h_states, h_ctl, h_rlist;
    size_t hint_arena = 0, hint_fin = 0;       // sizes learnt from earlier runs
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    std::vector<int64_t> graph_key;            // shape of the captured run
    std::vector<int64_t> last_key;             // shape of the previous run
    double info[7] = {0, 0, 0, 0, 0, 0, 0};    // last run (see cr_batch_info)
};

struct DeviceGuard {
    int prev = -1;
    bool changed = false;
    explicit DeviceGuard(int dev)
    {
        if (hipGetDevice(&prev) == hipSuccess && prev != dev)
            changed = hipSetDevice(dev) == hipSuccess;
    }
    ~DeviceGuard()
    {
        if (changed)
            (void)hipSetDevice(prev);
    }
};

int class_of(int mode, int lt)
{
    if (mode == kModeBorderFree)
        return 0;
    if (mode == kModeTable)
        return lt == 2 ? 1 : lt == 4 ? 2 : lt == 8 ? 3 : 4;
    return kOther;
}

// One document of a round through the single-buffer kernels: empty pattern, or the
// bitmap path (candidate bitmap, serial greedy, emit).  Returns the output length
// via *out_len; dst_for(len) supplies the destination.
template <class DstFn>
int run_other(Engine *E, const DocState &ds, const uint8_t *pat, int L, const uint8_t *rep, int r,
              int64_t *out_len, DstFn &&dst_for)
{
    hipStream_t s = E->stream;
    const int64_t n = ds.len;
    // device copies of pattern and replacement at the front of the scratch
    const size_t pr = align_up((size_t)L + (size_t)r + 16, 256);
    const int64_t tiles = (n + kTile - 1) / kTile;
    const int64_t groups = (tiles + kGroupTiles - 1) / kGroupTiles;
    const size_t bitmap_bytes = align_up((size_t)(tiles > 0 ? tiles : 1) * (kTile / 8), 256);
    const size_t info_bytes = align_up((size_t)(tiles + 1) * sizeof(TileInfo), 256);
    // scratch of the parallel kept-start selection (0: the serial greedy runs)
    const bool bf = L > 0 && border_free(pat, L);
    const size_t sel_bytes = L > 0 && tiles > 0 ? align_up(bitmap_select_bytes(tiles, L, bf), 256) : 0;
    const size_t need = pr + 256 + (L > 0 ? bitmap_bytes + info_bytes + sel_bytes : 0);
    CK(dev_reserve(E->d_other, need, false, 0, s));
    CK(host_reserve(E->h_misc, pr + 256));
    if (L > 0)
        memcpy(E->h_misc.p, pat, (size_t)L);
    if (r > 0)
        memcpy(E->h_misc.p + L, rep, (size_t)r);
    if (L + r > 0)
        CK(hipMemcpyAsync(E->d_other.p, E->h_misc.p, (size_t)L + r, hipMemcpyHostToDevice, s));
    ReplaceParams p{};
    p.src = ds.cur;
    p.n = n;
    p.floor = 0;
    p.out_len = -1;
    p.pat = E->d_other.p;
    p.rep = E->d_other.p + L;
    p.L = L;
    p.r = r;
    pattern_head(pat, L, &p.pat4, &p.mask4);
    pattern_qsad(pat, L, p.qref, &p.qx);
    p.stats = reinterpret_cast<LaunchStats *>(E->d_other.p + pr);
    p.num_tiles = tiles;
    p.num_groups = groups;
    p.epoch = 1;
    if (L == 0) {
        const int64_t len = n + (n + 1) * (int64_t)r;
        uint8_t *dst = nullptr;
        int rc = dst_for(len, &dst);
        if (rc != 0)
            return rc;
        p.dst = dst;
        if (r == 0) {
            if (n > 0)
                CK(hipMemcpyAsync(dst, ds.cur, (size_t)n, hipMemcpyDeviceToDevice, s));
        } else {
            CK(launch_empty_pattern(p, s));
        }
        *out_len = len;
        return (int)hipStreamSynchronize(s);   // the pinned staging is reused next call
    }
    if (n == 0) {
        uint8_t *dst = nullptr;
        *out_len = 0;
        return dst_for(0, &dst);
    }
    p.bitmap = reinterpret_cast<uint32_t *>(E->d_other.p + pr + 256);
    p.info = reinterpret_cast<TileInfo *>(E->d_other.p + pr + 256 + bitmap_bytes);
    CK(launch_fill_u64(reinterpret_cast<uint64_t *>(p.stats), 0, 4, s));
    if (sel_bytes) {
        CK(launch_bitmap_select(p, bf, E->d_other.p + pr + 256 + bitmap_bytes + info_bytes, s));
    } else {
        CK(launch_bitmap_mark(p, s));
        CK(launch_bitmap_greedy(p, s));
    }
    CK(hipMemcpyAsync(E->h_misc.p + pr, p.stats, sizeof(LaunchStats), hipMemcpyDeviceToHost, s));
    CK(hipStreamSynchronize(s));
    const uint64_t kt = reinterpret_cast<const LaunchStats *>(E->h_misc.p + pr)->k_total;
    if ((uint32_t)(kt >> kEpochShift) != 1u)
        return (int)hipErrorUnknown;
    const int64_t kept = (int64_t)(kt & ((1ull << kEpochShift) - 1));
    const int64_t len = n + kept * ((int64_t)r - L);
    uint8_t *dst = nullptr;
    int rc = dst_for(len, &dst);
    if (rc != 0)
        return rc;
    p.dst = dst;
    p.out_len = len;
    p.write_output = 1;
    CK(launch_emit(p, kModeBitmap, 0, kGroupTiles, 0, groups, s));
    *out_len = len;
    return 0;
}

// Final slots of finished documents, packed into the engine's final arena.  The
// arena may move when it grows (its contents are kept); offsets stay valid.
struct FinArena {
    Engine *E;
    size_t used = 0;
    // room for `bytes` more, made before any address of this round is handed out
    int reserve(size_t bytes)
    {
        if (used + bytes + kDocAlign > E->fin.cap)
            return (int)dev_reserve(E->fin, used + bytes + kDocAlign, true, used, E->stream);
        return 0;
    }
    int slot(int64_t len, uint8_t **dst, size_t *off)
    {
        const size_t need = used + align_up((size_t)len, kDocAlign) + kDocAlign;
        if (need > E->fin.cap) {
            const hipError_t e = dev_reserve(E->fin, need, true, used, E->stream);
            if (e != hipSuccess)
                return (int)e;
        }
        *off = used;
        *dst = E->fin.p + used;
        used += align_up((size_t)len, kDocAlign);
        return 0;
    }
};

// One document's step of a round: apply (pat, rep) to its current contents.
struct RoundItem {
    DocState *d;
    const uint8_t *pat;
    int L;
    const uint8_t *rep;
    int r;
    bool last;            // the document's final pair: its output goes to a final slot
    int64_t new_len;      // output length (set by the round)
};

// One host-driven round over the items, each document at its own pair: class-grouped
// batched counts and spines, one readback of the kept counts, destinations
// (arena[par ^ 1] for documents that go on, final slots from fin for the
// others), batched emits, then empty-pattern and bitmap-path documents one at a
// time.  Afterwards every item's document holds its output (d->cur, d->len).  par:
// the round's parity (documents are read from arena[par]; it also picks the
// descriptor staging copy filled this round).
int run_round(Engine *E, std::vector<RoundItem> &items, int par, FinArena &fin, double *algo_bytes)
{
    hipStream_t s = E->stream;
    const int gt = kGroupTiles;
    const size_t A = items.size();
    // ---- classify, build descriptors, groups and the pattern pool ----------
    std::vector<int> cls(A), lt(A);
    size_t pool = 0;
    int64_t cls_groups[kClasses + 1] = {0};
    int64_t cls_docs[kClasses + 1] = {0};
    for (size_t a = 0; a < A; ++a) {
        const DocState &d = *items[a].d;
        int t = 0;
        const int mode = pattern_mode(items[a].pat, items[a].L, &t);
        cls[a] = class_of(mode, t);
        lt[a] = t;
        pool += align_up((size_t)items[a].L + 16, 16) + align_up((size_t)items[a].r + 32, 16);
        if (cls[a] < kClasses) {
            const int64_t tiles = (d.len + kTile - 1) / kTile;
            cls_groups[cls[a]] += (tiles + gt - 1) / gt;
            cls_docs[cls[a]] += 1;
        }
    }
    int64_t g_off[kClasses + 1], d_off[kClasses + 1];
    int64_t G = 0, D = 0;
    for (int c = 0; c < kClasses; ++c) {
        g_off[c] = G;
        d_off[c] = D;
        G += cls_groups[c];
        D += cls_docs[c];
    }
    CK(host_reserve(E->h_docs[par], A * sizeof(BatchDoc) + 64));
    CK(host_reserve(E->h_groups[par], (size_t)G * sizeof(BatchGroup) + 64));
    CK(host_reserve(E->h_cdocs[par], (size_t)D * sizeof(int32_t) + 64));
    CK(host_reserve(E->h_pool[par], pool + 64));
    CK(host_reserve(E->h_kept, A * sizeof(int64_t) + 64));
    BatchDoc *hd = reinterpret_cast<BatchDoc *>(E->h_docs[par].p);
    BatchGroup *hg = reinterpret_cast<BatchGroup *>(E->h_groups[par].p);
    int32_t *hc = reinterpret_cast<int32_t *>(E->h_cdocs[par].p);
    uint8_t *hpool = E->h_pool[par].p;
    int64_t g_fill[kClasses], d_fill[kClasses];
    for (int c = 0; c < kClasses; ++c) {
        g_fill[c] = g_off[c];
        d_fill[c] = d_off[c];
    }
    size_t pool_fill = 0;
    for (size_t a = 0; a < A; ++a) {
        const DocState &d = *items[a].d;
        const int L = items[a].L, r = items[a].r;
        BatchDoc bd{};
        bd.src = d.cur;
        bd.dst = nullptr;
        bd.n = d.len;
        bd.L = L;
        bd.r = r;
        pattern_head(items[a].pat, L, &bd.pat4, &bd.mask4);
        bd.pat_off = (int32_t)pool_fill;
        if (L > 0)
            memcpy(hpool + pool_fill, items[a].pat, (size_t)L);
        pool_fill += align_up((size_t)L + 16, 16);
        bd.rep_off = (int32_t)pool_fill;
        if (r > 0)
            memcpy(hpool + pool_fill, items[a].rep, (size_t)r);
        pool_fill += align_up((size_t)r + 32, 16);
        const int c = cls[a];
        if (c < kClasses) {
            const int64_t tiles = (d.len + kTile - 1) / kTile;
            const int64_t ng = (tiles + gt - 1) / gt;
            bd.group0 = g_fill[c] - g_off[c];
            bd.ngroups = (int32_t)ng;
            for (int64_t g = 0; g < ng; ++g)
                hg[g_fill[c] + g] = BatchGroup{(int32_t)a, (int32_t)g};
            g_fill[c] += ng;
            hc[d_fill[c]++] = (int32_t)a;
        }
        hd[a] = bd;
    }
    CK(dev_reserve(E->d_docs, A * sizeof(BatchDoc) + 64, false, 0, s));
    CK(dev_reserve(E->d_groups, (size_t)G * sizeof(BatchGroup) + 64, false, 0, s));
    CK(dev_reserve(E->d_cdocs, (size_t)D * sizeof(int32_t) + 64, false, 0, s));
    CK(dev_reserve(E->d_pool, pool + 64, false, 0, s));
    CK(dev_reserve(E->d_gsum, (size_t)G * kMaxSummaryWords * sizeof(uint32_t) + 64, false, 0, s));
    CK(dev_reserve(E->d_pre, (size_t)G * sizeof(ChunkState) + 64, false, 0, s));
    CK(dev_reserve(E->d_kept, A * sizeof(int64_t) + 64, false, 0, s));
    if (A > 0)
        CK(hipMemcpyAsync(E->d_docs.p, hd, A * sizeof(BatchDoc), hipMemcpyHostToDevice, s));
    if (G > 0)
        CK(hipMemcpyAsync(E->d_groups.p, hg, (size_t)G * sizeof(BatchGroup), hipMemcpyHostToDevice, s));
    if (D > 0)
        CK(hipMemcpyAsync(E->d_cdocs.p, hc, (size_t)D * sizeof(int32_t), hipMemcpyHostToDevice, s));
    if (pool_fill > 0)
        CK(hipMemcpyAsync(E->d_pool.p, hpool, pool_fill, hipMemcpyHostToDevice, s));

    auto args_for = [&](int c) {
        BatchArgs ba{};
        ba.docs = reinterpret_cast<const BatchDoc *>(E->d_docs.p);
        ba.groups = reinterpret_cast<const BatchGroup *>(E->d_groups.p) + g_off[c];
        ba.class_docs = reinterpret_cast<const int32_t *>(E->d_cdocs.p) + d_off[c];
        ba.pool = E->d_pool.p;
        ba.gsum = reinterpret_cast<uint32_t *>(E->d_gsum.p) + g_off[c] * kMaxSummaryWords;
        ba.pre = reinterpret_cast<ChunkState *>(E->d_pre.p) + g_off[c];
        ba.doc_kept = reinterpret_cast<int64_t *>(E->d_kept.p);
        return ba;
    };
    static const int kClassMode[kClasses] = {kModeBorderFree, kModeTable, kModeTable, kModeTable, kModeTable};
    static const int kClassLt[kClasses] = {1, 2, 4, 8, 16};
    for (int c = 0; c < kClasses; ++c) {
        if (cls_docs[c] == 0)
            continue;
        const BatchArgs ba = args_for(c);
        CK(launch_batch_count(ba, kClassMode[c], kClassLt[c], gt, cls_groups[c], s));
        CK(launch_batch_spine(ba, kClassMode[c], kClassLt[c], cls_docs[c], s));
    }
    if (A > 0)
        CK(hipMemcpyAsync(E->h_kept.p, E->d_kept.p, A * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    CK(hipStreamSynchronize(s));

    // ---- new lengths, destinations -----------------------------------------
    const int nxt = par ^ 1;
    const int64_t *kept = reinterpret_cast<const int64_t *>(E->h_kept.p);
    std::vector<int64_t> new_len(A, 0);
    size_t arena_need = 0;
    for (size_t a = 0; a < A; ++a) {
        if (cls[a] >= kClasses)
            continue;
        const DocState &d = *items[a].d;
        new_len[a] = d.len + kept[a] * ((int64_t)items[a].r - items[a].L);
        if (!items[a].last)
            arena_need += align_up((size_t)new_len[a], kDocAlign);
    }
    CK(dev_reserve(E->arena[nxt], arena_need + kDocAlign, false, 0, s));
    // reserve the final arena for every class document finishing this round
    // before handing out addresses, so no address below moves before its emit
    size_t fin_need = 0;
    for (size_t a = 0; a < A; ++a)
        if (cls[a] < kClasses && items[a].last)
            fin_need += align_up((size_t)new_len[a], kDocAlign);
    {
        const int rc = fin.reserve(fin_need);
        if (rc != 0)
            return rc;
    }
    size_t arena_used = 0;
    for (size_t a = 0; a < A; ++a) {
        if (cls[a] >= kClasses)
            continue;
        DocState &d = *items[a].d;
        uint8_t *dst;
        if (items[a].last) {
            int rc = fin.slot(new_len[a], &dst, &d.fin_off);
            if (rc != 0)
                return rc;
        } else {
            dst = E->arena[nxt].p + arena_used;
            arena_used += align_up((size_t)new_len[a], kDocAlign);
        }
        hd[a].dst = dst;
    }
    if (A > 0)
        CK(hipMemcpyAsync(E->d_docs.p, hd, A * sizeof(BatchDoc), hipMemcpyHostToDevice, s));
    for (int c = 0; c < kClasses; ++c) {
        if (cls_docs[c] == 0)
            continue;
        CK(launch_batch_emit(args_for(c), kClassMode[c], kClassLt[c], gt, cls_groups[c], s));
    }
    // ---- empty-pattern and bitmap-path documents, one at a time ------------
    // (after a sync: the loop edits descriptors the upload above reads)
    if (std::find_if(cls.begin(), cls.end(), [](int c) { return c >= kClasses; }) != cls.end())
        CK(hipStreamSynchronize(s));
    for (size_t a = 0; a < A; ++a) {
        if (cls[a] < kClasses)
            continue;
        DocState &d = *items[a].d;
        const bool last = items[a].last;
        auto dst_for = [&](int64_t len, uint8_t **dst) -> int {
            if (last)
                return fin.slot(len, dst, &d.fin_off);
            const size_t need = arena_used + align_up((size_t)len, kDocAlign) + kDocAlign;
            if (need > E->arena[nxt].cap) {
                // arena[nxt] already holds this round's outputs: move them (queued
                // after their emits) and re-point their documents
                DevBuf old = E->arena[nxt];
                E->arena[nxt] = DevBuf{};
                hipError_t e = dev_reserve(E->arena[nxt], need, false, 0, s);
                if (e != hipSuccess) {
                    E->arena[nxt] = old;
                    return (int)e;
                }
                if (arena_used > 0) {
                    e = hipMemcpyAsync(E->arena[nxt].p, old.p, arena_used, hipMemcpyDeviceToDevice, s);
                    if (e != hipSuccess)
                        return (int)e;
                }
                for (size_t b = 0; b < A; ++b)
                    if (hd[b].dst >= old.p && hd[b].dst < old.p + old.cap)
                        hd[b].dst = E->arena[nxt].p + (hd[b].dst - old.p);
                e = hipStreamSynchronize(s);
                (void)hipFree(old.p);
                if (e != hipSuccess)
                    return (int)e;
            }
            *dst = E->arena[nxt].p + arena_used;
            arena_used += align_up((size_t)len, kDocAlign);
            return 0;
        };
        int64_t len = 0;
        int rc = run_other(E, d, items[a].pat, items[a].L, items[a].rep, items[a].r, &len, dst_for);
        if (rc != 0)
            return rc;
        new_len[a] = len;
        // record where the output went (dst_for may have moved other documents)
        hd[a].dst = last ? E->fin.p + d.fin_off : E->arena[nxt].p + (arena_used - align_up((size_t)len, kDocAlign));
    }
    for (size_t a = 0; a < A; ++a) {
        DocState &d = *items[a].d;
        *algo_bytes += (double)d.len + (double)new_len[a];
        d.len = new_len[a];
        d.cur = items[a].last ? E->fin.p + d.fin_off : hd[a].dst;
        items[a].new_len = new_len[a];
    }
    return 0;
}

int engine_run(Engine *E, int num_docs, const uint8_t *const *docs, const size_t *doc_lens, const int *pair_counts,
               const uint8_t *const *pats, const int *pat_lens, const uint8_t *const *reps, const int *rep_lens,
               const uint8_t **out_base, size_t *out_offsets, size_t *out_lens, double *stats)
{
    const auto t_start = std::chrono::steady_clock::now();
    hipStream_t s = E->stream;
    if (num_docs < 0 || !out_base || (num_docs > 0 && (!docs || !doc_lens || !pair_counts || !out_offsets ||
                                                         !out_lens)))
        return -1;
    std::vector<DocState> st((size_t)num_docs);
    int64_t total_pairs = 0;
    int max_pairs = 0;
    size_t arena_in = 0;
    std::vector<Piece> pieces;
    pieces.reserve((size_t)num_docs);
    for (int i = 0; i < num_docs; ++i) {
        if (pair_counts[i] < 0 || (doc_lens[i] > 0 && !docs[i]))
            return -1;
        st[(size_t)i].pair0 = total_pairs;
        st[(size_t)i].npairs = pair_counts[i];
        st[(size_t)i].len = (int64_t)doc_lens[i];
        total_pairs += pair_counts[i];
        max_pairs = std::max(max_pairs, pair_counts[i]);
        pieces.push_back(Piece{arena_in, docs[i], doc_lens[i]});
        arena_in += align_up(doc_lens[i], kDocAlign);
    }
    if (total_pairs > 0 && (!pats || !pat_lens || !reps || !rep_lens))
        return -1;
    for (int64_t q = 0; q < total_pairs; ++q) {
        if (pat_lens[q] < 0 || rep_lens[q] < 0 || (pat_lens[q] > 0 && !pats[q]) || (rep_lens[q] > 0 && !reps[q]))
            return -1;
    }

    // ---- upload every document, packed at 256-byte boundaries -------------------
    CK(host_reserve(E->h_in, arena_in + 1));
    CK(dev_reserve(E->arena[0], arena_in + 1, false, 0, s));
    CK(dev_reserve(E->fin, align_up(arena_in + arena_in / 4, kDocAlign) + kGranule, false, 0, s));
    CK(stage_and_upload(E->h_in.p, E->arena[0].p, pieces, arena_in, s));
    for (int i = 0; i < num_docs; ++i)
        st[(size_t)i].cur = E->arena[0].p + pieces[(size_t)i].dst_off;

    FinArena fin{E};
    // documents without pairs are final as uploaded
    for (int i = 0; i < num_docs; ++i) {
        DocState &d = st[(size_t)i];
        if (d.npairs > 0)
            continue;
        uint8_t *dst = nullptr;
        int rc = fin.slot(d.len, &dst, &d.fin_off);
        if (rc != 0)
            return rc;
        if (d.len > 0)
            CK(hipMemcpyAsync(dst, d.cur, (size_t)d.len, hipMemcpyDeviceToDevice, s));
    }

    CK(hipEventRecord(E->ev0, s));
    double algo_bytes = 0;
    std::vector<RoundItem> items;
    for (int j = 0; j < max_pairs; ++j) {
        items.clear();
        for (int i = 0; i < num_docs; ++i) {
            DocState &d = st[(size_t)i];
            if (d.npairs <= j)
                continue;
            const int64_t q = d.pair0 + j;
            items.push_back(RoundItem{&d, pats[q], pat_lens[q], reps[q], rep_lens[q], d.npairs == j + 1, 0});
        }
        const int rc = run_round(E, items, j & 1, fin, &algo_bytes);
        if (rc != 0)
            return rc;
    }
    CK(hipEventRecord(E->ev1, s));

    // ---- results ----------------------------------------------------------------
    CK(host_reserve(E->h_out, fin.used + 64));
    if (fin.used > 0)
        CK(hipMemcpyAsync(E->h_out.p, E->fin.p, fin.used, hipMemcpyDeviceToHost, s));
    CK(hipStreamSynchronize(s));
    for (int i = 0; i < num_docs; ++i) {
        out_offsets[i] = st[(size_t)i].fin_off;
        out_lens[i] = (size_t)st[(size_t)i].len;
    }
    *out_base = E->h_out.p;
    if (stats) {
        float ms = 0.f;
        CK(hipEventElapsedTime(&ms, E->ev0, E->ev1));
        stats[0] = ms;
        stats[1] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
        stats[2] = algo_bytes;
        stats[3] = max_pairs;
    }
    return 0;
}

// ---- live runs (continuous batching) ----------------------------------------
struct LiveDoc {
    DocState d;
    LiveReq req;
    int next = 0;   // the pair this round applies
};

int engine_live(Engine *E, const std::function<bool(std::vector<LiveReq> &, bool)> &admit,
                const std::function<void(void *, int, const uint8_t *, size_t)> &finish)
{
    hipStream_t s = E->stream;
    std::vector<LiveDoc> live;
    std::vector<LiveReq> incoming;
    std::vector<RoundItem> items;
    std::vector<Piece> pieces;
    int cur = 0;           // arena holding the live documents
    size_t cur_used = 0;   // their bytes, packed from its start
    auto fail_all = [&](int rc) {
        for (LiveDoc &ld : live)
            finish(ld.req.tag, rc, nullptr, 0);
        live.clear();
        cur_used = 0;
        return rc;
    };
    for (;;) {
        incoming.clear();
        if (!admit(incoming, live.empty()) && live.empty())
            return 0;
        // ---- newcomers: uploaded behind the live documents --------------------
        // (every request with pairs joins the live set, an empty document too)
        pieces.clear();
        size_t add = 0;
        bool joins = false;
        for (const LiveReq &r : incoming) {
            if (r.npairs == 0) {
                finish(r.tag, 0, r.doc, r.len);
                continue;
            }
            joins = true;
            pieces.push_back(Piece{add, r.doc, r.len});
            add += align_up(r.len, kDocAlign);
        }
        if (joins) {
            uint8_t *const old = E->arena[cur].p;
            hipError_t e = dev_reserve(E->arena[cur], cur_used + add + kDocAlign, true, cur_used, s);
            if (e == hipSuccess)
                e = host_reserve(E->h_in, add + 1);
            if (e != hipSuccess) {
                for (const LiveReq &r : incoming)
                    if (r.npairs > 0)
                        finish(r.tag, (int)e, nullptr, 0);
                if (live.empty())
                    continue;
                return fail_all((int)e);
            }
            if (E->arena[cur].p != old)
                for (LiveDoc &ld : live)
                    ld.d.cur = E->arena[cur].p + (ld.d.cur - old);
            if (add > 0) {
                e = stage_and_upload(E->h_in.p, E->arena[cur].p + cur_used, pieces, add, s);
                if (e != hipSuccess) {
                    for (const LiveReq &r : incoming)
                        if (r.npairs > 0)
                            finish(r.tag, (int)e, nullptr, 0);
                    return fail_all((int)e);
                }
            }
            size_t k = 0;
            for (const LiveReq &r : incoming) {
                if (r.npairs == 0)
                    continue;
                LiveDoc ld;
                ld.d.cur = E->arena[cur].p + cur_used + pieces[k++].dst_off;
                ld.d.len = (int64_t)r.len;
                ld.d.npairs = r.npairs;
                ld.req = r;
                live.push_back(ld);
            }
            cur_used += add;
        }
        if (live.empty())
            continue;
        // ---- one round: every live document applies its next pair --------------
        items.clear();
        for (LiveDoc &ld : live) {
            const int q = ld.next;
            items.push_back(RoundItem{&ld.d, ld.req.pats[q], ld.req.pat_lens[q], ld.req.reps[q], ld.req.rep_lens[q],
                                      q + 1 == ld.req.npairs, 0});
        }
        FinArena fin{E};
        double algo = 0;
        int rc = run_round(E, items, cur, fin, &algo);
        if (rc != 0)
            return fail_all(rc);
        // ---- finished documents go back; the others continue from arena[cur ^ 1]
        // (a round without finished documents needs no wait here: the next round's
        // readback orders everything after it)
        hipError_t e = hipSuccess;
        if (fin.used > 0) {
            e = host_reserve(E->h_out, fin.used + 64);
            if (e == hipSuccess)
                e = hipMemcpyAsync(E->h_out.p, E->fin.p, fin.used, hipMemcpyDeviceToHost, s);
            if (e == hipSuccess)
                e = hipStreamSynchronize(s);
        }
        if (e != hipSuccess)
            return fail_all((int)e);
        const int nxt = cur ^ 1;
        size_t used = 0;
        size_t w = 0;
        for (size_t i = 0; i < live.size(); ++i) {
            LiveDoc &ld = live[i];
            if (ld.next + 1 == ld.req.npairs) {
                finish(ld.req.tag, 0, E->h_out.p + ld.d.fin_off, (size_t)ld.d.len);
                continue;
            }
            ld.next += 1;
            const size_t end = (size_t)(ld.d.cur - E->arena[nxt].p) + align_up((size_t)ld.d.len, kDocAlign);
            used = std::max(used, end);
            live[w++] = ld;
        }
        live.resize(w);
        cur = nxt;
        cur_used = used;
    }
}

// ---- device-driven runs -------------------------------------------------------
// Every round of the run is planned, counted, placed and emitted on the GPU
// (launch_batch_round), so the host issues the whole run at once: captured as a HIP
// graph, replayed while the run's shape and the engine's buffers stay the same.
// The work lists and arenas are sized from the input (and from earlier runs); a run
// that would not fit sets the device's overflow flag, the host-driven rounds redo
// it, and the next run sizes up.
constexpr int kFallback = 1 << 20;   // internal: redo the run with host-driven rounds

int env_int(const char *name, int dflt)
{
    const char *e = getenv(name);
    return e && *e ? atoi(e) : dflt;
}

// Tiles per group for device-driven runs, per device and input size bucket: the
// first run of a bucket executes round 0 for each candidate (1, 2, 4) from the same
// starting state, in three interleaved sweeps timed with HIP events, and keeps the
// fastest (another size than 2 only when more than 3% faster).
std::mutex g_tune_mtx;
std::map<std::pair<int, int>, std::pair<int, std::vector<float>>> g_tune;

int batch_bucket(size_t bytes)
{
    if (bytes < (4u << 20))
        return 0;
    if (bytes < (64u << 20))
        return 1;
    if (bytes < ((size_t)1 << 30))
        return 2;
    return 3;
}

// Round j's classes, document counts and estimated items per class: the groups of
// each document at the smaller of its worst-case length (every possible occurrence
// of every pair so far kept) and 1.5x its input length (at least the input
// length).  Items past the estimate are still done, by the round's persistent rest
// kernels.
BatchRoundShape round_shape(const std::vector<DocState> &st, const std::vector<int64_t> &ub_round,
                            const std::vector<int8_t> &pcls, int j, int gt, int cus)
{
    BatchRoundShape sh{};
    int64_t est[kBatchClasses] = {0, 0, 0, 0, 0};
    for (size_t i = 0; i < st.size(); ++i) {
        const DocState &d = st[i];
        if (d.npairs <= j)
            continue;
        const int c = pcls[(size_t)(d.pair0 + j)];
        sh.present[c] = 1;
        sh.class_docs[c] += 1;
        if (c < kBatchClasses) {
            const int64_t len = std::min(ub_round[i], d.len + d.len / 2);
            est[c] += ((len + kTile - 1) / kTile + gt - 1) / gt;
        }
    }
    for (int c = 0; c < kBatchClasses; ++c)
        sh.work_grid[c] = (int)std::max<int64_t>(1, std::min<int64_t>(est[c], 1 << 22));
    sh.cus = cus;
    return sh;
}

// Worst-case length of every document before round j (ub[j][i]), capped at 2^50.
std::vector<std::vector<int64_t>> length_bounds(const std::vector<DocState> &st, int max_pairs, const int *pat_lens,
                                                const int *rep_lens)
{
    std::vector<std::vector<int64_t>> ub((size_t)max_pairs, std::vector<int64_t>(st.size(), 0));
    const int64_t kCap = 1ll << 50;
    for (size_t i = 0; i < st.size(); ++i) {
        int64_t u = st[i].len;
        for (int j = 0; j < max_pairs; ++j) {
            ub[(size_t)j][i] = u;
            if (j >= st[i].npairs)
                continue;
            const int64_t q = st[i].pair0 + j;
            const int64_t L = pat_lens[q], r = rep_lens[q];
            if (L == 0)
                u = u > kCap / (r + 1) ? kCap : u + (u + 1) * r;
            else if (r > L)
                u = u / L > kCap / (r - L) ? kCap : u + (u / L) * (r - L);
            u = std::min(u, kCap);
        }
    }
    return ub;
}

bool valid_run(int num_docs, const uint8_t *const *docs, const size_t *doc_lens, const int *pair_counts,
               const uint8_t *const *pats, const int *pat_lens, const uint8_t *const *reps, const int *rep_lens,
               const uint8_t **out_base, size_t *out_offsets, size_t *out_lens)
{
    if (num_docs < 0 || !out_base || (num_docs > 0 && (!docs || !doc_lens || !pair_counts || !out_offsets ||
                                                         !out_lens)))
        return false;
    int64_t total_pairs = 0;
    for (int i = 0; i < num_docs; ++i) {
        if (pair_counts[i] < 0 || (doc_lens[i] > 0 && !docs[i]))
            return false;
        total_pairs += pair_counts[i];
    }
    if (total_pairs > 0 && (!pats || !pat_lens || !reps || !rep_lens))
        return false;
    for (int64_t q = 0; q < total_pairs; ++q)
        if (pat_lens[q] < 0 || rep_lens[q] < 0 || (pat_lens[q] > 0 && !pats[q]) || (rep_lens[q] > 0 && !reps[q]))
            return false;
    return true;
}

int engine_run_device(Engine *E, int num_docs, const uint8_t *const *docs, const size_t *doc_lens,
                      const int *pair_counts, const uint8_t *const *pats, const int *pat_lens,
                      const uint8_t *const *reps, const int *rep_lens, const uint8_t **out_base,
                      size_t *out_offsets, size_t *out_lens, double *stats)
{
    const auto t_start = std::chrono::steady_clock::now();
    hipStream_t s = E->stream;
    std::vector<DocState> st((size_t)num_docs);
    int64_t total_pairs = 0;
    int max_pairs = 0;
    size_t arena_in = 0, fin0 = 0;
    std::vector<Piece> pieces;
    pieces.reserve((size_t)num_docs);
    for (int i = 0; i < num_docs; ++i) {
        DocState &d = st[(size_t)i];
        d.pair0 = total_pairs;
        d.npairs = pair_counts[i];
        d.len = (int64_t)doc_lens[i];
        total_pairs += pair_counts[i];
        max_pairs = std::max(max_pairs, pair_counts[i]);
        pieces.push_back(Piece{arena_in, docs[i], doc_lens[i]});
        arena_in += align_up(doc_lens[i], kDocAlign);
        if (d.npairs == 0) {
            d.fin_off = fin0;
            fin0 += align_up(doc_lens[i], kDocAlign);
        }
    }
    // pair classes; the bitmap path is host-driven only
    std::vector<int8_t> pcls((size_t)total_pairs);
    size_t pool = 0;
    for (int64_t q = 0; q < total_pairs; ++q) {
        int lt = 0;
        const int mode = pattern_mode(pats[q], pat_lens[q], &lt);
        if (mode == kModeEmpty) {
            pcls[(size_t)q] = kBatchEmpty;
        } else {
            const int c = class_of(mode, lt);
            if (c >= kClasses)
                return kFallback;
            pcls[(size_t)q] = (int8_t)c;
        }
        pool += align_up((size_t)pat_lens[q] + 16, 16) + align_up((size_t)rep_lens[q] + 32, 16);
    }

    // ---- buffers ------------------------------------------------------------------
    const size_t grow_hint = arena_in + arena_in / 4 + (1u << 20);
    const size_t cap_a = std::max(grow_hint, E->hint_arena);
    CK(dev_reserve(E->arena[0], cap_a, false, 0, s));
    CK(dev_reserve(E->arena[1], cap_a, false, 0, s));
    CK(dev_reserve(E->fin, fin0 + std::max(grow_hint, E->hint_fin), false, 0, s));
    int64_t group_cap = (int64_t)(std::max(E->arena[0].cap, E->arena[1].cap) / kTile) + num_docs + 1;
    static const int kWords[kBatchClasses] = {1, 2, 4, 8, 16};
    auto reserve_lists = [&]() -> hipError_t {
        for (int c = 0; c < kBatchClasses; ++c) {
            hipError_t e = dev_reserve(E->g_groups[c], (size_t)group_cap * sizeof(BatchGroup), false, 0, s);
            if (e == hipSuccess)
                e = dev_reserve(E->g_gsum[c], (size_t)group_cap * kWords[c] * sizeof(uint32_t), false, 0, s);
            if (e == hipSuccess)
                e = dev_reserve(E->g_pre[c], (size_t)group_cap * sizeof(ChunkState), false, 0, s);
            if (e != hipSuccess)
                return e;
        }
        return hipSuccess;
    };
    CK(reserve_lists());
    const size_t D = (size_t)std::max(num_docs, 1);
    CK(dev_reserve(E->g_rlist, (size_t)std::max<int64_t>(total_pairs, 1) * sizeof(int32_t), false, 0, s));
    CK(dev_reserve(E->g_docs, D * sizeof(BatchDoc), false, 0, s));
    CK(dev_reserve(E->g_kept, D * sizeof(int64_t), false, 0, s));
    CK(dev_reserve(E->g_states, D * sizeof(BatchDocState), false, 0, s));
    CK(dev_reserve(E->g_states0, D * sizeof(BatchDocState), false, 0, s));
    CK(dev_reserve(E->g_ctl, sizeof(BatchCtl), false, 0, s));
    CK(dev_reserve(E->g_ctl0, sizeof(BatchCtl), false, 0, s));
    CK(dev_reserve(E->g_pairs, (size_t)std::max<int64_t>(total_pairs, 1) * sizeof(BatchPair), false, 0, s));
    CK(dev_reserve(E->g_pool, pool + 64, false, 0, s));

    // ---- staging and uploads --------------------------------------------------------
    CK(host_reserve(E->h_in, arena_in + 1));
    CK(host_reserve(E->h_pairs, (size_t)std::max<int64_t>(total_pairs, 1) * sizeof(BatchPair) + pool + 64));
    CK(host_reserve(E->h_states, D * sizeof(BatchDocState) + sizeof(BatchCtl) + 64));
    CK(host_reserve(E->h_ctl, sizeof(BatchCtl) + 64));
    CK(host_reserve(E->h_rlist, (size_t)std::max<int64_t>(total_pairs, 1) * sizeof(int32_t) + 64));
    // each round's active documents grouped by class (classes and activity follow
    // from the pairs alone)
    std::vector<int64_t> rbase((size_t)max_pairs + 1, 0);
    std::vector<std::array<int32_t, kBatchClasses + 2>> rseg((size_t)max_pairs);
    {
        // counting sort of (round, class, document) over every pair
        const int nb = kBatchClasses + 1;
        std::vector<int64_t> cnt((size_t)max_pairs * nb + 1, 0);
        for (int i = 0; i < num_docs; ++i)
            for (int j = 0; j < st[(size_t)i].npairs; ++j)
                ++cnt[(size_t)j * nb + (size_t)pcls[(size_t)(st[(size_t)i].pair0 + j)] + 1];
        for (size_t b = 1; b < cnt.size(); ++b)
            cnt[b] += cnt[b - 1];
        for (int j = 0; j < max_pairs; ++j) {
            rbase[(size_t)j] = cnt[(size_t)j * nb];
            for (int c = 0; c < nb; ++c)
                rseg[(size_t)j][(size_t)c] = (int32_t)(cnt[(size_t)j * nb + (size_t)c] - rbase[(size_t)j]);
            rseg[(size_t)j][kBatchClasses + 1] = (int32_t)(cnt[(size_t)(j + 1) * nb] - rbase[(size_t)j]);
        }
        rbase[(size_t)max_pairs] = cnt[(size_t)max_pairs * nb];
        int32_t *hl = reinterpret_cast<int32_t *>(E->h_rlist.p);
        for (int i = 0; i < num_docs; ++i)   // documents in increasing order within each bucket
            for (int j = 0; j < st[(size_t)i].npairs; ++j)
                hl[cnt[(size_t)j * nb + (size_t)pcls[(size_t)(st[(size_t)i].pair0 + j)]]++] = i;
    }
    BatchPair *hp = reinterpret_cast<BatchPair *>(E->h_pairs.p);
    uint8_t *hpool = E->h_pairs.p + (size_t)std::max<int64_t>(total_pairs, 1) * sizeof(BatchPair);
    size_t pool_fill = 0;
    for (int64_t q = 0; q < total_pairs; ++q) {
        BatchPair bp{};
        bp.L = pat_lens[q];
        bp.r = rep_lens[q];
        pattern_head(pats[q], bp.L, &bp.pat4, &bp.mask4);
        bp.cls = pcls[(size_t)q];
        bp.pat_off = (int32_t)pool_fill;
        if (bp.L > 0)
            memcpy(hpool + pool_fill, pats[q], (size_t)bp.L);
        pool_fill += align_up((size_t)bp.L + 16, 16);
        bp.rep_off = (int32_t)pool_fill;
        if (bp.r > 0)
            memcpy(hpool + pool_fill, reps[q], (size_t)bp.r);
        pool_fill += align_up((size_t)bp.r + 32, 16);
        hp[q] = bp;
    }
    BatchDocState *hs = reinterpret_cast<BatchDocState *>(E->h_states.p);
    for (int i = 0; i < num_docs; ++i) {
        const DocState &d = st[(size_t)i];
        hs[i] = BatchDocState{d.len, (int64_t)pieces[(size_t)i].dst_off, (int64_t)d.fin_off, (int32_t)d.pair0,
                              d.npairs};
    }
    BatchCtl *hc = reinterpret_cast<BatchCtl *>(E->h_ctl.p);
    memset(hc, 0, sizeof(BatchCtl));
    hc->fin_used = (int64_t)fin0;
    CK(stage_and_upload(E->h_in.p, E->arena[0].p, pieces, arena_in, s));
    if (total_pairs > 0)
        CK(hipMemcpyAsync(E->g_pairs.p, hp, (size_t)total_pairs * sizeof(BatchPair), hipMemcpyHostToDevice, s));
    if (pool_fill > 0)
        CK(hipMemcpyAsync(E->g_pool.p, hpool, pool_fill, hipMemcpyHostToDevice, s));
    if (num_docs > 0)
        CK(hipMemcpyAsync(E->g_states.p, hs, (size_t)num_docs * sizeof(BatchDocState), hipMemcpyHostToDevice, s));
    CK(hipMemcpyAsync(E->g_ctl.p, hc, sizeof(BatchCtl), hipMemcpyHostToDevice, s));
    if (total_pairs > 0)
        CK(hipMemcpyAsync(E->g_rlist.p, E->h_rlist.p, (size_t)total_pairs * sizeof(int32_t), hipMemcpyHostToDevice,
                          s));
    for (int i = 0; i < num_docs; ++i) {
        const DocState &d = st[(size_t)i];
        if (d.npairs == 0 && d.len > 0)
            CK(hipMemcpyAsync(E->fin.p + d.fin_off, E->arena[0].p + pieces[(size_t)i].dst_off, (size_t)d.len,
                              hipMemcpyDeviceToDevice, s));
    }

    // ---- rounds -------------------------------------------------------------------
    BatchRound base{};
    auto refresh = [&]() {   // buffer addresses and sizes (after any growth)
        for (int c = 0; c < kBatchClasses; ++c) {
            base.groups[c] = reinterpret_cast<BatchGroup *>(E->g_groups[c].p);
            base.gsum[c] = reinterpret_cast<uint32_t *>(E->g_gsum[c].p);
            base.pre[c] = reinterpret_cast<ChunkState *>(E->g_pre[c].p);
        }
        base.fin = E->fin.p;
        base.fin_cap = (int64_t)E->fin.cap;
        base.group_cap = group_cap;
    };
    base.ctl = reinterpret_cast<BatchCtl *>(E->g_ctl.p);
    base.states = reinterpret_cast<BatchDocState *>(E->g_states.p);
    base.pairs = reinterpret_cast<const BatchPair *>(E->g_pairs.p);
    base.pool = E->g_pool.p;
    base.docs = reinterpret_cast<BatchDoc *>(E->g_docs.p);
    base.doc_kept = reinterpret_cast<int64_t *>(E->g_kept.p);
    base.num_docs = num_docs;
    refresh();
    auto round_args = [&](int j, int gt) {
        BatchRound a = base;
        a.round = j;
        a.gt = gt;
        a.rlist = reinterpret_cast<const int32_t *>(E->g_rlist.p) + rbase[(size_t)j];
        a.rcount = (int32_t)(rbase[(size_t)j + 1] - rbase[(size_t)j]);
        for (int c = 0; c < kBatchClasses + 2; ++c)
            a.cseg[c] = rseg[(size_t)j][(size_t)c];
        a.src_arena = E->arena[j & 1].p;
        a.dst_arena = E->arena[(j + 1) & 1].p;
        a.dst_cap = (int64_t)E->arena[(j + 1) & 1].cap;
        return a;
    };
    const int grid_max = E->num_cus;
    const std::vector<std::vector<int64_t>> ub = length_bounds(st, max_pairs, pat_lens, rep_lens);

    int gt = env_int("CR_GROUP_TILES", 0);
    if (gt != 1 && gt != 2 && gt != 4)
        gt = 0;
    bool tuned_now = false;
    const auto tkey = std::make_pair(E->device, batch_bucket(arena_in));
    if (gt == 0) {
        std::lock_guard<std::mutex> lk(g_tune_mtx);
        auto it = g_tune.find(tkey);
        if (it != g_tune.end())
            gt = it->second.first;
    }
    if (gt == 0 && (env_int("CR_AUTOTUNE", 1) == 0 || tkey.second == 0 || max_pairs == 0))
        gt = kGroupTiles;
    if (gt == 0) {
        // round 0 once per candidate, each from the uploaded state
        const size_t sbytes = (size_t)num_docs * sizeof(BatchDocState);
        if (sbytes > 0)
            CK(hipMemcpyAsync(E->g_states0.p, E->g_states.p, sbytes, hipMemcpyDeviceToDevice, s));
        CK(hipMemcpyAsync(E->g_ctl0.p, E->g_ctl.p, sizeof(BatchCtl), hipMemcpyDeviceToDevice, s));
        std::vector<float> ms(3, 0.f);
        static const int kCands[3] = {1, 2, 4};
        for (int sweep = 0; sweep < 3; ++sweep) {   // interleaved, best time per candidate
            for (int ci = 0; ci < 3; ++ci) {
                if (sbytes > 0)
                    CK(hipMemcpyAsync(E->g_states.p, E->g_states0.p, sbytes, hipMemcpyDeviceToDevice, s));
                CK(hipMemcpyAsync(E->g_ctl.p, E->g_ctl0.p, sizeof(BatchCtl), hipMemcpyDeviceToDevice, s));
                CK(hipEventRecord(E->ev0, s));
                CK(launch_batch_round(round_args(0, kCands[ci]),
                                      round_shape(st, ub[0], pcls, 0, kCands[ci], grid_max), s));
                CK(hipEventRecord(E->ev1, s));
                CK(hipEventSynchronize(E->ev1));
                float t = 0.f;
                CK(hipEventElapsedTime(&t, E->ev0, E->ev1));
                ms[(size_t)ci] = sweep == 0 || t < ms[(size_t)ci] ? t : ms[(size_t)ci];
            }
        }
        gt = kGroupTiles;   // another size only when more than 3% faster
        for (int ci = 0; ci < 3; ++ci)
            if (ms[(size_t)ci] < 0.97f * ms[1] && ms[(size_t)ci] < ms[(size_t)(gt == 1 ? 0 : gt == 2 ? 1 : 2)])
                gt = kCands[ci];
        if (sbytes > 0)
            CK(hipMemcpyAsync(E->g_states.p, E->g_states0.p, sbytes, hipMemcpyDeviceToDevice, s));
        CK(hipMemcpyAsync(E->g_ctl.p, E->g_ctl0.p, sizeof(BatchCtl), hipMemcpyDeviceToDevice, s));
        std::lock_guard<std::mutex> lk(g_tune_mtx);
        g_tune[tkey] = std::make_pair(gt, ms);
        tuned_now = true;
    }

    std::vector<BatchRoundShape> shapes((size_t)max_pairs);
    std::vector<int64_t> key = {num_docs, max_pairs, gt, group_cap, (int64_t)E->fin.cap, (int64_t)E->arena[0].cap,
                                (int64_t)E->arena[1].cap};
    for (int j = 0; j < max_pairs; ++j) {
        shapes[(size_t)j] = round_shape(st, ub[(size_t)j], pcls, j, gt, grid_max);
        const BatchRoundShape &sh = shapes[(size_t)j];
        for (int c = 0; c <= kBatchClasses; ++c)
            key.push_back(sh.present[c] ? sh.class_docs[c] : 0);
        for (int c = 0; c < kBatchClasses; ++c)
            key.push_back(sh.work_grid[c]);
    }
    const void *ptrs[] = {E->arena[0].p, E->arena[1].p, E->fin.p, E->g_pairs.p, E->g_pool.p, E->g_states.p,
                          E->g_ctl.p, E->g_docs.p, E->g_kept.p};
    for (const void *q : ptrs)
        key.push_back((int64_t)(uintptr_t)q);
    for (int c = 0; c < kBatchClasses; ++c) {
        key.push_back((int64_t)(uintptr_t)E->g_groups[c].p);
        key.push_back((int64_t)(uintptr_t)E->g_gsum[c].p);
        key.push_back((int64_t)(uintptr_t)E->g_pre[c].p);
    }
    key.push_back((int64_t)(uintptr_t)E->g_rlist.p);
    for (int j = 0; j <= max_pairs; ++j)
        key.push_back(rbase[(size_t)j]);
    for (int j = 0; j < max_pairs; ++j)
        for (int c = 0; c < kBatchClasses + 2; ++c)
            key.push_back(rseg[(size_t)j][(size_t)c]);

    auto launch_rounds = [&]() -> hipError_t {
        for (int j = 0; j < max_pairs; ++j) {
            hipError_t e = launch_batch_round(round_args(j, gt), shapes[(size_t)j], s);
            if (e != hipSuccess)
                return e;
        }
        return hipSuccess;
    };
    // a shape is captured when it repeats (a one-off run pays no capture), then
    // replayed for as long as it keeps coming
    bool reused = false;
    const bool graph_ok = env_int("CR_BATCH_GRAPH", 1) != 0 && max_pairs > 0;
    const bool replay = graph_ok && E->graph_exec && key == E->graph_key;
    const bool use_graph = replay || (graph_ok && key == E->last_key);
    E->last_key = key;
    if (use_graph && !replay) {
        if (E->graph_exec)
            (void)hipGraphExecDestroy(E->graph_exec);
        if (E->graph)
            (void)hipGraphDestroy(E->graph);
        E->graph_exec = nullptr;
        E->graph = nullptr;
        E->graph_key.clear();
        CK(hipStreamSynchronize(s));
        CK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        const hipError_t le = launch_rounds();
        hipGraph_t g = nullptr;
        const hipError_t ee = hipStreamEndCapture(s, &g);
        if (le != hipSuccess || ee != hipSuccess) {
            if (g)
                (void)hipGraphDestroy(g);
            return (int)(le != hipSuccess ? le : ee);
        }
        hipGraphExec_t x = nullptr;
        hipError_t ie = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
        if (ie != hipSuccess) {
            (void)hipGraphDestroy(g);
            return (int)ie;
        }
        E->graph = g;
        E->graph_exec = x;
        E->graph_key = key;
    } else if (use_graph) {
        reused = true;
    }
    CK(hipEventRecord(E->ev0, s));
    if (use_graph)
        CK(hipGraphLaunch(E->graph_exec, s));
    else
        CK(launch_rounds());
    CK(hipEventRecord(E->ev1, s));

    // ---- outcome: a round that did not fit is resumed with grown buffers ----------------
    int resumes = 0;
    for (;; ++resumes) {
        CK(hipMemcpyAsync(hc, E->g_ctl.p, sizeof(BatchCtl), hipMemcpyDeviceToHost, s));
        CK(hipStreamSynchronize(s));
        if (!hc->overflow)
            break;
        if (resumes >= 32)
            return (int)hipErrorOutOfMemory;
        const int j = hc->overflow - 1;
        if (hc->need_groups > group_cap) {
            group_cap = hc->need_groups + hc->need_groups / 4 + num_docs + 1;
            CK(reserve_lists());
        }
        DevBuf &dst = E->arena[(j + 1) & 1];   // round j's destination: nothing live in it
        if ((size_t)hc->need_arena > dst.cap)
            CK(dev_reserve(dst, (size_t)hc->need_arena + (size_t)hc->need_arena / 4 + kGranule, false, 0, s));
        if ((size_t)hc->need_fin > E->fin.cap)
            CK(dev_reserve(E->fin, (size_t)hc->need_fin + (size_t)hc->need_fin / 4 + kGranule, true,
                           (size_t)hc->fin_used, s));
        refresh();
        hc->overflow = 0;
        CK(hipMemcpyAsync(E->g_ctl.p, hc, sizeof(BatchCtl), hipMemcpyHostToDevice, s));
        for (int jj = j; jj < max_pairs; ++jj)
            CK(launch_batch_round(round_args(jj, gt), shapes[(size_t)jj], s));
        CK(hipEventRecord(E->ev1, s));
    }
    E->hint_arena = std::max(E->arena[0].cap, E->arena[1].cap);
    E->hint_fin = E->fin.cap - fin0;
    if (num_docs > 0)
        CK(hipMemcpyAsync(hs, E->g_states.p, (size_t)num_docs * sizeof(BatchDocState), hipMemcpyDeviceToHost, s));
    const size_t fin_used = (size_t)hc->fin_used;
    CK(host_reserve(E->h_out, fin_used + 64));
    if (fin_used > 0)
        CK(hipMemcpyAsync(E->h_out.p, E->fin.p, fin_used, hipMemcpyDeviceToHost, s));
    CK(hipStreamSynchronize(s));
    for (int i = 0; i < num_docs; ++i) {
        out_offsets[i] = (size_t)hs[i].fin_off;
        out_lens[i] = (size_t)hs[i].len;
    }
    *out_base = E->h_out.p;
    if (stats) {
        float ms = 0.f;
        CK(hipEventElapsedTime(&ms, E->ev0, E->ev1));
        stats[0] = ms;
        stats[1] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
        stats[2] = (double)hc->algo_bytes;
        stats[3] = max_pairs;
    }
    E->info[0] = 2;
    E->info[1] = gt;
    E->info[2] = reused ? 1 : 0;
    E->info[3] = tuned_now ? 1 : 0;
    E->info[4] = max_pairs;
    E->info[5] = use_graph ? 1 : 0;
    E->info[6] = resumes;
    return 0;
}

void engine_free(Engine *E)
{
    if (E->graph_exec)
        (void)hipGraphExecDestroy(E->graph_exec);
    if (E->graph)
        (void)hipGraphDestroy(E->graph);
    DevBuf *dev[] = {&E->arena[0], &E->arena[1], &E->fin, &E->d_docs, &E->d_groups, &E->d_cdocs,
                     &E->d_pool, &E->d_gsum, &E->d_pre, &E->d_kept, &E->d_other, &E->g_pairs, &E->g_pool,
                     &E->g_states, &E->g_states0, &E->g_ctl, &E->g_ctl0, &E->g_docs, &E->g_kept};
    for (DevBuf *b : dev)
        if (b->p)
            (void)hipFree(b->p);
    for (int c = 0; c < kBatchClasses; ++c) {
        DevBuf *per[] = {&E->g_groups[c], &E->g_gsum[c], &E->g_pre[c]};
        for (DevBuf *b : per)
            if (b->p)
                (void)hipFree(b->p);
    }
    if (E->g_rlist.p)
        (void)hipFree(E->g_rlist.p);
    HostBuf *host[] = {&E->h_in, &E->h_docs[0], &E->h_docs[1], &E->h_groups[0], &E->h_groups[1],
                       &E->h_cdocs[0], &E->h_cdocs[1], &E->h_pool[0], &E->h_pool[1], &E->h_kept,
                       &E->h_out, &E->h_misc, &E->h_pairs, &E->h_states, &E->h_ctl, &E->h_rlist};
    for (HostBuf *b : host)
        if (b->p)
            (void)hipHostFree(b->p);
    if (E->ev0)
        (void)hipEventDestroy(E->ev0);
    if (E->ev1)
        (void)hipEventDestroy(E->ev1);
    if (E->stream)
        (void)hipStreamDestroy(E->stream);
}

} // namespace

extern "C" {

CR_API int cr_batch_create(void **pb)
{
    if (!pb)
        return -1;
    *pb = nullptr;
    Engine *E = new (std::nothrow) Engine();
    if (!E)
        return -2;
    hipError_t e = hipGetDevice(&E->device);
    if (e == hipSuccess)
        e = hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking);
    if (e == hipSuccess)
        e = hipEventCreate(&E->ev0);
    if (e == hipSuccess)
        e = hipEventCreate(&E->ev1);
    if (e == hipSuccess) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, E->device) == hipSuccess && cus > 0)
            E->num_cus = cus;
    }
    if (e != hipSuccess) {
        engine_free(E);
        delete E;
        return (int)e;
    }
    const int m = env_int("CR_BATCH_DEVICE", 1);
    E->mode = m == 0 ? 1 : 0;
    *pb = E;
    return 0;
}

CR_API int cr_batch_run(void *b, int num_docs, const uint8_t *const *docs, const size_t *doc_lens,
                        const int *pair_counts, const uint8_t *const *pats, const int *pat_lens,
                        const uint8_t *const *reps, const int *rep_lens, const uint8_t **out_base,
                        size_t *out_offsets, size_t *out_lens, double *stats)
{
    if (!b)
        return -1;
    Engine *E = static_cast<Engine *>(b);
    std::lock_guard<std::mutex> lock(E->mtx);
    DeviceGuard dg(E->device);
    if (E->mode != 1 && valid_run(num_docs, docs, doc_lens, pair_counts, pats, pat_lens, reps, rep_lens, out_base,
                                  out_offsets, out_lens)) {
        const int rc = engine_run_device(E, num_docs, docs, doc_lens, pair_counts, pats, pat_lens, reps, rep_lens,
                                         out_base, out_offsets, out_lens, stats);
        if (rc != kFallback)
            return rc;
    }
    const int rc = engine_run(E, num_docs, docs, doc_lens, pair_counts, pats, pat_lens, reps, rep_lens, out_base,
                              out_offsets, out_lens, stats);
    if (rc == 0) {
        E->info[0] = 1;
        E->info[1] = kGroupTiles;
        E->info[2] = E->info[3] = E->info[5] = E->info[6] = 0;
        E->info[4] = stats ? stats[3] : 0;
    }
    return rc;
}

// mode: 0 automatic (device-driven rounds whenever no pair needs the bitmap path),
// 1 host-driven rounds only.
CR_API int cr_batch_set_mode(void *b, int mode)
{
    if (!b || (mode != 0 && mode != 1))
        return -1;
    Engine *E = static_cast<Engine *>(b);
    std::lock_guard<std::mutex> lock(E->mtx);
    E->mode = mode;
    return 0;
}

// The last run: info[0] path (1 host-driven rounds, 2 device-driven), [1] tiles per
// group, [2] HIP graph replayed without capture, [3] group size measured in this
// run, [4] rounds, [5] run as a HIP graph, [6] rounds resumed after growing the
// buffers.  n: entries wanted (up to 7).
CR_API int cr_batch_info(void *b, double *info, int n)
{
    if (!b || !info || n < 0)
        return -1;
    Engine *E = static_cast<Engine *>(b);
    std::lock_guard<std::mutex> lock(E->mtx);
    for (int i = 0; i < n && i < 7; ++i)
        info[i] = E->info[i];
    return 0;
}

CR_API void cr_batch_destroy(void *b)
{
    if (!b)
        return;
    Engine *E = static_cast<Engine *>(b);
    {
        std::lock_guard<std::mutex> lock(E->mtx);
        DeviceGuard dg(E->device);
        if (E->stream)
            (void)hipStreamSynchronize(E->stream);
        engine_free(E);
    }
    delete E;
}

} // extern "C"

namespace cr {

int batch_live(void *eng, const std::function<bool(std::vector<LiveReq> &, bool)> &admit,
               const std::function<void(void *, int, const uint8_t *, size_t)> &finish)
{
    if (!eng)
        return -1;
    Engine *E = static_cast<Engine *>(eng);
    std::lock_guard<std::mutex> lock(E->mtx);
    DeviceGuard dg(E->device);
    return engine_live(E, admit, finish);
}

} // namespace cr
