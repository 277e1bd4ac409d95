// cr_runtime.hip -- host runtime and C ABI of the gfx950 byte-pattern replace library.
//
// A session keeps the buffer resident in HBM as a ping-pong pair (d_buf holds the
// current contents, d_tmp receives the next pass) plus the scratch of the chunked
// kernels: one summary per 16 KiB group, one state per chunk, per-tile records and a
// small stats block.  Every apply is: upload pattern + replacement (pinned staging,
// async), a count and an emit launch per chunk (mark + greedy + emit for the bitmap
// fallback), one 32-byte stats readback, swap.
//
// Mirrors the reference's session model (cuda_replace.cu:139-216 CRContext,
// 997-1140 ensure_* allocators, 2097-2159 apply_once, 2161-2603 exports) with
// 64-bit lengths and no per-call device allocation of pattern buffers.
#include "cr_internal.h"
#include "cr_pattern.h"
#include "cr_host.h"
#include "../../include/cuda_replace.h"

#include <limits.h>
#include <stdio.h>
#include <mutex>
#include <new>
#include <stdlib.h>
#include <string.h>
#include <vector>
#include <algorithm>
#include <chrono>
#include <map>
#include <string>
#include <tuple>

using namespace cr;

namespace {

constexpr size_t kAllocGranule = 2u << 20;   // device buffers grow in 2 MiB steps

// ---- autotuner: tiles per group ------------------------------------------------
// The group size (1, 2 or 4 tiles of 8 KiB per workgroup) trades the spine's work
// and the per-workgroup fixed costs against LDS per workgroup (occupancy) and tail
// balance.  The best one is measured, not assumed: the first pass of a pattern class
// (border-free, carry tables of 2/4/8/16 states) in an input size bucket on a device
// runs count + spine + emit per candidate after its normal pass (three interleaved
// sweeps, HIP events, best time per candidate) and records the fastest for every
// later pass in the bucket; another size replaces the default 2 only when it is
// more than 3% faster.  Inputs under 4 MiB are not tuned.  CR_GROUP_TILES=1|2|4 fixes the choice;
// CR_AUTOTUNE=0 keeps the default.
constexpr int kTuneCands[3] = {1, 2, 4};

// Border-free passes have more forms than a group size; the tuner measures them all:
// the single persistent launch (64 source bytes per thread, 16 KiB groups, counts
// two rounds ahead of the emit), the same kernels as three launches (the emit
// recomputes candidates and takes its carry from a left halo), and the 32-byte-per-
// thread count and emit with 1, 2 or 4 tiles per group (the count hands every
// segment's candidate mask to the emit).
enum BfVariant { kBfSweep = 0, kBfLean = 1, kBfTiles1 = 2, kBfTiles2 = 3, kBfTiles4 = 4, kBfKeep = 5, kBfVariants = 6 };
constexpr int kBfVariantTiles[kBfVariants] = {kGroupTiles, kGroupTiles, 1, 2, 4, kGroupTiles};
// the single-launch forms (kBfSweep re-stages a group for its emit, kBfKeep holds it in LDS)
constexpr bool single_launch(int v) { return v == kBfSweep || v == kBfKeep; }

struct TuneEntry {
    int gt = 0;
    float ms[3] = {0, 0, 0};   // per candidate, best of three
    int single = -1;           // 2-state tables: 1 the single launch, 0 the three launches, -1 not timed
    float single_ms = -1.f;    // ... the single launch's best time (1e30: it gave up)
    int variant = -1;          // border-free class: the chosen BfVariant
    float vms[kBfVariants] = {0, 0, 0, 0, 0, 0};
};

struct Tuner {
    std::mutex mtx;
    std::map<std::tuple<int, int, int>, TuneEntry> best;   // (device, class, bucket)
};

Tuner &tuner()
{
    static Tuner t;
    return t;
}

int tune_class(int mode, int lt)
{
    if (mode == kModeBorderFree)
        return 0;
    if (mode == kModeTable)
        return lt == 2 ? 1 : lt == 4 ? 2 : lt == 8 ? 3 : 4;
    return -1;
}

// bucket 0 (under 4 MiB) is never measured, only pinned (cr_autotune_set)
int size_bucket(int64_t n)
{
    if (n < (4ll << 20))
        return 0;
    if (n < (32ll << 20))
        return 1;
    if (n < (256ll << 20))
        return 2;
    if (n < (2048ll << 20))
        return 3;
    return 4;
}

int forced_group_tiles()
{
    static const int v = [] {
        const char *e = getenv("CR_GROUP_TILES");
        const int g = e ? atoi(e) : 0;
        return g == 1 || g == 2 || g == 4 ? g : 0;
    }();
    return v;
}

bool autotune_enabled()
{
    static const bool v = [] {
        const char *e = getenv("CR_AUTOTUNE");
        return !(e && e[0] == '0');
    }();
    return v;
}

// Group size for a pass; *tune is set when this pass should measure the candidates.
int choose_group_tiles(int device, int mode, int lt, int64_t n, bool *tune)
{
    *tune = false;
    if (mode == kModeBitmap)
        return kGroupTiles;
    if (const int f = forced_group_tiles())
        return f;
    const int c = tune_class(mode, lt), b = size_bucket(n);
    if (c < 0)
        return kGroupTiles;
    Tuner &t = tuner();
    std::lock_guard<std::mutex> lk(t.mtx);
    auto it = t.best.find(std::make_tuple(device, c, b));
    if (it != t.best.end() && it->second.gt > 0)
        return it->second.gt;
    *tune = b > 0 && autotune_enabled();
    return kGroupTiles;
}

void record_tuning(int device, int mode, int lt, int64_t n, const TuneEntry &e)
{
    Tuner &t = tuner();
    std::lock_guard<std::mutex> lk(t.mtx);
    t.best[std::make_tuple(device, tune_class(mode, lt), size_bucket(n))] = e;
}

int64_t groups_of(int64_t num_tiles, int gt) { return (num_tiles + gt - 1) / gt; }

struct Session {
    std::mutex mtx;
    bool alive = true;
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;

    uint8_t *d_buf = nullptr;
    size_t cap_buf = 0;
    uint8_t *d_tmp = nullptr;
    size_t cap_tmp = 0;
    size_t len = 0;

    uint32_t *d_gsum = nullptr;      // group summaries (count or carry table per group)
    size_t gsum_words = 0;
    ChunkState *d_chunks = nullptr;  // state after each chunk (or entering each spine block)
    size_t chunk_cap = 0;
    uint64_t *d_blocks = nullptr;    // spine block summaries
    size_t blocks_cap = 0;
    ChunkState *d_pre = nullptr;     // state before each group
    size_t pre_cap = 0;
    TileInfo *d_info = nullptr;
    size_t info_tiles = 0;
    uint32_t *d_bitmap = nullptr;
    size_t bitmap_words = 0;
    uint64_t *d_sweep = nullptr;     // single-pass sweep: per-group counts and prefixes, abort word
    size_t sweep_words = 0;
    uint8_t *d_bmsel = nullptr;      // parallel bitmap selection: tile tables, levels, entries
    size_t bmsel_cap = 0;
    uint8_t *d_patrep = nullptr;
    size_t patrep_cap = 0;
    uint8_t *h_patrep = nullptr;     // pinned staging for pattern + replacement
    size_t h_patrep_cap = 0;
    uint64_t *d_presence = nullptr;
    size_t presence_words = 0;

    // small device block: stats and query output
    uint8_t *d_misc = nullptr;
    uint8_t *h_misc = nullptr;       // pinned mirror
    LaunchStats *d_stats = nullptr;
    int64_t *d_query = nullptr;
    uint32_t epoch = 0;

    // streaming pipeline (cr_stream_replace): two chunk slots in and out on the
    // device, pinned staging for pageable input, upload / download streams
    uint8_t *d_sin[2] = {nullptr, nullptr};
    size_t cap_sin[2] = {0, 0};
    uint8_t *d_sout[2] = {nullptr, nullptr};
    size_t cap_sout[2] = {0, 0};
    uint8_t *h_sin[2] = {nullptr, nullptr};
    size_t cap_hsin[2] = {0, 0};
    hipStream_t cs_in = nullptr, cs_out = nullptr;
    hipEvent_t ev_in[2] = {nullptr, nullptr}, ev_out[2] = {nullptr, nullptr};

    // fused apply_batch: device table; the pairs and the batch's source (kept in
    // d_tmp) let a later prefix query replay the pairs one by one
    FusedTable *d_fused = nullptr;
    FusedTable *h_fused = nullptr;   // pinned staging
    bool fused_pending = false;
    std::vector<std::vector<uint8_t>> fused_pats, fused_reps;
    size_t fused_src_len = 0;
    // the last pair list offered to apply_batch (lengths and bytes) and what the
    // fusability check said of it: 1 fusable (h_fused and d_fused hold its table),
    // 2 not fusable, 0 nothing cached
    std::vector<uint8_t> fused_key;
    int fused_key_state = 0;
    int fused_max_grow = 0;   // largest r - L of the cached fusable list
    uint8_t *d_q[2] = {nullptr, nullptr};
    size_t cap_q[2] = {0, 0};

    // record of the most recent apply, for prefix queries
    bool q_valid = false;
    int q_mode = 0;
    ReplaceParams q_params{};
};

struct DeviceGuard {
    int prev = -1;
    bool changed = false;
    explicit DeviceGuard(int dev)
    {
        if (hipGetDevice(&prev) == hipSuccess && prev != dev) {
            changed = hipSetDevice(dev) == hipSuccess;
        }
    }
    ~DeviceGuard()
    {
        if (changed)
            (void)hipSetDevice(prev);
    }
};

#define CK(expr)                                  \
    do {                                          \
        hipError_t e_ = (expr);                   \
        if (e_ != hipSuccess)                     \
            return (int)e_;                       \
    } while (0)

size_t round_up(size_t v, size_t g) { return (v + g - 1) / g * g; }

// grow a device allocation without preserving contents
// Test hook: CR_FAIL_GROW_OVER=bytes makes every device buffer growth to more than
// that many bytes fail as an out-of-memory hipMalloc would (read at each growth), so
// tests can check that a failed growth leaves the session's data and buffers as
// they were.
static size_t fail_grow_over()
{
    const char *e = getenv("CR_FAIL_GROW_OVER");
    return e ? (size_t)strtoull(e, nullptr, 10) : SIZE_MAX;
}

hipError_t grow(void **ptr, size_t *cap, size_t need_bytes, bool zero, hipStream_t s)
{
    if (*cap >= need_bytes && *ptr)
        return hipSuccess;
    if (need_bytes > fail_grow_over())
        return hipErrorOutOfMemory;
    size_t nc = round_up(need_bytes > 0 ? need_bytes : 1, kAllocGranule);
    if (*cap > 0 && nc < *cap + *cap / 2)
        nc = round_up(*cap + *cap / 2, kAllocGranule);
    void *np = nullptr;
    hipError_t e = hipMalloc(&np, nc);
    if (e != hipSuccess) {
        // retry with the exact size before giving up
        (void)hipGetLastError();
        nc = round_up(need_bytes > 0 ? need_bytes : 1, kAllocGranule);
        e = hipMalloc(&np, nc);
        if (e != hipSuccess)
            return e;
    }
    if (zero) {
        e = hipMemsetAsync(np, 0, nc, s);
        if (e != hipSuccess) {
            (void)hipFree(np);
            return e;
        }
    }
    if (*ptr) {
        // the old buffer may still be read by queued work on this stream
        hipError_t se = hipStreamSynchronize(s);
        (void)hipFree(*ptr);
        if (se != hipSuccess) {
            *ptr = np;
            *cap = nc;
            return se;
        }
    }
    *ptr = np;
    *cap = nc;
    return hipSuccess;
}

hipError_t grow_host(uint8_t **ptr, size_t *cap, size_t need)
{
    if (*cap >= need && *ptr)
        return hipSuccess;
    size_t nc = round_up(need > 0 ? need : 1, 4096);
    void *np = nullptr;
    hipError_t e = hipHostMalloc(&np, nc, hipHostMallocDefault);
    if (e != hipSuccess)
        return e;
    if (*ptr)
        (void)hipHostFree(*ptr);
    *ptr = (uint8_t *)np;
    *cap = nc;
    return hipSuccess;
}

hipError_t ensure_misc(Session *s)
{
    if (s->d_misc)
        return hipSuccess;
    const size_t bytes = 256;
    hipError_t e = hipMalloc((void **)&s->d_misc, bytes);
    if (e != hipSuccess)
        return e;
    e = hipMemsetAsync(s->d_misc, 0, bytes, s->stream);
    if (e != hipSuccess)
        return e;
    e = hipHostMalloc((void **)&s->h_misc, bytes, hipHostMallocDefault);
    if (e != hipSuccess)
        return e;
    s->d_stats = (LaunchStats *)(s->d_misc + 64);
    s->d_query = (int64_t *)(s->d_misc + 128);
    return hipStreamSynchronize(s->stream);
}

// next epoch; clear the stats words when the 16-bit epoch wraps
hipError_t next_epoch(Session *s)
{
    if (s->epoch >= 0xffffu) {
        hipError_t e = launch_fill_u64(reinterpret_cast<uint64_t *>(s->d_stats), 0,
                                       sizeof(LaunchStats) / sizeof(uint64_t), s->stream);
        if (e != hipSuccess)
            return e;
        // the sweep's tagged words too, so no word of an earlier round of epochs matches
        e = launch_fill_u64(s->d_sweep, 0, s->sweep_words, s->stream);
        if (e != hipSuccess)
            return e;
        s->epoch = 0;
    }
    ++s->epoch;
    return hipSuccess;
}

// Groups per chunk.  By default one chunk holds the whole pass (count, spine and
// emit each launch once); CR_CHUNK_GROUPS=g splits the pass into chunks of g groups.
int64_t chunk_groups()
{
    static const int64_t v = [] {
        const char *e = getenv("CR_CHUNK_GROUPS");
        long long g = e ? atoll(e) : 0;
        return (int64_t)(g > 0 ? g : INT64_MAX / 4);
    }();
    return v;
}

// Single-pass sweep for border-free passes of at least CR_SWEEP_MIN bytes (default
// 64 MiB; CR_SWEEP=0 turns it off); CR_SWEEP_AHEAD rounds of count-ahead (default 2).
// Read at every pass, so a test can force either path.
// Bitmap-mode passes select kept starts in parallel (CR_BITMAP_SERIAL=1 forces the
// one-wave serial greedy, for comparison).  Read at every pass.
bool bitmap_parallel_enabled()
{
    const char *e = getenv("CR_BITMAP_SERIAL");
    return !(e && atoi(e) != 0);
}

int64_t sweep_min_bytes()
{
    const char *off = getenv("CR_SWEEP");
    if (off && atoi(off) == 0)
        return INT64_MAX;
    const char *e = getenv("CR_SWEEP_MIN");
    return e ? (int64_t)atoll(e) : (64ll << 20);
}

// 2-state carry-table passes of this many bytes or more run as one persistent launch
// (CR_TABLE_SWEEP=0 keeps them on the three launches, CR_TABLE_SWEEP_MIN moves the
// threshold; read per pass)
int64_t table_sweep_min_bytes()
{
    const char *off = getenv("CR_TABLE_SWEEP");
    if (off && atoi(off) == 0)
        return INT64_MAX;
    const char *e = getenv("CR_TABLE_SWEEP_MIN");
    return e ? (int64_t)atoll(e) : (64ll << 20);
}

// Whether a 2-state carry-table pass of n bytes runs as the single launch: never
// with CR_TABLE_SWEEP=0, by the threshold alone when CR_TABLE_SWEEP_MIN is set, else
// the tuned choice for the size bucket, else from 64 MiB.
bool table_single_chosen(int device, int64_t n)
{
    const int64_t smin = table_sweep_min_bytes();
    if (smin == INT64_MAX)
        return false;
    if (getenv("CR_TABLE_SWEEP_MIN"))
        return n >= smin;
    {
        Tuner &t = tuner();
        std::lock_guard<std::mutex> lk(t.mtx);
        auto it = t.best.find(std::make_tuple(device, tune_class(kModeTable, 2), size_bucket(n)));
        if (it != t.best.end() && it->second.single >= 0)
            return it->second.single == 1;
    }
    return n >= smin;
}

// Fused multi-pattern passes of this many bytes or more run as one persistent launch
// (CR_FUSED_SWEEP=0 keeps them on three launches; read per pass)
int64_t fused_sweep_min_bytes()
{
    const char *off = getenv("CR_FUSED_SWEEP");
    if (off && atoi(off) == 0)
        return INT64_MAX;
    const char *e = getenv("CR_FUSED_SWEEP_MIN");
    return e ? (int64_t)atoll(e) : (64ll << 20);
}

int sweep_ahead()
{
    const char *e = getenv("CR_SWEEP_AHEAD");
    const int v = e ? atoi(e) : 2;
    return v < 0 ? 0 : (v > 64 ? 64 : v);
}

int sweep_spin_limit()
{
    const char *e = getenv("CR_SWEEP_SPIN");
    const int v = e ? atoi(e) : 200000;
    return v < 1 ? 1 : v;
}

// The untuned form of a border-free pass of n bytes: from CR_SWEEP_MIN (64 MiB) the
// single launch holding groups in LDS, below it the three launches of the
// re-staging launch's kernels.
int default_bf_variant(int64_t n)
{
    return n < sweep_min_bytes() ? kBfLean : kBfKeep;
}

// The form of a border-free pass (BfVariant); *tune is set when this pass should
// measure every form.  CR_BF_VARIANT=0..5 fixes it; CR_SWEEP=0 keeps the single
// launches out; otherwise the tuned choice for the size bucket, or the default.
int choose_bf_variant(int device, int64_t n, bool *tune)
{
    *tune = false;
    const int64_t smin = sweep_min_bytes();
    const bool sweep_ok = smin != INT64_MAX;
    if (const char *e = getenv("CR_BF_VARIANT")) {
        const int v = atoi(e);
        if (v >= 0 && v < kBfVariants)
            return single_launch(v) && !sweep_ok ? kBfLean : v;
    }
    const int dflt = default_bf_variant(n);
    const int b = size_bucket(n);
    Tuner &t = tuner();
    std::lock_guard<std::mutex> lk(t.mtx);
    auto it = t.best.find(std::make_tuple(device, 0, b));
    if (it != t.best.end() && it->second.variant >= 0)
        return single_launch(it->second.variant) && !sweep_ok ? kBfLean : it->second.variant;
    *tune = b > 0 && autotune_enabled();
    return dflt;
}

// The sweep's words, cleared when (re)allocated.
hipError_t ensure_sweep(Session *s, int64_t groups)
{
    const size_t need = (size_t)sweep_words(groups);
    if (s->sweep_words >= need && s->d_sweep)
        return hipSuccess;
    size_t cap_bytes = s->sweep_words * 8;
    hipError_t e = grow((void **)&s->d_sweep, &cap_bytes, need * 8, false, s->stream);
    if (e != hipSuccess)
        return e;
    s->sweep_words = cap_bytes / 8;
    return launch_fill_u64(s->d_sweep, 0, s->sweep_words, s->stream);
}

hipError_t ensure_scratch(Session *s, int64_t num_tiles, int mode, int table_size)
{
    hipError_t e;
    // one word (border-free) or one table of table_size words per group, sized for
    // the smallest group (one tile) so every group size fits
    const int64_t groups = num_tiles;
    const size_t per = mode == kModeTable ? (size_t)table_size : 1u;
    const size_t words = per * (size_t)groups;
    if (s->gsum_words < words) {
        size_t cap_bytes = s->gsum_words * 4;
        e = grow((void **)&s->d_gsum, &cap_bytes, words * 4, false, s->stream);
        if (e != hipSuccess)
            return e;
        s->gsum_words = cap_bytes / 4;
    }
    const size_t chunks = std::max((size_t)((groups + chunk_groups() - 1) / chunk_groups()) + 1,
                                   (size_t)spine_blocks_needed(groups) + 1);
    if (s->blocks_cap < chunks) {
        size_t cap_bytes = s->blocks_cap * kMaxTablePat * sizeof(uint64_t);
        e = grow((void **)&s->d_blocks, &cap_bytes, chunks * kMaxTablePat * sizeof(uint64_t), false, s->stream);
        if (e != hipSuccess)
            return e;
        s->blocks_cap = cap_bytes / (kMaxTablePat * sizeof(uint64_t));
    }
    if (s->chunk_cap < chunks) {
        size_t cap_bytes = s->chunk_cap * sizeof(ChunkState);
        e = grow((void **)&s->d_chunks, &cap_bytes, chunks * sizeof(ChunkState), false, s->stream);
        if (e != hipSuccess)
            return e;
        s->chunk_cap = cap_bytes / sizeof(ChunkState);
    }
    if (s->pre_cap < (size_t)groups) {
        size_t cap_bytes = s->pre_cap * sizeof(ChunkState);
        e = grow((void **)&s->d_pre, &cap_bytes, (size_t)groups * sizeof(ChunkState), false, s->stream);
        if (e != hipSuccess)
            return e;
        s->pre_cap = cap_bytes / sizeof(ChunkState);
    }
    if (s->info_tiles < (size_t)num_tiles + 1) {
        size_t cap_bytes = s->info_tiles * sizeof(TileInfo);
        e = grow((void **)&s->d_info, &cap_bytes, ((size_t)num_tiles + 1) * sizeof(TileInfo), false,
                 s->stream);
        if (e != hipSuccess)
            return e;
        s->info_tiles = cap_bytes / sizeof(TileInfo);
    }
    // one word per 32-byte segment: the bitmap of a bitmap-mode pass, or the
    // candidate masks the count hands to the emit (optional: without them the emit
    // recomputes the masks, so a failed allocation only costs time)
    size_t need = (size_t)num_tiles * (kTile / 32);
    if (s->bitmap_words < need) {
        size_t cap_bytes = s->bitmap_words * 4;
        e = grow((void **)&s->d_bitmap, &cap_bytes, need * 4, false, s->stream);
        if (e != hipSuccess) {
            if (mode == kModeBitmap)
                return e;
            (void)hipGetLastError();
            return hipSuccess;
        }
        s->bitmap_words = cap_bytes / 4;
    }
    return hipSuccess;
}

hipError_t upload_pattern(Session *s, const uint8_t *pat, int L, const uint8_t *rep, int r)
{
    const size_t need = (size_t)L + (size_t)r + 16;
    hipError_t e = grow((void **)&s->d_patrep, &s->patrep_cap, need, false, s->stream);
    if (e != hipSuccess)
        return e;
    e = grow_host(&s->h_patrep, &s->h_patrep_cap, need);
    if (e != hipSuccess)
        return e;
    if (L > 0)
        memcpy(s->h_patrep, pat, (size_t)L);
    if (r > 0)
        memcpy(s->h_patrep + L, rep, (size_t)r);
    if (L + r > 0)
        return hipMemcpyAsync(s->d_patrep, s->h_patrep, (size_t)L + (size_t)r, hipMemcpyHostToDevice,
                              s->stream);
    return hipSuccess;
}

int64_t floor_from_seed(long long seed, int L)
{
    if (seed > (long long)(INT64_MAX - L))
        return INT64_MAX;
    const int64_t f = (int64_t)seed + L;
    return f < 0 ? 0 : f;
}

struct PassResult {
    int64_t kept = 0;
    int64_t last = -1;
    int64_t out_len = 0;
};

// Core pass: d_src (n bytes) -> d_dst.  dst_cap_fn(needed) must make d_dst hold at
// least `needed` bytes and return its address (or nullptr with an error).
// Core pass: d_src (n bytes) -> destination from dst_for(needed, &ptr).
// fixed_cap < 0: the destination can grow to any size; otherwise its capacity.
// want_last: also compute the last kept start (one extra small kernel).
template <class DstFn>
int run_pass(Session *s, const uint8_t *d_src, int64_t n, const uint8_t *pat, int L, const uint8_t *rep,
             int r, int64_t floor_pos, int64_t fixed_cap, bool want_last, DstFn &&dst_for, PassResult *res,
             float *kernel_ms, int *mode_out, ReplaceParams *params_out)
{
    CK(ensure_misc(s));
    int table_size = 0;
    int mode = pattern_mode(pat, L, &table_size);

    if (mode == kModeEmpty) {
        if (mode_out)
            *mode_out = mode;
        const int64_t out = n + (n + 1) * (int64_t)r;
        uint8_t *dst = nullptr;
        int rc = dst_for(out, &dst);
        if (rc != 0) {
            res->out_len = out;
            return rc;
        }
        res->kept = n + 1;
        res->last = n;
        res->out_len = out;
        if (kernel_ms)
            CK(hipEventRecord(s->ev0, s->stream));
        if (r == 0) {
            if (n > 0 && dst != d_src)
                CK(hipMemcpyAsync(dst, d_src, (size_t)n, hipMemcpyDeviceToDevice, s->stream));
        } else {
            CK(upload_pattern(s, pat, 0, rep, r));
            ReplaceParams p{};
            p.src = d_src;
            p.dst = dst;
            p.n = n;
            p.rep = s->d_patrep;
            p.r = r;
            CK(launch_empty_pattern(p, s->stream));
        }
        if (kernel_ms)
            CK(hipEventRecord(s->ev1, s->stream));
        CK(hipStreamSynchronize(s->stream));
        if (kernel_ms)
            CK(hipEventElapsedTime(kernel_ms, s->ev0, s->ev1));
        return 0;
    }

    if (n == 0) {
        if (mode_out)
            *mode_out = mode;
        uint8_t *dst = nullptr;
        int rc = dst_for(0, &dst);
        if (rc != 0)
            return rc;
        res->kept = 0;
        res->last = -1;
        res->out_len = 0;
        if (kernel_ms)
            *kernel_ms = 0.f;
        return 0;
    }

    CK(upload_pattern(s, pat, L, rep, r));
    const int64_t num_tiles = (n + kTile - 1) / kTile;
    bool tune = false;
    // border-free patterns: the form the tuner chose (its group size follows);
    // other patterns: the tuned group size
    const int bf_var = mode == kModeBorderFree ? choose_bf_variant(s->device, n, &tune) : -1;
    const bool bf_lean = single_launch(bf_var) || bf_var == kBfLean;   // 64 bytes per thread
    const int gt = mode == kModeBorderFree ? kBfVariantTiles[bf_var]
                                           : choose_group_tiles(s->device, mode, table_size, n, &tune);
    const int64_t num_groups = groups_of(num_tiles, gt);
    const int64_t cg = chunk_groups();
    const int64_t num_chunks = (num_groups + cg - 1) / cg;

    ReplaceParams p{};
    p.src = d_src;
    p.n = n;
    p.floor = floor_pos;
    p.out_len = -1;
    p.pat = s->d_patrep;
    p.rep = s->d_patrep + L;
    p.L = L;
    p.r = r;
    uint32_t pat4 = 0, mask4 = 0;
    for (int i = 0; i < 4 && i < L; ++i) {
        pat4 |= (uint32_t)pat[i] << (8 * i);
        mask4 |= 0xffu << (8 * i);
    }
    p.pat4 = pat4;
    p.mask4 = mask4;
    pattern_qsad(pat, L, p.qref, &p.qx);
    p.stats = s->d_stats;
    p.num_tiles = num_tiles;
    p.num_groups = num_groups;

    // read the pass's stats block (kept count, and the bitmap greedy's last kept)
    auto read_stats = [&](int64_t *kept, int64_t *last) -> int {
        hipError_t e = hipMemcpyAsync(s->h_misc + 64, s->d_stats, sizeof(LaunchStats), hipMemcpyDeviceToHost,
                                      s->stream);
        if (e != hipSuccess)
            return (int)e;
        e = hipStreamSynchronize(s->stream);
        if (e != hipSuccess)
            return (int)e;
        const LaunchStats *h = (const LaunchStats *)(s->h_misc + 64);
        const uint64_t kt = h->k_total, lk = h->last_kept;
        if ((uint32_t)(kt >> kEpochShift) != s->epoch)
            return (int)hipErrorUnknown;   // the final group did not report
        *kept = (int64_t)(kt & ((1ull << kEpochShift) - 1));
        if (last)
            *last = ((uint32_t)(lk >> kEpochShift) == s->epoch)
                        ? (int64_t)(lk & ((1ull << kEpochShift) - 1)) - 1
                        : -1;
        return 0;
    };

    const int64_t diff = (int64_t)r - L;
    int64_t kept = 0, last = -1;
    CK(ensure_scratch(s, num_tiles, mode, table_size));
    p.gsum = s->d_gsum;
    p.chunks = s->d_chunks;
    p.block_tabs = s->d_blocks;
    p.chunks_pre = s->d_pre;
    // per-tile records serve later prefix queries and the last kept start; a device
    // pass that asks for neither drops them (bitmap mode always reads its own)
    const bool records = fixed_cap < 0 || want_last || mode == kModeBitmap;
    p.info = records ? s->d_info : nullptr;
    p.bitmap = s->d_bitmap;
    // the 32-byte-per-thread count hands its candidate masks to the emit
    const bool have_mask_buf = s->bitmap_words >= (size_t)num_tiles * (kTile / 32);
    p.masks = (mode == kModeTable || (mode == kModeBorderFree && !bf_lean)) && have_mask_buf ? s->d_bitmap : nullptr;

    if (mode == kModeBitmap) {
        // candidates -> serial greedy (gives the exact count) -> emit
        CK(next_epoch(s));
        p.epoch = s->epoch;
        if (kernel_ms)
            CK(hipEventRecord(s->ev0, s->stream));
        // the parallel selection when its scratch is there, else the serial greedy
        const bool bf = border_free(pat, L);
        const size_t sel = bitmap_select_bytes(num_tiles, L, bf);
        bool parallel = false;
        if (sel && bitmap_parallel_enabled()) {
            size_t cap = s->bmsel_cap;
            if (grow((void **)&s->d_bmsel, &cap, sel, false, s->stream) == hipSuccess) {
                s->bmsel_cap = cap;
                parallel = true;
            } else {
                (void)hipGetLastError();
            }
        }
        if (parallel) {
            CK(launch_bitmap_select(p, bf, s->d_bmsel, s->stream));
        } else {
            CK(launch_bitmap_mark(p, s->stream));
            CK(launch_bitmap_greedy(p, s->stream));
        }
        int rc = read_stats(&kept, &last);
        if (rc != 0)
            return rc;
        const int64_t exact = n + kept * diff;
        uint8_t *dst = nullptr;
        rc = dst_for(exact, &dst);
        if (rc != 0) {
            res->out_len = exact;
            return rc;
        }
        p.dst = dst;
        p.out_len = exact;
        p.write_output = 1;
        CK(launch_emit(p, kModeBitmap, 0, gt, 0, num_groups, s->stream));
        if (kernel_ms)
            CK(hipEventRecord(s->ev1, s->stream));
        CK(hipStreamSynchronize(s->stream));
        if (kernel_ms)
            CK(hipEventElapsedTime(kernel_ms, s->ev0, s->ev1));
    } else {
        // every chunk: count, spine, emit (stream order hands each chunk's final
        // state to the next chunk's spine)
        auto run_chunks = [&](bool emit) -> hipError_t {
            for (int64_t c = 0; c < num_chunks; ++c) {
                const int64_t g0 = c * cg;
                const int64_t gn = num_groups - g0 < cg ? num_groups - g0 : cg;
                hipError_t e = bf_lean ? launch_bf_count(p, g0, gn, s->stream)
                                       : launch_count(p, mode, table_size, gt, g0, gn, s->stream);
                if (e != hipSuccess)
                    return e;
                e = num_chunks == 1 ? launch_spine_blocks(p, mode, table_size, gn, s->stream)
                                    : launch_spine(p, mode, table_size, g0, gn, c, c == num_chunks - 1, s->stream);
                if (e != hipSuccess)
                    return e;
                if (emit) {
                    e = bf_lean ? launch_bf_emit(p, g0, gn, s->stream)
                                : launch_emit(p, mode, table_size, gt, g0, gn, s->stream);
                    if (e != hipSuccess)
                        return e;
                }
            }
            return hipSuccess;
        };
        CK(next_epoch(s));
        p.epoch = s->epoch;
        p.write_output = 1;

        // output size: shrink never grows; expand uses the non-overlapping worst
        // case when affordable, otherwise the exact size from count + spine first
        int64_t need = n;
        bool counted = false;
        if (diff > 0) {
            const int64_t span = floor_pos < n ? n - floor_pos : 0;
            const int64_t worst = n + (span / L) * diff;
            int64_t afford = 4 * n > (256ll << 20) ? 4 * n : (256ll << 20);
            if (fixed_cap >= 0)
                afford = fixed_cap;
            if (worst <= afford) {
                need = worst;
            } else {
                if (kernel_ms)
                    CK(hipEventRecord(s->ev0, s->stream));
                CK(run_chunks(false));
                int rc = read_stats(&kept, nullptr);
                if (rc != 0)
                    return rc;
                need = n + kept * diff;
                counted = true;
            }
        }
        uint8_t *dst = nullptr;
        int rc = dst_for(need, &dst);
        if (rc != 0) {
            res->out_len = need;
            return rc;
        }
        p.dst = dst;
        // border-free passes of a single chunk, and 2-state carry-table passes: one
        // persistent launch when large enough
        bool swept = false;
        const bool tbl_single = !counted && mode == kModeTable && table_size == 2 && num_chunks == 1 &&
                                table_single_chosen(s->device, n);
        const int64_t sweep_groups = tbl_single ? groups_of(num_tiles, kGroupTiles) : num_groups;
        if ((tbl_single || (!counted && single_launch(bf_var) && num_chunks == 1)) &&
            ensure_sweep(s, sweep_groups) == hipSuccess) {
            ReplaceParams q = p;
            q.num_groups = sweep_groups;
            if (kernel_ms)
                CK(hipEventRecord(s->ev0, s->stream));
            CK(tbl_single ? launch_tbl_sweep(q, s->d_sweep, sweep_spin_limit(), s->stream)
                          : launch_bf_sweep(q, s->d_sweep, sweep_ahead(), sweep_spin_limit(), bf_var == kBfKeep,
                                            s->stream));
            if (kernel_ms)
                CK(hipEventRecord(s->ev1, s->stream));
            CK(hipMemcpyAsync(s->h_misc + 96, s->d_sweep + 2 * sweep_groups, 8, hipMemcpyDeviceToHost, s->stream));
            if (want_last)
                CK(launch_last_kept(p, s->d_query, s->stream));
            rc = read_stats(&kept, nullptr);   // synchronizes
            const uint64_t fv = *(const uint64_t *)(s->h_misc + 96);
            swept = fv != (((uint64_t)s->epoch << kEpochShift) | 1u);
            if (swept && rc != 0)
                return rc;
            if (!swept) {
                // a wait gave up (the grid was not all resident): the three launches,
                // under a new epoch so nothing of the abandoned pass is read
                CK(next_epoch(s));
                p.epoch = s->epoch;
            }
        }
        if (!swept) {
            if (counted) {
                CK(bf_lean ? launch_bf_emit(p, 0, num_groups, s->stream)
                           : launch_emit(p, mode, table_size, gt, 0, num_groups, s->stream));
            } else {
                if (kernel_ms)
                    CK(hipEventRecord(s->ev0, s->stream));
                CK(run_chunks(true));
            }
            if (kernel_ms)
                CK(hipEventRecord(s->ev1, s->stream));
            if (want_last)
                CK(launch_last_kept(p, s->d_query, s->stream));
            rc = read_stats(&kept, nullptr);
        }
        if (rc != 0)
            return rc;
        if (want_last) {
            CK(hipMemcpyAsync(s->h_misc + 128, s->d_query, 8, hipMemcpyDeviceToHost, s->stream));
            CK(hipStreamSynchronize(s->stream));
            last = ((const int64_t *)(s->h_misc + 128))[0] - 1;
        }
        if (kernel_ms)
            CK(hipEventElapsedTime(kernel_ms, s->ev0, s->ev1));
        if (tune && num_chunks == 1 && mode == kModeBorderFree) {
            // every form recomputes the same summaries, states, records and output
            // (a single launch that gives up leaves what it wrote equal to the
            // rest), so the pass's results above stay valid.  Three interleaved
            // sweeps, best time per form; another form replaces the default only
            // when it is more than 3% faster.
            TuneEntry te;
            te.gt = kGroupTiles;
            const ReplaceParams p0 = p;
            constexpr float kGaveUp = 1e30f;
            for (int sweep = 0; sweep < 3; ++sweep) {
                for (int v = 0; v < kBfVariants; ++v) {
                    ReplaceParams q = p0;
                    const int vg = kBfVariantTiles[v];
                    const int64_t ng = groups_of(num_tiles, vg);
                    q.num_groups = ng;
                    q.masks = v >= kBfTiles1 && have_mask_buf ? s->d_bitmap : nullptr;
                    float ms = kGaveUp;
                    if (single_launch(v)) {
                        if (sweep_min_bytes() == INT64_MAX || ensure_sweep(s, ng) != hipSuccess) {
                            (void)hipGetLastError();
                            te.vms[v] = kGaveUp;
                            continue;
                        }
                        CK(next_epoch(s));
                        q.epoch = s->epoch;
                        CK(hipEventRecord(s->ev0, s->stream));
                        CK(launch_bf_sweep(q, s->d_sweep, sweep_ahead(), sweep_spin_limit(), v == kBfKeep,
                                           s->stream));
                        CK(hipEventRecord(s->ev1, s->stream));
                        CK(hipMemcpyAsync(s->h_misc + 96, s->d_sweep + 2 * ng, 8, hipMemcpyDeviceToHost,
                                          s->stream));
                        CK(hipStreamSynchronize(s->stream));
                        const uint64_t fv = *(const uint64_t *)(s->h_misc + 96);
                        if (fv != (((uint64_t)s->epoch << kEpochShift) | 1u))
                            CK(hipEventElapsedTime(&ms, s->ev0, s->ev1));
                    } else {
                        const bool lean = v == kBfLean;
                        CK(hipEventRecord(s->ev0, s->stream));
                        CK(lean ? launch_bf_count(q, 0, ng, s->stream)
                                : launch_count(q, mode, table_size, vg, 0, ng, s->stream));
                        CK(launch_spine_blocks(q, mode, table_size, ng, s->stream));
                        CK(lean ? launch_bf_emit(q, 0, ng, s->stream)
                                : launch_emit(q, mode, table_size, vg, 0, ng, s->stream));
                        CK(hipEventRecord(s->ev1, s->stream));
                        CK(hipEventSynchronize(s->ev1));
                        CK(hipEventElapsedTime(&ms, s->ev0, s->ev1));
                    }
                    te.vms[v] = sweep == 0 || ms < te.vms[v] ? ms : te.vms[v];
                }
            }
            const int dflt = default_bf_variant(n);
            te.variant = dflt;
            float best = te.vms[dflt];
            for (int v = 0; v < kBfVariants; ++v) {
                if (te.vms[v] < best && te.vms[v] < 0.97f * te.vms[dflt]) {
                    te.variant = v;
                    best = te.vms[v];
                }
            }
            record_tuning(s->device, mode, table_size, n, te);
        } else if (tune && num_chunks == 1) {
            // every candidate recomputes the same summaries, states, records and
            // output, so the pass's results above stay valid.  Three interleaved
            // sweeps (so clock and cache drift hit every candidate alike), best
            // time per candidate; another size replaces the default only when it
            // is more than 3% faster.
            TuneEntry te;
            // 2-state tables: the single launch is a candidate too
            const int64_t g16 = groups_of(num_tiles, kGroupTiles);
            const bool time_single = mode == kModeTable && table_size == 2 && !counted &&
                                     table_sweep_min_bytes() != INT64_MAX && ensure_sweep(s, g16) == hipSuccess;
            for (int sweep = 0; sweep < 3; ++sweep) {
                if (time_single) {
                    ReplaceParams q = p;
                    q.num_groups = g16;
                    CK(next_epoch(s));
                    q.epoch = s->epoch;
                    CK(hipEventRecord(s->ev0, s->stream));
                    CK(launch_tbl_sweep(q, s->d_sweep, sweep_spin_limit(), s->stream));
                    CK(hipEventRecord(s->ev1, s->stream));
                    CK(hipMemcpyAsync(s->h_misc + 96, s->d_sweep + 2 * g16, 8, hipMemcpyDeviceToHost, s->stream));
                    CK(hipStreamSynchronize(s->stream));
                    float ms = 1e30f;
                    if (*(const uint64_t *)(s->h_misc + 96) != (((uint64_t)s->epoch << kEpochShift) | 1u))
                        CK(hipEventElapsedTime(&ms, s->ev0, s->ev1));
                    te.single_ms = sweep == 0 || ms < te.single_ms ? ms : te.single_ms;
                }
                for (int ci = 0; ci < 3; ++ci) {
                    const int cand = kTuneCands[ci];
                    const int64_t ng = groups_of(num_tiles, cand);
                    CK(hipEventRecord(s->ev0, s->stream));
                    CK(launch_count(p, mode, table_size, cand, 0, ng, s->stream));
                    CK(launch_spine_blocks(p, mode, table_size, ng, s->stream));
                    CK(launch_emit(p, mode, table_size, cand, 0, ng, s->stream));
                    CK(hipEventRecord(s->ev1, s->stream));
                    CK(hipEventSynchronize(s->ev1));
                    float ms = 0.f;
                    CK(hipEventElapsedTime(&ms, s->ev0, s->ev1));
                    te.ms[ci] = sweep == 0 || ms < te.ms[ci] ? ms : te.ms[ci];
                }
            }
            te.gt = kGroupTiles;
            float best = te.ms[1];
            for (int ci = 0; ci < 3; ++ci) {
                if (te.ms[ci] < best && te.ms[ci] < 0.97f * te.ms[1]) {
                    te.gt = kTuneCands[ci];
                    best = te.ms[ci];
                }
            }
            if (time_single) {
                // the default (the single launch from the threshold on, else the three
                // launches) stays unless the other is more than 3% faster
                const bool dflt = n >= table_sweep_min_bytes();
                te.single = dflt ? !(best < 0.97f * te.single_ms) : te.single_ms < 0.97f * best;
            }
            record_tuning(s->device, mode, table_size, n, te);
        }
    }
    if (mode_out)
        *mode_out = mode;
    res->kept = kept;
    res->last = last;
    res->out_len = n + kept * diff;
    if (params_out) {
        *params_out = p;
        params_out->epoch = s->epoch;
    }
    return 0;
}

// session apply: d_buf -> d_tmp, swap
int session_apply(Session *s, const uint8_t *pat, int L, const uint8_t *rep, int r, int64_t floor_pos,
                  size_t *new_len, long long *out_last)
{
    if ((L > 0 && !pat) || (r > 0 && !rep) || L < 0 || r < 0)
        return -1;
    PassResult res;
    int mode = 0;
    ReplaceParams qp{};
    auto dst_for = [&](int64_t need, uint8_t **dst) -> int {
        hipError_t e = grow((void **)&s->d_tmp, &s->cap_tmp, (size_t)(need > 0 ? need : 1), false, s->stream);
        if (e != hipSuccess)
            return (int)e;
        *dst = s->d_tmp;
        return 0;
    };
    s->q_valid = false;
    s->fused_pending = false;
    int rc = run_pass(s, s->d_buf, (int64_t)s->len, pat, L, rep, r, floor_pos, -1, out_last != nullptr, dst_for,
                      &res, nullptr, &mode, &qp);
    if (rc != 0)
        return rc;
    std::swap(s->d_buf, s->d_tmp);
    std::swap(s->cap_buf, s->cap_tmp);
    const size_t src_len = s->len;
    s->len = (size_t)res.out_len;
    // the previous contents now live in d_tmp: queries recompute from there
    s->q_valid = true;
    s->q_mode = mode;
    s->q_params = qp;
    s->q_params.src = s->d_tmp;
    s->q_params.n = (int64_t)src_len;
    if (new_len)
        *new_len = s->len;
    if (out_last)
        *out_last = res.last;
    return 0;
}

int replay_fused(Session *s);

int session_query(Session *s, size_t limit, int64_t *count, int64_t *last)
{
    *count = 0;
    *last = -1;
    if (s->fused_pending) {
        const int rc = replay_fused(s);
        if (rc != 0)
            return rc;
    }
    if (!s->q_valid)
        return 0;
    const int64_t n = s->q_params.n;
    const int64_t lim = (int64_t)limit < n ? (int64_t)limit : n;
    if (lim <= 0)
        return 0;
    if (s->q_mode == kModeEmpty) {
        *count = lim;
        *last = lim - 1;
        return 0;
    }
    CK(launch_query(s->q_params, s->q_mode, lim, s->d_query, s->stream));
    CK(hipMemcpyAsync(s->h_misc + 128, s->d_query, 16, hipMemcpyDeviceToHost, s->stream));
    CK(hipStreamSynchronize(s->stream));
    const int64_t *h = (const int64_t *)(s->h_misc + 128);
    *count = h[0];
    *last = h[1];
    return 0;
}

int session_upload(Session *s, const uint8_t *src, size_t len)
{
    // over-read slack for full 16-byte tile loads is not needed (loads are bounded)
    CK(grow((void **)&s->d_buf, &s->cap_buf, len > 0 ? len : 1, false, s->stream));
    if (len > 0)
        CK(hipMemcpyAsync(s->d_buf, src, len, hipMemcpyHostToDevice, s->stream));
    CK(hipStreamSynchronize(s->stream));
    s->len = len;
    s->q_valid = false;
    s->fused_pending = false;
    return 0;
}


// ---- fused multi-pattern apply_batch -----------------------------------------------
// Builds the fused table when one pass over the source equals applying the pairs in
// order: 2..kMaxFusedPats pairs; patterns of 2..kMaxFusedPat bytes sharing their first
// byte, border-free, distinct, none a substring of another, no proper suffix of one
// equal to a prefix of another (so no two occurrences of any of them overlap and every
// occurrence is kept); replacements of 1..kMaxFusedPat bytes that contain no byte of
// any pattern (so no replacement forms or completes an occurrence of a later pattern,
// and none is deleted to join bytes around it).  Returns false otherwise.
bool build_fused(const uint8_t *const *pats, const int *pat_lens, const uint8_t *const *reps, const int *rep_lens,
                 int count, FusedTable *t)
{
    if (count < 2 || count > kMaxFusedPats || !pats || !pat_lens || !reps || !rep_lens)
        return false;
    bool in_pat[256] = {false};
    int min_len = INT_MAX, max_len = 0;
    size_t pool = 0;
    for (int i = 0; i < count; ++i) {
        const int L = pat_lens[i], r = rep_lens[i];
        if (L < 2 || L > kMaxFusedPat || r < 1 || r > kMaxFusedPat || !pats[i] || !reps[i])
            return false;
        if (pats[i][0] != pats[0][0] || !border_free(pats[i], L))
            return false;
        for (int k = 0; k < L; ++k)
            in_pat[pats[i][k]] = true;
        min_len = std::min(min_len, L);
        max_len = std::max(max_len, L);
        pool += (size_t)L + (size_t)r;
    }
    if (pool > (size_t)kFusedPool)
        return false;
    for (int i = 0; i < count; ++i)
        for (int k = 0; k < rep_lens[i]; ++k)
            if (in_pat[reps[i][k]])
                return false;
    const uint8_t c0 = pats[0][0];
    for (int a = 0; a < count; ++a) {
        const uint8_t *A = pats[a];
        const int La = pat_lens[a];
        for (int b = 0; b < count; ++b) {
            if (a == b)
                continue;
            const uint8_t *B = pats[b];
            const int Lb = pat_lens[b];
            // A inside B (every occurrence of A starts with c0) ...
            for (int x = 0; x + La <= Lb; ++x)
                if (B[x] == c0 && memcmp(B + x, A, (size_t)La) == 0)
                    return false;
            // ... or a proper suffix of A equal to a prefix of B
            for (int x = 1; x < La; ++x) {
                const int l = La - x;
                if (l < Lb && A[x] == c0 && memcmp(A + x, B, (size_t)l) == 0)
                    return false;
            }
        }
    }
    memset(t, 0, sizeof(FusedTable));
    t->count = count;
    t->key_len = std::min(4, min_len - 1);
    t->key_mask = t->key_len == 4 ? 0xffffffffu : (1u << (8 * t->key_len)) - 1u;
    t->min_len = min_len;
    t->max_len = max_len;
    uint32_t off = 0;
    for (int i = 0; i < count; ++i) {
        t->pat_len[i] = (uint16_t)pat_lens[i];
        t->rep_len[i] = (uint16_t)rep_lens[i];
        t->pat_off[i] = (uint16_t)off;
        memcpy(t->pool + off, pats[i], (size_t)pat_lens[i]);
        off += (uint32_t)pat_lens[i];
        t->rep_off[i] = (uint16_t)off;
        memcpy(t->pool + off, reps[i], (size_t)rep_lens[i]);
        off += (uint32_t)rep_lens[i];
        memcpy(&t->head[i], pats[i], (size_t)std::min(16, pat_lens[i]));
        memcpy(&t->rep_head[i], reps[i], (size_t)std::min(16, rep_lens[i]));
        uint32_t key = 0;
        for (int b = 0; b < t->key_len; ++b)
            key |= (uint32_t)pats[i][1 + b] << (8 * b);
        uint32_t h = fused_hash(key);
        while (t->slot[h].z != 0)
            h = (h + 1) & (kFusedSlots - 1);
        uint8_t body[kFusedSlotBytes] = {0};
        memcpy(body, pats[i] + 1, (size_t)std::min(kFusedSlotBytes, pat_lens[i] - 1));
        uint32_t rep4 = 0;
        memcpy(&rep4, reps[i], (size_t)std::min(4, rep_lens[i]));
        uint32_t lo, hi;
        memcpy(&lo, body, 4);
        memcpy(&hi, body + 4, 4);
        t->slot[h].x = lo;
        t->slot[h].y = hi;
        t->slot[h].z = (uint32_t)i | (uint32_t)pat_lens[i] << 8 | (uint32_t)rep_lens[i] << 16 | kFusedUsed;
        t->slot[h].w = rep4;
    }
    return true;
}

// One fused pass over the session buffer with the table in s->d_fused (uploaded from
// s->h_fused first when `upload`).
int session_apply_fused(Session *s, bool upload)
{
    CK(ensure_misc(s));
    if (upload) {
        if (!s->d_fused)
            CK(hipMalloc((void **)&s->d_fused, sizeof(FusedTable)));
        CK(hipMemcpyAsync(s->d_fused, s->h_fused, sizeof(FusedTable), hipMemcpyHostToDevice, s->stream));
    }
    const int64_t n = (int64_t)s->len;
    int64_t out = n;
    if (n > 0) {
        const int64_t num_tiles = (n + kTile - 1) / kTile;
        const int64_t num_groups = groups_of(num_tiles, kGroupTiles);
        CK(ensure_scratch(s, num_tiles, kModeBorderFree, 0));
        CK(next_epoch(s));
        ReplaceParams p{};
        p.src = s->d_buf;
        p.n = n;
        p.floor = 0;
        p.out_len = -1;
        p.L = 1;
        p.pat4 = s->h_fused->pool[0];
        p.mask4 = 0xffu;
        p.gsum = s->d_gsum;
        p.chunks = s->d_chunks;
        p.block_tabs = s->d_blocks;
        p.chunks_pre = s->d_pre;
        p.masks = s->bitmap_words >= (size_t)num_tiles * (kTile / 32) ? s->d_bitmap : nullptr;
        p.stats = s->d_stats;
        p.epoch = s->epoch;
        p.write_output = 1;
        p.num_tiles = num_tiles;
        p.num_groups = num_groups;
        // output size: at most n + (n / shortest pattern) * largest growth; when that
        // bound is affordable (the single-pair rule) the emit follows the spine with
        // no readback in between, otherwise the exact size is read first
        const int64_t min_len = s->h_fused->min_len;
        const int64_t worst = s->fused_max_grow <= 0 ? n : n + (n / min_len) * s->fused_max_grow;
        const int64_t afford = 4 * n > (256ll << 20) ? 4 * n : (256ll << 20);
        const bool bounded = worst <= afford;
        // the signed total length change, from the stats block copied to h_misc + 64
        auto parse_total = [&](int64_t *total) -> int {
            const uint64_t kt = ((const LaunchStats *)(s->h_misc + 64))->k_total;
            if ((uint32_t)(kt >> kEpochShift) != s->epoch)
                return (int)hipErrorUnknown;
            *total = n + ((int64_t)(kt << (64 - kEpochShift)) >> (64 - kEpochShift));   // sign-extend
            return 0;
        };
        auto read_total = [&](int64_t *total) -> int {
            CK(hipMemcpyAsync(s->h_misc + 64, s->d_stats, sizeof(LaunchStats), hipMemcpyDeviceToHost, s->stream));
            CK(hipStreamSynchronize(s->stream));
            return parse_total(total);
        };
        if (bounded)
            CK(grow((void **)&s->d_tmp, &s->cap_tmp, (size_t)(worst > 0 ? worst : 1), false, s->stream));
        // with the output bounded, large passes run as one persistent launch
        bool swept = false;
        if (bounded && n >= fused_sweep_min_bytes() && ensure_sweep(s, num_groups) == hipSuccess) {
            p.dst = s->d_tmp;
            CK(launch_fused_sweep(p, s->d_fused, s->d_sweep, sweep_spin_limit(), s->stream));
            // the give-up flag and the stats block under one synchronization
            CK(hipMemcpyAsync(s->h_misc + 96, s->d_sweep + 2 * num_groups, 8, hipMemcpyDeviceToHost, s->stream));
            CK(hipMemcpyAsync(s->h_misc + 64, s->d_stats, sizeof(LaunchStats), hipMemcpyDeviceToHost, s->stream));
            CK(hipStreamSynchronize(s->stream));
            swept = *(const uint64_t *)(s->h_misc + 96) != (((uint64_t)s->epoch << kEpochShift) | 1u);
            if (!swept) {
                // a wait gave up: the three launches under a new epoch
                CK(next_epoch(s));
                p.epoch = s->epoch;
            }
        }
        if (!swept) {
            CK(launch_fused(p, s->d_fused, 0, 0, num_groups, s->stream));
            CK(launch_fused(p, s->d_fused, 1, 0, num_groups, s->stream));
            if (!bounded) {
                int rc = read_total(&out);
                if (rc != 0)
                    return rc;
                CK(grow((void **)&s->d_tmp, &s->cap_tmp, (size_t)(out > 0 ? out : 1), false, s->stream));
            }
            p.dst = s->d_tmp;
            CK(launch_fused(p, s->d_fused, 2, 0, num_groups, s->stream));
        }
        int rc = swept ? parse_total(&out) : read_total(&out);
        if (rc != 0)
            return rc;
        if (bounded && out > worst)
            return (int)hipErrorUnknown;
    } else {
        CK(grow((void **)&s->d_tmp, &s->cap_tmp, 1, false, s->stream));
    }
    std::swap(s->d_buf, s->d_tmp);
    std::swap(s->cap_buf, s->cap_tmp);
    s->fused_src_len = (size_t)n;
    s->len = (size_t)out;
    return 0;
}

// Pairs in order; fused into one pass when build_fused allows (and allow_fuse).
int session_apply_batch(Session *s, const uint8_t *const *pats, const int *pat_lens, const uint8_t *const *reps,
                        const int *rep_lens, int count, bool allow_fuse, int *fused_out)
{
    if (fused_out)
        *fused_out = 0;
    if (allow_fuse && count >= 2 && count <= kMaxFusedPats && pats && pat_lens && reps && rep_lens) {
        // the same pair list as the last call reuses that call's check and table
        std::vector<uint8_t> key;
        bool keyable = true;
        for (int i = 0; i < count && keyable; ++i) {
            const int L = pat_lens[i], r = rep_lens[i];
            if (L < 0 || r < 0 || L > kMaxFusedPat || r > kMaxFusedPat || (L && !pats[i]) || (r && !reps[i])) {
                keyable = false;
                break;
            }
            key.push_back((uint8_t)L);
            key.push_back((uint8_t)r);
            key.insert(key.end(), pats[i], pats[i] + L);
            key.insert(key.end(), reps[i], reps[i] + r);
        }
        const bool cached = keyable && s->fused_key_state != 0 && key == s->fused_key;
        bool fusable = cached && s->fused_key_state == 1;
        if (keyable && !cached) {
            if (!s->h_fused)
                CK(hipHostMalloc((void **)&s->h_fused, sizeof(FusedTable), hipHostMallocDefault));
            CK(hipStreamSynchronize(s->stream));   // the staging may still feed an earlier upload
            fusable = build_fused(pats, pat_lens, reps, rep_lens, count, s->h_fused);
            s->fused_key.swap(key);
            s->fused_key_state = fusable ? 1 : 2;
            if (fusable) {
                int g = INT_MIN;
                for (int i = 0; i < count; ++i)
                    g = std::max(g, rep_lens[i] - pat_lens[i]);
                s->fused_max_grow = g;
            }
        }
        if (fusable) {
            s->q_valid = false;
            s->fused_pending = false;
            const bool same_pairs = cached && (int)s->fused_pats.size() == count;
            int rc = session_apply_fused(s, !cached);
            if (rc != 0) {
                s->fused_key_state = 0;
                return rc;
            }
            if (!same_pairs) {
                s->fused_pats.assign(count, {});
                s->fused_reps.assign(count, {});
                for (int i = 0; i < count; ++i) {
                    s->fused_pats[i].assign(pats[i], pats[i] + pat_lens[i]);
                    s->fused_reps[i].assign(reps[i], reps[i] + rep_lens[i]);
                }
            }
            s->fused_pending = true;
            if (fused_out)
                *fused_out = 1;
            return 0;
        }
    }
    for (int i = 0; i < count; ++i) {
        const uint8_t *r = (reps && rep_lens) ? reps[i] : nullptr;
        const int rl = (reps && rep_lens) ? rep_lens[i] : 0;
        int rc = session_apply(s, pats[i], pat_lens[i], r, rl, 0, nullptr, nullptr);
        if (rc != 0)
            return rc;
    }
    return 0;
}

// A prefix query after a fused batch refers to the batch's last pair, applied to the
// output of the ones before it: replay the pairs one by one from the batch's source
// (still in d_tmp) into two scratch buffers, keeping the last pass's record.
int replay_fused(Session *s)
{
    const int count = (int)s->fused_pats.size();
    const uint8_t *src = s->d_tmp;
    int64_t n = (int64_t)s->fused_src_len;
    for (int i = 0; i < count; ++i) {
        const int b = i & 1;
        auto dst_for = [&](int64_t need, uint8_t **dst) -> int {
            hipError_t e = grow((void **)&s->d_q[b], &s->cap_q[b], (size_t)(need > 0 ? need : 1), false, s->stream);
            if (e != hipSuccess)
                return (int)e;
            *dst = s->d_q[b];
            return 0;
        };
        PassResult res;
        int mode = 0;
        ReplaceParams qp{};
        const std::vector<uint8_t> &P = s->fused_pats[(size_t)i], &R = s->fused_reps[(size_t)i];
        int rc = run_pass(s, src, n, P.data(), (int)P.size(), R.data(), (int)R.size(), 0, -1, false, dst_for, &res,
                          nullptr, &mode, &qp);
        if (rc != 0)
            return rc;
        if (i == count - 1) {
            s->q_params = qp;
            s->q_params.src = src;
            s->q_params.n = n;
            s->q_mode = mode;
            s->q_valid = true;
        }
        src = s->d_q[b];
        n = res.out_len;
    }
    s->fused_pending = false;
    return 0;
}


// ---- streaming pipeline ---------------------------------------------------------
// Pinned host output that grows; in-flight downloads into it finish first.
struct PinnedOut {
    uint8_t *p = nullptr;
    size_t cap = 0;
    size_t len = 0;
};

hipError_t pinned_reserve(PinnedOut &o, size_t need, hipStream_t pending)
{
    if (o.p && o.cap >= need)
        return hipSuccess;
    hipError_t e = hipStreamSynchronize(pending);
    if (e != hipSuccess)
        return e;
    const size_t nc = round_up(std::max(need, o.cap + o.cap / 2), 1u << 20);
    void *np = nullptr;
    e = hipHostMalloc(&np, nc, hipHostMallocDefault);
    if (e != hipSuccess)
        return e;
    if (o.p) {
        parallel_memcpy(np, o.p, o.len);
        (void)hipHostFree(o.p);
    }
    o.p = (uint8_t *)np;
    o.cap = nc;
    return hipSuccess;
}

hipError_t ensure_stream_pipeline(Session *s)
{
    if (s->cs_in)
        return hipSuccess;
    hipError_t e = hipStreamCreateWithFlags(&s->cs_in, hipStreamNonBlocking);
    if (e == hipSuccess)
        e = hipStreamCreateWithFlags(&s->cs_out, hipStreamNonBlocking);
    for (int b = 0; b < 2 && e == hipSuccess; ++b) {
        e = hipEventCreateWithFlags(&s->ev_in[b], hipEventDisableTiming);
        if (e == hipSuccess)
            e = hipEventCreateWithFlags(&s->ev_out[b], hipEventDisableTiming);
        if (e == hipSuccess)
            e = hipEventRecord(s->ev_in[b], s->cs_in);
        if (e == hipSuccess)
            e = hipEventRecord(s->ev_out[b], s->cs_out);
    }
    return e;
}

// One pair over X (n bytes) in chunks of C source bytes: chunk k covers source
// [kC, (k+1)C) plus an (L-1)-byte overlap, its seed is the last kept start before
// kC, and its output runs from the end of a match straddling kC (dropped bytes)
// to the end of a match straddling (k+1)C (cuda_replace_wrapper.py:305-366, with
// the cut moved to the end of a straddling match so no byte is lost).  Chunk
// k+1 is staged and uploaded while chunk k computes; chunk k's output is
// downloaded into Y while chunk k+1 computes.
int stream_pair(Session *s, const uint8_t *X, size_t n, bool x_pinned, const uint8_t *pat, int L,
                const uint8_t *rep, int r, size_t C, PinnedOut &Y, int64_t *chunks)
{
    const int64_t diff = (int64_t)r - L;
    Y.len = 0;
    CK(pinned_reserve(Y, (diff <= 0 ? n : n + n / 8) + (size_t)r + 64, s->cs_out));
    if (n == 0) {
        if (L == 0 && r > 0) {
            memcpy(Y.p, rep, (size_t)r);
            Y.len = (size_t)r;
        }
        return 0;
    }
    const size_t H = L > 0 ? (size_t)L - 1 : 0;
    const size_t K = (n + C - 1) / C;
    constexpr int64_t kNone = INT64_MIN;
    int64_t last_global = kNone;   // last kept start of the chunks so far
    size_t cut_prev = 0;           // source position where the previous chunk's output ended

    auto upload = [&](size_t k) -> int {
        const int b = (int)(k & 1);
        const size_t start = k * C;
        const size_t len = std::min(n, start + C + H) - start;
        // the slot's previous upload (chunk k-2) was consumed by a finished pass
        CK(grow((void **)&s->d_sin[b], &s->cap_sin[b], len + 16, false, s->stream));
        const uint8_t *src = X + start;
        if (!x_pinned) {
            CK(hipEventSynchronize(s->ev_in[b]));   // staging slot b is free again
            CK(grow_host(&s->h_sin[b], &s->cap_hsin[b], len + 16));
            parallel_memcpy(s->h_sin[b], src, len);
            src = s->h_sin[b];
        }
        CK(hipMemcpyAsync(s->d_sin[b], src, len, hipMemcpyHostToDevice, s->cs_in));
        CK(hipEventRecord(s->ev_in[b], s->cs_in));
        return 0;
    };

    int rc = upload(0);
    if (rc != 0)
        return rc;
    for (size_t k = 0; k < K; ++k) {
        const int b = (int)(k & 1);
        const size_t start = k * C;
        const size_t len = std::min(n, start + C + H) - start;
        const bool last_chunk = k + 1 == K;
        if (!last_chunk && (rc = upload(k + 1)) != 0)
            return rc;
        CK(hipStreamWaitEvent(s->stream, s->ev_in[b], 0));    // chunk k is on the device
        CK(hipStreamWaitEvent(s->stream, s->ev_out[b], 0));   // slot b's last download is done
        int64_t floor_pos = 0;
        if (last_global != kNone) {
            const int64_t f = last_global + L - (int64_t)start;
            floor_pos = f > 0 ? f : 0;
        }
        PassResult res;
        int mode = 0;
        ReplaceParams qp{};
        auto dst_for = [&](int64_t need, uint8_t **dst) -> int {
            hipError_t e = hipEventSynchronize(s->ev_out[b]);
            if (e == hipSuccess)
                e = grow((void **)&s->d_sout[b], &s->cap_sout[b], (size_t)(need > 0 ? need : 1) + 16, false,
                         s->stream);
            if (e != hipSuccess)
                return (int)e;
            *dst = s->d_sout[b];
            return 0;
        };
        rc = run_pass(s, s->d_sin[b], (int64_t)len, pat, L, rep, r, floor_pos, -1, false, dst_for, &res, nullptr,
                      &mode, &qp);
        if (rc != 0)
            return rc;
        // output slice of this chunk
        const size_t drop = cut_prev > start ? cut_prev - start : 0;
        size_t end;
        size_t cut;
        if (L == 0) {
            // rep b0 rep b1 ... rep: the trailing rep belongs to the next chunk
            end = last_chunk ? (size_t)res.out_len : len * ((size_t)r + 1);
            cut = start + len;
        } else if (last_chunk) {
            end = (size_t)res.out_len;
            cut = n;
        } else {
            CK(launch_query(qp, mode, (int64_t)C, s->d_query, s->stream));
            CK(hipMemcpyAsync(s->h_misc + 128, s->d_query, 16, hipMemcpyDeviceToHost, s->stream));
            CK(hipStreamSynchronize(s->stream));
            const int64_t kept_before = ((const int64_t *)(s->h_misc + 128))[0];
            const int64_t last_rel = ((const int64_t *)(s->h_misc + 128))[1];
            if (last_rel >= 0)
                last_global = (int64_t)start + last_rel;
            // the latest kept match may cover past the chunk end (also one from an
            // earlier chunk when L exceeds the chunk size)
            cut = start + C;
            if (last_global != kNone)
                cut = std::max(cut, (size_t)(last_global + L));
            // no kept start lies in [C, cut - start): it would overlap the straddling one
            end = (size_t)((int64_t)(cut - start) + kept_before * diff);
        }
        if (end > drop) {
            CK(pinned_reserve(Y, Y.len + (end - drop) + 64, s->cs_out));
            CK(hipMemcpyAsync(Y.p + Y.len, s->d_sout[b] + drop, end - drop, hipMemcpyDeviceToHost, s->cs_out));
            Y.len += end - drop;
        }
        CK(hipEventRecord(s->ev_out[b], s->cs_out));
        cut_prev = cut;
        if (chunks)
            ++*chunks;
    }
    CK(hipStreamSynchronize(s->cs_out));
    return 0;
}

// Pairs in order, each streamed over the previous pair's output.
int session_stream(Session *s, const uint8_t *src, size_t n, const uint8_t *const *pats, const int *pat_lens,
                   const uint8_t *const *reps, const int *rep_lens, int count, size_t chunk, uint8_t **out,
                   size_t *out_len, double *stats)
{
    const auto t0 = std::chrono::steady_clock::now();
    CK(ensure_stream_pipeline(s));
    if (chunk == 0)
        chunk = 256u << 20;
    PinnedOut cur;
    const uint8_t *X = src;
    size_t xn = n;
    bool x_pinned = false;
    double bytes_in = 0, bytes_out = 0;
    int64_t chunks = 0;
    for (int i = 0; i < count; ++i) {
        PinnedOut Y;
        int rc = stream_pair(s, X, xn, x_pinned, pats[i], pat_lens[i], reps ? reps[i] : nullptr,
                             reps ? rep_lens[i] : 0, chunk, Y, &chunks);
        if (rc != 0) {
            if (Y.p)
                (void)hipHostFree(Y.p);
            if (cur.p)
                (void)hipHostFree(cur.p);
            return rc;
        }
        bytes_in += (double)xn;
        bytes_out += (double)Y.len;
        if (cur.p)
            (void)hipHostFree(cur.p);
        cur = Y;
        X = cur.p;
        xn = cur.len;
        x_pinned = true;
    }
    if (count == 0) {
        CK(pinned_reserve(cur, n + 64, s->cs_out));
        parallel_memcpy(cur.p, src, n);
        cur.len = n;
    }
    *out = cur.len ? cur.p : nullptr;
    *out_len = cur.len;
    if (!cur.len && cur.p)
        (void)hipHostFree(cur.p);
    if (stats) {
        stats[0] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        stats[1] = bytes_in;
        stats[2] = bytes_out;
        stats[3] = (double)chunks;
    }
    return 0;
}

void destroy(Session *s)
{
    if (!s)
        return;
    DeviceGuard g(s->device);
    if (s->stream)
        (void)hipStreamSynchronize(s->stream);
    for (hipStream_t cs : {s->cs_in, s->cs_out})
        if (cs)
            (void)hipStreamSynchronize(cs);
    void *dev[] = {s->d_buf, s->d_tmp, s->d_gsum, s->d_chunks, s->d_pre, s->d_info, s->d_bitmap, s->d_patrep,
                   s->d_presence, s->d_misc, s->d_sin[0], s->d_sin[1], s->d_sout[0], s->d_sout[1], s->d_blocks,
                   s->d_sweep, s->d_bmsel};
    for (void *p : dev)
        if (p)
            (void)hipFree(p);
    for (uint8_t *p : {s->h_sin[0], s->h_sin[1]})
        if (p)
            (void)hipHostFree(p);
    for (void *p : {(void *)s->d_fused, (void *)s->d_q[0], (void *)s->d_q[1]})
        if (p)
            (void)hipFree(p);
    if (s->h_fused)
        (void)hipHostFree(s->h_fused);
    for (hipEvent_t e : {s->ev_in[0], s->ev_in[1], s->ev_out[0], s->ev_out[1]})
        if (e)
            (void)hipEventDestroy(e);
    for (hipStream_t cs : {s->cs_in, s->cs_out})
        if (cs)
            (void)hipStreamDestroy(cs);
    if (s->h_patrep)
        (void)hipHostFree(s->h_patrep);
    if (s->h_misc)
        (void)hipHostFree(s->h_misc);
    if (s->ev0)
        (void)hipEventDestroy(s->ev0);
    if (s->ev1)
        (void)hipEventDestroy(s->ev1);
    if (s->stream)
        (void)hipStreamDestroy(s->stream);
    delete s;
}

int open_session(Session **out)
{
    Session *s = new (std::nothrow) Session();
    if (!s)
        return -2;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        delete s;
        return (int)e;
    }
    s->device = dev;
    e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e == hipSuccess)
        e = hipEventCreate(&s->ev0);
    if (e == hipSuccess)
        e = hipEventCreate(&s->ev1);
    if (e != hipSuccess) {
        destroy(s);
        return (int)e;
    }
    int rc = (int)ensure_misc(s);
    if (rc != 0) {
        destroy(s);
        return rc;
    }
    *out = s;
    return 0;
}

} // namespace

// =============================================================================
// C ABI
// =============================================================================
extern "C" {

int cr_stream_replace(void *h, const uint8_t *src, size_t n, const uint8_t *const *pats, const int *pat_lens,
                      const uint8_t *const *reps, const int *rep_lens, int count, size_t chunk_bytes,
                      uint8_t **out_host, size_t *out_len, double *stats)
{
    if (!h || !out_host || !out_len || (n > 0 && !src) || count < 0 || (count > 0 && (!pats || !pat_lens)))
        return -1;
    for (int i = 0; i < count; ++i) {
        if (pat_lens[i] < 0 || (pat_lens[i] > 0 && !pats[i]))
            return -1;
        if (reps && (!rep_lens || rep_lens[i] < 0 || (rep_lens[i] > 0 && !reps[i])))
            return -1;
    }
    Session *s = (Session *)h;
    std::lock_guard<std::mutex> lock(s->mtx);
    if (!s->alive)
        return -3;
    DeviceGuard g(s->device);
    *out_host = nullptr;
    *out_len = 0;
    s->q_valid = false;   // the pipeline's passes are not the session's applies
    s->fused_pending = false;
    return session_stream(s, src, n, pats, pat_lens, reps, rep_lens, count, chunk_bytes, out_host, out_len, stats);
}

const char *cuda_replace_version(void) { return "cuda_replace_amd 1.0 gfx950 count-spine-emit"; }

int cuda_replace_open(void **ph, const uint8_t *src, size_t len)
{
    if (!ph || (len > 0 && !src))
        return -1;
    *ph = nullptr;
    Session *s = nullptr;
    int rc = open_session(&s);
    if (rc != 0)
        return rc;
    rc = session_upload(s, src, len);
    if (rc != 0) {
        destroy(s);
        return rc;
    }
    *ph = s;
    return 0;
}

int cuda_replace_reset(void *h, const uint8_t *src, size_t len)
{
    if (!h || (len > 0 && !src))
        return -1;
    Session *s = (Session *)h;
    std::lock_guard<std::mutex> lock(s->mtx);
    if (!s->alive)
        return -3;
    DeviceGuard g(s->device);
    return session_upload(s, src, len);
}

int cuda_replace_build_index(void *h, size_t chunk_size)
{
    if (!h)
        return -1;
    Session *s = (Session *)h;
    std::lock_guard<std::mutex> lock(s->mtx);
    if (!s->alive)
        return -3;
    DeviceGuard g(s->device);
    if (chunk_size == 0)
        chunk_size = 1u << 20;
    if (s->len == 0)
        return 0;
    const int64_t chunks = (int64_t)((s->len + chunk_size - 1) / chunk_size);
    size_t cap_bytes = s->presence_words * 8;
    CK(grow((void **)&s->d_presence, &cap_bytes, (size_t)chunks * 32, false, s->stream));
    s->presence_words = cap_bytes / 8;
    CK(launch_presence_index(s->d_buf, (int64_t)s->len, (int64_t)chunk_size, s->d_presence, chunks, s->stream));
    CK(hipStreamSynchronize(s->stream));
    return 0;
}

int cuda_replace_apply(void *h, const uint8_t *pat, int pat_len, const uint8_t *rep, int rep_len,
                       size_t *new_len)
{
    if (!h)
        return -1;
    Session *s = (Session *)h;
    std::lock_guard<std::mutex> lock(s->mtx);
    if (!s->alive)
        return -3;
    DeviceGuard g(s->device);
    return session_apply(s, pat, pat_len, rep, rep_len, 0, new_len, nullptr);
}

int cuda_replace_apply_seeded(void *h, const uint8_t *pat, int pat_len, const uint8_t *rep, int rep_len,
                              long long prev_last_rel, long long *out_last_rel, size_t *new_len)
{
    if (!h)
        return -1;
    Session *s = (Session *)h;
    std::lock_guard<std::mutex> lock(s->mtx);
    if (!s->alive)
        return -3;
    DeviceGuard g(s->device);
    const int64_t f = pat_len > 0 ? floor_from_seed(prev_last_rel, pat_len) : 0;
    return session_apply(s, pat, pat_len, rep, rep_len, f, new_len, out_last_rel);
}

int cr_query_prefix64(void *h, size_t limit, long long *kept_before_limit, long long *last_kept_before_limit)
{
    if (!h)
        return -1;
    Session *s = (Session *)h;
    std::lock_guard<std::mutex> lock(s->mtx);
    if (!s->alive)
        return -3;
    DeviceGuard g(s->device);
    int64_t c = 0, l = -1;
    int rc = session_query(s, limit, &c, &l);
    if (rc != 0)
        return rc;
    if (kept_before_limit)
        *kept_before_limit = c;
    if (last_kept_before_limit)
        *last_kept_before_limit = l;
    return 0;
}

int cuda_replace_query_prefix(void *h, size_t limit, int *kept_before_limit, int *last_kept_before_limit)
{
    long long c = 0, l = -1;
    int rc = cr_query_prefix64(h, limit, &c, &l);
    if (rc != 0)
        return rc;
    if (c > INT_MAX || l > INT_MAX)
        return -4;
    if (kept_before_limit)
        *kept_before_limit = (int)c;
    if (last_kept_before_limit)
        *last_kept_before_limit = (int)l;
    return 0;
}

int cuda_replace_apply_batch(void *h, const uint8_t **pats, const int *pat_lens, const uint8_t **reps,
                             const int *rep_lens, int count, size_t *new_len)
{
    return cr_apply_batch_ex(h, pats, pat_lens, reps, rep_lens, count, new_len, 1, nullptr);
}

int cr_apply_batch_ex(void *h, const uint8_t *const *pats, const int *pat_lens, const uint8_t *const *reps,
                      const int *rep_lens, int count, size_t *new_len, int allow_fuse, int *fused)
{
    if (fused)
        *fused = 0;
    if (!h)
        return -1;
    Session *s = (Session *)h;
    std::lock_guard<std::mutex> lock(s->mtx);
    if (!s->alive)
        return -3;
    DeviceGuard g(s->device);
    if (pats && pat_lens && count > 0) {   // no patterns: every pair is skipped
        int rc = session_apply_batch(s, pats, pat_lens, reps, rep_lens, count, allow_fuse != 0, fused);
        if (rc != 0)
            return rc;
    }
    if (new_len)
        *new_len = s->len;
    return 0;
}

int cr_batch_fusable(const uint8_t *const *pats, const int *pat_lens, const uint8_t *const *reps, const int *rep_lens,
                     int count)
{
    FusedTable *t = new (std::nothrow) FusedTable;
    if (!t)
        return -2;
    const bool ok = build_fused(pats, pat_lens, reps, rep_lens, count, t);
    delete t;
    return ok ? 1 : 0;
}

int cuda_replace_result(void *h, uint8_t **out_host, size_t *outlen_host)
{
    if (!h || !out_host || !outlen_host)
        return -1;
    Session *s = (Session *)h;
    std::lock_guard<std::mutex> lock(s->mtx);
    if (!s->alive)
        return -3;
    DeviceGuard g(s->device);
    *out_host = nullptr;
    *outlen_host = s->len;
    if (s->len == 0)
        return 0;
    void *hp = nullptr;
    hipError_t e = hipHostMalloc(&hp, s->len, hipHostMallocDefault);
    if (e == hipSuccess)
        e = hipMemcpyAsync(hp, s->d_buf, s->len, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) {
        if (hp)
            (void)hipHostFree(hp);
        *out_host = nullptr;
        *outlen_host = 0;
        return (int)e;
    }
    *out_host = (uint8_t *)hp;
    return 0;
}

void cuda_free_host(void *p)
{
    if (p)
        (void)hipHostFree(p);
}

void cuda_replace_close(void *h)
{
    if (!h)
        return;
    Session *s = (Session *)h;
    {
        std::lock_guard<std::mutex> lock(s->mtx);
        if (!s->alive)
            return;
        s->alive = false;
    }
    destroy(s);
}

int cuda_replace_unified(const uint8_t *src, size_t src_len, const uint8_t *pat, int pat_len, const uint8_t *rep,
                         int rep_len, uint8_t **out_host, size_t *out_len)
{
    if (!out_host || !out_len)
        return -1;
    *out_host = nullptr;
    *out_len = 0;
    void *h = nullptr;
    int rc = cuda_replace_open(&h, src, src_len);
    if (rc != 0)
        return rc;
    rc = cuda_replace_apply(h, pat, pat_len, rep, rep_len, out_len);
    if (rc == 0)
        rc = cuda_replace_result(h, out_host, out_len);
    cuda_replace_close(h);
    return rc;
}

int cr_session_length(void *h, size_t *len)
{
    if (!h || !len)
        return -1;
    Session *s = (Session *)h;
    std::lock_guard<std::mutex> lock(s->mtx);
    if (!s->alive)
        return -3;
    *len = s->len;
    return 0;
}

int cr_replace_device(void *h, const uint8_t *d_src, size_t n, uint8_t *d_dst, size_t dst_cap, const uint8_t *pat,
                      int pat_len, const uint8_t *rep, int rep_len, long long seed, size_t *out_len,
                      long long *out_last, float *kernel_ms)
{
    if (!h || (n > 0 && !d_src) || (pat_len > 0 && !pat) || (rep_len > 0 && !rep) || pat_len < 0 || rep_len < 0)
        return -1;
    if (((uintptr_t)d_src & 15u) || ((uintptr_t)d_dst & 15u))
        return (int)hipErrorInvalidValue;
    Session *s = (Session *)h;
    std::lock_guard<std::mutex> lock(s->mtx);
    if (!s->alive)
        return -3;
    DeviceGuard g(s->device);
    // the pass reuses the session's pattern buffer and tile records, so the record
    // of the session's last apply is dropped: a later prefix query answers 0 / -1
    // until the next apply (the session's buffer itself is untouched)
    s->q_valid = false;
    s->fused_pending = false;
    const int64_t f = pat_len > 0 ? floor_from_seed(seed, pat_len) : 0;
    PassResult res;
    auto dst_for = [&](int64_t need, uint8_t **dst) -> int {
        if ((size_t)need > dst_cap || (need > 0 && !d_dst))
            return (int)hipErrorInvalidValue;
        *dst = d_dst;
        return 0;
    };
    int rc = run_pass(s, d_src, (int64_t)n, pat, pat_len, rep, rep_len, f, (int64_t)dst_cap, out_last != nullptr,
                      dst_for, &res, kernel_ms, nullptr, nullptr);
    if (out_len)
        *out_len = (size_t)res.out_len;
    if (rc != 0)
        return rc;
    if (out_last)
        *out_last = res.last;
    return 0;
}

int cr_autotune_set(const uint8_t *pat, int pat_len, size_t n, int group_tiles)
{
    if (pat_len < 0 || (pat_len > 0 && !pat))
        return -1;
    int lt = 0;
    const int c = tune_class(pattern_mode(pat, pat_len, &lt), lt);
    // border-free patterns: 1, 2 or 4 pin the 32-byte-per-thread form with that group
    // size, 8 the three-launch 64-byte form, 16 the single launch, 32 the single launch
    // that holds each group in LDS
    const bool bf_codes = c == 0 && (group_tiles == 8 || group_tiles == 16 || group_tiles == 32);
    if (c < 0 || !(group_tiles == 0 || group_tiles == 1 || group_tiles == 2 || group_tiles == 4 || bf_codes))
        return -1;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess)
        return (int)e;
    Tuner &t = tuner();
    std::lock_guard<std::mutex> lk(t.mtx);
    const auto key = std::make_tuple(dev, c, size_bucket((int64_t)n));
    if (group_tiles == 0) {
        t.best.erase(key);
    } else if (c == 0) {
        TuneEntry te;
        te.gt = kGroupTiles;
        te.variant = group_tiles == 32 ? kBfKeep : group_tiles == 16 ? kBfSweep : group_tiles == 8 ? kBfLean
                   : group_tiles == 1 ? kBfTiles1 : group_tiles == 2 ? kBfTiles2 : kBfTiles4;
        t.best[key] = te;
    } else {
        TuneEntry te;
        te.gt = group_tiles;
        t.best[key] = te;
    }
    return 0;
}

void cr_autotune_reset(void)
{
    Tuner &t = tuner();
    std::lock_guard<std::mutex> lk(t.mtx);
    t.best.clear();
}

int cr_autotune_report(char *buf, size_t cap)
{
    static const char *kClassName[5] = {"border_free", "table2", "table4", "table8", "table16"};
    static const char *kBucket[5] = {"<4MiB", "4-32MiB", "32-256MiB", "256MiB-2GiB", ">=2GiB"};
    std::string out = "[";
    {
        Tuner &t = tuner();
        std::lock_guard<std::mutex> lk(t.mtx);
        bool first = true;
        static const char *kVariant[kBfVariants] = {"single_launch", "three_launch_64B", "tiles1_32B",
                                                    "tiles2_32B", "tiles4_32B", "single_launch_lds"};
        for (const auto &kv : t.best) {
            char line[512];
            const TuneEntry &te = kv.second;
            if (te.variant >= 0) {
                auto f = [](float v) { return v >= 1e29f ? -1.0 : (double)v; };   // -1: gave up
                snprintf(line, sizeof line,
                         "%s{\"device\": %d, \"class\": \"%s\", \"size\": \"%s\", \"form\": \"%s\", "
                         "\"ms\": {\"single_launch\": %.4f, \"three_launch_64B\": %.4f, \"tiles1_32B\": %.4f, "
                         "\"tiles2_32B\": %.4f, \"tiles4_32B\": %.4f, \"single_launch_lds\": %.4f}}",
                         first ? "" : ", ", std::get<0>(kv.first), kClassName[std::get<1>(kv.first)],
                         kBucket[std::get<2>(kv.first)], kVariant[te.variant], f(te.vms[0]), f(te.vms[1]),
                         f(te.vms[2]), f(te.vms[3]), f(te.vms[4]), f(te.vms[5]));
            } else if (te.single >= 0) {
                const double sm = te.single_ms >= 1e29f ? -1.0 : (double)te.single_ms;   // -1: gave up
                snprintf(line, sizeof line,
                         "%s{\"device\": %d, \"class\": \"%s\", \"size\": \"%s\", \"form\": \"%s\", "
                         "\"group_tiles\": %d, \"ms\": {\"1\": %.4f, \"2\": %.4f, \"4\": %.4f, "
                         "\"single_launch_lds\": %.4f}}",
                         first ? "" : ", ", std::get<0>(kv.first), kClassName[std::get<1>(kv.first)],
                         kBucket[std::get<2>(kv.first)], te.single ? "single_launch_lds" : "three_launch", te.gt,
                         te.ms[0], te.ms[1], te.ms[2], sm);
            } else {
                snprintf(line, sizeof line,
                         "%s{\"device\": %d, \"class\": \"%s\", \"size\": \"%s\", \"group_tiles\": %d, "
                         "\"ms\": {\"1\": %.4f, \"2\": %.4f, \"4\": %.4f}}",
                         first ? "" : ", ", std::get<0>(kv.first), kClassName[std::get<1>(kv.first)],
                         kBucket[std::get<2>(kv.first)], te.gt, te.ms[0], te.ms[1], te.ms[2]);
            }
            out += line;
            first = false;
        }
    }
    out += "]";
    if (buf && cap > 0) {
        const size_t k = out.size() < cap - 1 ? out.size() : cap - 1;
        memcpy(buf, out.data(), k);
        buf[k] = 0;
    }
    return (int)out.size();
}

int cr_pattern_mode(const uint8_t *pat, int pat_len)
{
    if (pat_len < 0 || (pat_len > 0 && !pat))
        return -1;
    int m = pattern_mode(pat, pat_len, nullptr);
    return m == kModeEmpty ? 0 : m;
}

} // extern "C"
