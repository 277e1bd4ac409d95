// cr_noalloc.hip -- the device-to-device replace pass without any allocation.
//
// cr_replace_device_noalloc runs the same count / spine / emit kernels as
// cr_replace_device, but its scratch (group summaries, group states, spine block
// states and summaries, the stats words, the pattern and replacement bytes) lives
// in one file-scope device array, `scratch`, instead of session buffers, and it
// uses the caller's stream.  It is for callers that cannot allocate: code in tight
// serving loops, harnesses that forbid hipMalloc.  It calls no hipMalloc, hipFree
// or other allocating routine.
//
// Limits (the array is sized for them): input up to kNoAllocMaxBytes (4 GiB),
// pattern and replacement up to 256 bytes each, and border-free or carry-table
// patterns (a self-overlapping pattern longer than 16 bytes needs the bitmap path,
// whose bitmap is n/8 bytes).  Larger requests return -5.
//
// Every GPU holds its own copy of the array (about 20 MiB) from module load until
// the process ends.  Calls on one GPU share it and keep the last call's contents,
// so calls that overlap on one GPU (from several threads or streams) race.
//
// No counterpart in the reference (cuda_replace.cu allocates per context).
#include "cr_internal.h"
#include "cr_pattern.h"
#include "../../include/cuda_replace.h"

#include <atomic>
#include <mutex>
#include <string.h>

using namespace cr;

namespace {

constexpr int64_t kNoAllocMaxBytes = (int64_t)1 << 32;
constexpr int kNoAllocMaxPat = kMaxSinglePassPat;   // 256
constexpr int kNoAllocMaxRep = 256;
constexpr int kNoAllocGroupTiles = kGroupTiles;     // 16 KiB groups, the default
constexpr int64_t kNaGroups = kNoAllocMaxBytes / ((int64_t)kTile * kNoAllocGroupTiles);   // 262144
constexpr int64_t kNaBlocks = kNaGroups / 1024 + 3;   // spine blocks (>= spine_blocks_needed + 1)

constexpr size_t al256(size_t v) { return (v + 255) / 256 * 256; }

// byte layout of the scratch array
constexpr size_t kOffGsum = 0;                                                        // summaries
constexpr size_t kOffPre = kOffGsum + al256((size_t)kNaGroups * kMaxTablePat * 4);   // group states
constexpr size_t kOffChunks = kOffPre + al256((size_t)kNaGroups * sizeof(ChunkState));
constexpr size_t kOffTabs = kOffChunks + al256((size_t)kNaBlocks * sizeof(ChunkState));
constexpr size_t kOffStats = kOffTabs + al256((size_t)kNaBlocks * kMaxTablePat * sizeof(uint64_t));
constexpr size_t kOffPatRep = kOffStats + al256(sizeof(LaunchStats));
constexpr size_t kScratchBytes = kOffPatRep + al256(kNoAllocMaxPat + kNoAllocMaxRep);

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

// Pattern and replacement bytes travel as a kernel argument, so no host staging
// buffer or copy from pageable memory is needed.
struct PatRep {
    uint8_t b[kNoAllocMaxPat + kNoAllocMaxRep];
};

__global__ void put_patrep_kernel(PatRep a, int len, uint8_t *dst)
{
    for (int i = threadIdx.x; i < len; i += blockDim.x)
        dst[i] = a.b[i];
}

// The array's address on each device, looked up once per device with that device
// current (in host code the name `scratch` is only a stand-in, never an address).
struct DevScratch {
    std::once_flag once;
    hipError_t err = hipSuccess;
    uint8_t *ptr = nullptr;
    std::atomic<uint32_t> calls{0};
};
constexpr int kMaxDevices = 64;
DevScratch g_dev[kMaxDevices];

#define NA_CK(expr)                              \
    do {                                         \
        hipError_t e_ = (expr);                  \
        if (e_ != hipSuccess)                    \
            return (int)e_;                      \
    } while (0)

} // namespace

extern "C" {

CR_API int cr_replace_device_noalloc(const uint8_t *d_src, size_t n, uint8_t *d_dst, size_t dst_cap,
                                     const uint8_t *pat, int pat_len, const uint8_t *rep, int rep_len,
                                     long long seed, void *stream, size_t *out_len)
{
    if (!out_len || (n > 0 && !d_src) || (pat_len > 0 && !pat) || (rep_len > 0 && !rep) || pat_len < 0 ||
        rep_len < 0)
        return -1;
    *out_len = 0;
    if (((uintptr_t)d_src & 15u) || ((uintptr_t)d_dst & 15u))
        return (int)hipErrorInvalidValue;
    if ((int64_t)n > kNoAllocMaxBytes || n > (size_t)kNoAllocMaxBytes || pat_len > kNoAllocMaxPat ||
        rep_len > kNoAllocMaxRep)
        return -5;
    int table_size = 0;
    const int mode = pattern_mode(pat, pat_len, &table_size);
    if (mode == kModeBitmap)
        return -5;

    int dev = 0;
    NA_CK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        return (int)hipErrorInvalidDevice;
    DevScratch &ds = g_dev[dev];
    std::call_once(ds.once, [&] { ds.err = hipGetSymbolAddress((void **)&ds.ptr, HIP_SYMBOL(scratch)); });
    if (ds.err != hipSuccess)
        return (int)ds.err;
    uint8_t *const base = ds.ptr;
    hipStream_t s = (hipStream_t)stream;

    const int L = pat_len, r = rep_len;
    const int64_t N = (int64_t)n;
    // the stats words are tagged with a 16-bit epoch; the first call on a device
    // and every wrap of the epoch clear them (with a kernel, not a memset)
    const uint32_t call = ds.calls.fetch_add(1) % 65535u;
    const uint32_t epoch = call + 1;
    LaunchStats *stats = reinterpret_cast<LaunchStats *>(base + kOffStats);
    if (call == 0)
        NA_CK(launch_fill_u64(reinterpret_cast<uint64_t *>(stats), 0, sizeof(LaunchStats) / 8, s));

    if (L + r > 0) {
        PatRep a;
        memset(&a, 0, sizeof a);
        if (L > 0)
            memcpy(a.b, pat, (size_t)L);
        if (r > 0)
            memcpy(a.b + L, rep, (size_t)r);
        hipLaunchKernelGGL(put_patrep_kernel, dim3(1), dim3(256), 0, s, a, L + r, base + kOffPatRep);
        NA_CK(hipGetLastError());
    }

    ReplaceParams p{};
    p.src = d_src;
    p.dst = d_dst;
    p.n = N;
    p.out_len = -1;
    p.pat = base + kOffPatRep;
    p.rep = base + kOffPatRep + L;
    p.L = L;
    p.r = r;
    p.stats = stats;
    p.epoch = epoch;
    p.write_output = 1;

    if (mode == kModeEmpty) {
        const int64_t out = N + (N + 1) * (int64_t)r;
        *out_len = (size_t)out;
        if ((size_t)out > dst_cap || (out > 0 && !d_dst))
            return (int)hipErrorInvalidValue;
        if (r == 0) {
            if (N > 0)
                NA_CK(hipMemcpyAsync(d_dst, d_src, n, hipMemcpyDeviceToDevice, s));
        } else {
            NA_CK(launch_empty_pattern(p, s));
        }
        NA_CK(hipStreamSynchronize(s));
        return 0;
    }
    if (N == 0)
        return 0;

    // seed -> floor, as cuda_replace_apply_seeded
    int64_t floor_pos = 0;
    if (seed > (long long)(INT64_MAX - L))
        floor_pos = INT64_MAX;
    else
        floor_pos = (int64_t)seed + L < 0 ? 0 : (int64_t)seed + L;
    p.floor = floor_pos;
    pattern_head(pat, L, &p.pat4, &p.mask4);
    pattern_qsad(pat, L, p.qref, &p.qx);
    const int64_t tiles = (N + kTile - 1) / kTile;
    const int64_t groups = (tiles + kNoAllocGroupTiles - 1) / kNoAllocGroupTiles;
    p.num_tiles = tiles;
    p.num_groups = groups;
    p.gsum = reinterpret_cast<uint32_t *>(base + kOffGsum);
    p.chunks_pre = reinterpret_cast<ChunkState *>(base + kOffPre);
    p.chunks = reinterpret_cast<ChunkState *>(base + kOffChunks);
    p.block_tabs = reinterpret_cast<uint64_t *>(base + kOffTabs);
    p.info = nullptr;     // no per-tile records: no prefix query follows this pass
    p.bitmap = nullptr;
    p.masks = nullptr;    // the emit recomputes the candidate masks

    // the same kernels as cr_replace_device: the lean pass for border-free patterns
    auto count = [&](hipStream_t st) {
        return mode == kModeBorderFree ? launch_bf_count(p, 0, groups, st)
                                       : launch_count(p, mode, table_size, kNoAllocGroupTiles, 0, groups, st);
    };
    auto emit = [&](hipStream_t st) {
        return mode == kModeBorderFree ? launch_bf_emit(p, 0, groups, st)
                                       : launch_emit(p, mode, table_size, kNoAllocGroupTiles, 0, groups, st);
    };
    auto read_kept = [&](int64_t *kept) -> int {
        LaunchStats h;
        NA_CK(hipMemcpyAsync(&h, stats, sizeof h, hipMemcpyDeviceToHost, s));
        NA_CK(hipStreamSynchronize(s));
        if ((uint32_t)(h.k_total >> kEpochShift) != epoch)
            return (int)hipErrorUnknown;
        *kept = (int64_t)(h.k_total & ((1ull << kEpochShift) - 1));
        return 0;
    };

    const int64_t diff = (int64_t)r - L;
    int64_t need = N;
    if (diff > 0) {
        const int64_t span = floor_pos < N ? N - floor_pos : 0;
        need = N + (span / L) * diff;
    }
    int64_t kept = 0;
    if (need <= (int64_t)dst_cap && d_dst) {
        NA_CK(count(s));
        NA_CK(launch_spine_blocks(p, mode, table_size, groups, s));
        NA_CK(emit(s));
        const int rc = read_kept(&kept);
        if (rc != 0)
            return rc;
        *out_len = (size_t)(N + kept * diff);
        return 0;
    }
    if (diff <= 0) {
        *out_len = (size_t)N;
        return (int)hipErrorInvalidValue;
    }
    // expansion whose worst case exceeds the destination: count first
    NA_CK(count(s));
    NA_CK(launch_spine_blocks(p, mode, table_size, groups, s));
    int rc = read_kept(&kept);
    if (rc != 0)
        return rc;
    const int64_t exact = N + kept * diff;
    *out_len = (size_t)exact;
    if (exact > (int64_t)dst_cap || !d_dst)
        return (int)hipErrorInvalidValue;
    NA_CK(emit(s));
    NA_CK(hipStreamSynchronize(s));
    return 0;
}

} // extern "C"
