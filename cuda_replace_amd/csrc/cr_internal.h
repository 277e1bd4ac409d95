// cr_internal.h -- types shared by the HIP kernels and the host runtime of the
// byte-pattern replace library (gfx950).
//
// Replace semantics (leftmost, non-overlapping, bytes.replace):
//   candidates  i in [0, n-L] with src[i..i+L) == pat          (cuda_replace.cu:119-129)
//   kept        greedy ascending: keep i iff i - last >= L        (cuda_replace.cu:938-968)
//               a seed "last" only blocks candidates i < seed+L, so it is carried as a
//               floor F = max(0, seed + L) below which no candidate is kept
//   output      kept start -> rep, its L bytes skipped; other bytes copied
//               (cuda_replace.cu:1391-1467, 1520-1559); out_len = n + |K| * (r - L)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cr {

// ---- tiling ---------------------------------------------------------------
constexpr int kThreads = 256;                 // 4 wave64 per workgroup
constexpr int kBytesPerThread = 32;           // one thread owns 32 consecutive source bytes
constexpr int kTile = kThreads * kBytesPerThread;   // 8192 source bytes per tile
constexpr int kHalo = 256;                    // LDS halo on each side of a tile
constexpr int kMaxSinglePassPat = kHalo;      // single-pass kernels take L <= 256
constexpr int kMaxTablePat = 16;              // carry tables (self-overlapping patterns) take L <= 16
constexpr int kStage = kTile + kTile / 8 + 64; // LDS stage for a tile's output (bytes)

// ---- decoupled look-back state words ----------------------------------------
// One 64-bit word per published value, written and read with relaxed agent-scope
// atomics, so flag and payload travel together:
//   [63:48] epoch  [47:46] status  [45:0] payload
// A word whose epoch is not the launch's epoch reads as "not ready", so the state
// array never needs clearing between launches (it is cleared when the 16-bit epoch
// wraps).
constexpr uint64_t kStatusAgg = 1;   // tile aggregate published
constexpr uint64_t kStatusInc = 2;   // inclusive prefix published
constexpr int kEpochShift = 48;
constexpr int kStatusShift = 46;
constexpr uint64_t kPayloadMask = (1ull << 46) - 1;
// table entries: payload = carry(6 bits) << 40 | count(40 bits)
constexpr int kCarryShift = 40;
constexpr uint64_t kCountMask = (1ull << 40) - 1;

// Per-tile record kept after an apply so a later prefix query can be answered.
struct TileInfo {
    int64_t k_excl;     // kept starts before the tile
    int64_t last_abs;   // last kept start inside the tile, -1 if none
    int32_t carry_in;   // leading tile bytes covered by an earlier kept match
    int32_t mode;       // apply mode that wrote the record
};

enum Mode : int32_t {
    kModeBorderFree = 1,   // pattern cannot overlap itself: every candidate >= F is kept
    kModeTable = 2,        // self-overlapping, L <= 16: carry-state tables + look-back
    kModeBitmap = 3,       // general: candidate bitmap, serial greedy, scatter
};

// stats words written by a launch
struct LaunchStats {
    uint64_t k_total;      // epoch-tagged total kept count
    uint64_t last_kept;    // epoch-tagged (last kept + 1), written by the bitmap greedy
    uint64_t pad[2];       // pad[1]: epoch-tagged abort flag of a look-back launch
};

struct ReplaceParams {
    const uint8_t *src;
    uint8_t *dst;
    int64_t n;             // source length
    int64_t floor;         // no candidate below this position is kept
    int64_t out_len;       // exact output length (scatter kernels) or -1
    const uint8_t *pat;    // device copy, L bytes
    const uint8_t *rep;    // device copy, r bytes
    int32_t L;
    int32_t r;
    uint32_t pat4;         // first min(L,4) pattern bytes, little-endian
    uint32_t mask4;        // byte mask of those bytes
    uint64_t *state;       // look-back words
    TileInfo *info;        // per-tile records (num_tiles)
    uint32_t *bitmap;      // candidate / kept bitmap (bitmap mode)
    LaunchStats *stats;
    uint32_t epoch;        // 1..65535
    int32_t write_output;  // 0: count only
    int64_t num_tiles;
};

// ---- launchers (cr_kernels.hip) --------------------------------------------
hipError_t launch_border_free(const ReplaceParams &p, hipStream_t s);
hipError_t launch_table(const ReplaceParams &p, int table_size, hipStream_t s);
hipError_t launch_bitmap_mark(const ReplaceParams &p, hipStream_t s);
hipError_t launch_bitmap_greedy(const ReplaceParams &p, hipStream_t s);
hipError_t launch_bitmap_scatter(const ReplaceParams &p, hipStream_t s);
hipError_t launch_empty_pattern(const ReplaceParams &p, hipStream_t s);
hipError_t launch_fill_u64(uint64_t *ptr, uint64_t value, size_t count, hipStream_t s);
hipError_t launch_query(const ReplaceParams &p, int mode, int64_t limit,
                        int64_t *out2, hipStream_t s);
hipError_t launch_last_kept(const ReplaceParams &p, int64_t *out1, hipStream_t s);
hipError_t launch_presence_index(const uint8_t *src, int64_t n, int64_t chunk,
                                 uint64_t *masks, int64_t num_chunks, hipStream_t s);

} // namespace cr
