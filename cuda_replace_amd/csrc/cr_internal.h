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
constexpr int kHalo = 256;                    // LDS halo on each side of a group
constexpr int kMaxSinglePassPat = kHalo;      // chunked kernels take L <= 256
constexpr int kMaxTablePat = 16;              // carry tables (self-overlapping patterns) take L <= 16
constexpr int kStage = kTile + kTile / 8 + 64; // LDS image of a tile's output (bytes)

// A group is kGroupTiles consecutive tiles handled by one workgroup; a chunk is
// a run of consecutive groups handled by one count, one spine and one emit launch.
constexpr int kGroupTiles = 2;                    // default tiles per group (1, 2 or 4 may run)
constexpr int kMaxGroupTiles = 4;
constexpr int kGroupBytes = kGroupTiles * kTile;    // 16 KiB of source per workgroup

// Stats words carry the apply's 16-bit epoch in their top bits, so a stale word
// from an earlier apply is told apart from this one's.
constexpr int kEpochShift = 48;
// 64-bit carry tables used while composing group tables: carry << 40 | count
constexpr int kCarryShift = 40;
constexpr uint64_t kCountMask = (1ull << 40) - 1;

// Per-tile record kept after an apply so a later prefix query can be answered.
struct TileInfo {
    int64_t k_excl;     // kept starts before the tile
    int64_t last_abs;   // last kept start inside the tile, -1 if none
    int32_t carry_in;   // leading tile bytes covered by an earlier kept match
    int32_t mode;       // apply mode that wrote the record
};

// Kept state at a group or chunk boundary (the spine writes one per group and one
// after each chunk).
struct ChunkState {
    int64_t k_base;     // kept starts before the chunk
    int32_t carry;      // leading bytes of the chunk covered by an earlier kept match
    int32_t pad;
};

enum Mode : int32_t {
    kModeBorderFree = 1,   // pattern cannot overlap itself: every candidate >= F is kept
    kModeTable = 2,        // self-overlapping, L <= 16: per-carry-state tables
    kModeBitmap = 3,       // general: candidate bitmap, serial greedy, scatter
    kModeFused = 5,        // several pairs in one pass (see FusedTable)
};

// ---- fused multi-pattern pass ----------------------------------------------------
// Several (pattern, replacement) pairs applied in one pass.  The host only builds a
// table when that equals applying the pairs one after another: the patterns are
// border-free, distinct, none a substring of another and no proper suffix of one a
// prefix of another (so no two occurrences of any of them overlap and every
// occurrence is kept), every replacement is non-empty and made of bytes that occur in
// no pattern (so no replacement forms or joins a later occurrence), and all patterns
// share their first byte (the candidate filter).
constexpr int kMaxFusedPats = 256;
constexpr int kMaxFusedPat = 64;       // longest pattern or replacement in a fused pass
constexpr int kFusedPool = 16384;      // pattern + replacement bytes
constexpr int kFusedHashBits = 10;
constexpr int kFusedSlots = 1 << kFusedHashBits;

// A slot is 16 bytes, loaded at once: x, y = the pattern's bytes 1..8 (little-endian,
// zero past the pattern), z = meta (pattern index in bits 0-7, pattern length in
// 8-15, replacement length in 16-23, bit 31 set; 0 = empty slot), w = the
// replacement's first 4 bytes (zero padded).  A pattern of up to 9 bytes is verified
// and a replacement of up to 4 bytes written from the slot alone.
// A probe hit is the 64-bit entry meta | rep4 << 32.
constexpr uint32_t kFusedUsed = 1u << 31;
constexpr int kFusedSlotBytes = 8;     // pattern bytes after the first held in a slot

struct FusedTable {
    int32_t count;
    int32_t key_len;                   // pattern bytes after the first one that are hashed (1..4)
    int32_t min_len;
    int32_t max_len;
    uint32_t key_mask;                 // low key_len bytes set
    uint32_t pad[3];
    uint4 slot[kFusedSlots];           // linear probing on the hash of the key bytes
    uint4 head[kMaxFusedPats];         // first 16 bytes of each pattern, zero padded
    uint4 rep_head[kMaxFusedPats];     // first 16 bytes of each replacement, zero padded
    uint16_t pat_len[kMaxFusedPats];
    uint16_t rep_len[kMaxFusedPats];
    uint16_t pat_off[kMaxFusedPats];   // into pool
    uint16_t rep_off[kMaxFusedPats];
    uint8_t pool[kFusedPool + 16];     // 16 bytes of slack: pattern bytes are compared 16 at a time
};

__host__ __device__ inline int fused_entry_index(uint64_t e) { return (int)(e & 0xff); }
__host__ __device__ inline int fused_entry_pat_len(uint64_t e) { return (int)((e >> 8) & 0xff); }
__host__ __device__ inline int fused_entry_rep_len(uint64_t e) { return (int)((e >> 16) & 0xff); }
__host__ __device__ inline uint32_t fused_entry_rep4(uint64_t e) { return (uint32_t)(e >> 32); }

// Hash of the key bytes after the shared first byte (multiplicative, top bits).
__host__ __device__ inline uint32_t fused_hash(uint32_t key)
{
    return (key * 2654435761u) >> (32 - kFusedHashBits);
}

// stats words written by a pass
struct LaunchStats {
    uint64_t k_total;      // epoch-tagged total kept count
    uint64_t last_kept;    // epoch-tagged (last kept + 1), written by the bitmap greedy
    uint64_t pad[2];
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
    uint32_t qref[8];      // masked-SAD reference words: pattern bytes 4q..4q+3 xor qx's byte,
                           // 0 past the pattern (L <= 32; the SAD skips zero reference bytes)
    uint32_t qx;           // a byte value absent from the pattern, in all four bytes
    uint32_t *gsum;        // per-group kept count, or LT table entries per group
    ChunkState *chunks;    // num_chunks + 1 chunk states (or spine block states)
    uint64_t *block_tabs;  // spine blocks: each block's summary (kMaxTablePat words)
    ChunkState *chunks_pre; // per group: kept starts before it, carry into it
    TileInfo *info;        // per-tile records (num_tiles)
    uint32_t *bitmap;      // candidate / kept bitmap (bitmap mode)
    uint32_t *masks;       // candidate mask of every 32-byte segment (tile * 256 + segment),
                           // written by the count and read by the emit; null: the emit
                           // recomputes them
    LaunchStats *stats;
    uint32_t epoch;        // 1..65535
    int32_t write_output;  // 0: count only
    int64_t num_tiles;
    int64_t num_groups;
};

// ---- batches of independent documents ----------------------------------------
// One document of a batch round: its current bytes, its destination and the pair
// the round applies to it.
struct BatchDoc {
    const uint8_t *src;    // device, 16-byte aligned
    uint8_t *dst;          // device, 16-byte aligned (set once the round's sizes are known)
    int64_t n;             // current length
    int64_t group0;        // first of the document's groups in its class's group list
    int32_t ngroups;
    int32_t L;
    int32_t r;
    uint32_t pat4;
    uint32_t mask4;
    int32_t pat_off;       // pattern bytes at pool + pat_off, replacement at pool + rep_off
    int32_t rep_off;
    int32_t pad;
};

// One workgroup of a batched count or emit launch: a group of one document.
struct BatchGroup {
    int32_t doc;           // index into the round's BatchDoc array
    int32_t g;             // group index within the document
};

struct BatchArgs {
    const BatchDoc *docs;
    const BatchGroup *groups;   // this class's groups, document by document
    const int32_t *class_docs;  // this class's documents (spine: one wave each)
    const uint8_t *pool;        // patterns and replacements of the round
    uint32_t *gsum;             // one summary per group of the class
    ChunkState *pre;            // one state per group of the class
    int64_t *doc_kept;          // kept starts per document of the round
};

// gt: tiles per group (1, 2 or 4), the same for a round's count and emit
hipError_t launch_batch_count(const BatchArgs &a, int mode, int table_size, int gt, int64_t groups, hipStream_t s);
hipError_t launch_batch_spine(const BatchArgs &a, int mode, int table_size, int64_t docs, hipStream_t s);
hipError_t launch_batch_emit(const BatchArgs &a, int mode, int table_size, int gt, int64_t groups, hipStream_t s);

// ---- device-driven batch: every round planned and placed on the GPU ----------------
// Lengths never come back to the host between rounds, so a whole run is a fixed
// launch sequence (captured once as a HIP graph and replayed).  Per round: plan
// (one workgroup: each active document's descriptor and group range; the round's
// documents, grouped by class, come from the host, which knows every document's
// pairs), expand (the class work lists of (document, group) items), per class a
// count over its work list (one workgroup per item up to the host's estimate, a
// persistent grid for any items beyond it) and a persistent spine, place (one workgroup: new lengths and destinations, packed in
// the ping-pong arena or the final arena), per class a persistent emit, and the
// empty-pattern documents.  A plan or placement that would not fit the buffers
// commits nothing and records the round and the sizes it needs; every later kernel
// of the run returns at once, and the host grows the buffers and resumes there.
constexpr int kBatchClasses = 5;       // border-free, carry tables of 2/4/8/16 states
constexpr int kBatchEmpty = 5;         // class index of empty-pattern pairs

struct BatchPair {
    int32_t L;
    int32_t r;
    uint32_t pat4;
    uint32_t mask4;
    int32_t pat_off;       // into the run's pool
    int32_t rep_off;
    int32_t cls;           // 0..4 counted classes, kBatchEmpty
    int32_t pad;
};

struct BatchDocState {
    int64_t len;           // current length
    int64_t off;           // offset of the current bytes in the round's source arena
    int64_t fin_off;       // offset in the final arena once the last pair ran
    int32_t pair0;         // first pair of the document
    int32_t npairs;
};

struct BatchCtl {
    int64_t class_groups[kBatchClasses];
    int64_t fin_used;                   // final arena bytes handed out so far
    int64_t algo_bytes;                 // input + output bytes of every applied pair
    int64_t need_groups;                // on overflow: work-list entries the round needs
    int64_t need_arena;                 //   destination arena bytes
    int64_t need_fin;                   //   final arena bytes (all documents so far)
    int32_t overflow;                   // 0, or 1 + the round that did not fit
    int32_t pad;
};

struct BatchRound {
    BatchCtl *ctl;
    BatchDocState *states;
    const BatchPair *pairs;
    const uint8_t *pool;
    BatchDoc *docs;                         // indexed by document
    BatchGroup *groups[kBatchClasses];      // work lists, group_cap entries each
    const int32_t *rlist;                   // the round's active documents, grouped by class
    int32_t rcount;                         // entries of rlist
    int32_t cseg[kBatchClasses + 2];        // class c's documents are rlist[cseg[c], cseg[c + 1])
    uint32_t *gsum[kBatchClasses];          // group summaries (group_cap x words)
    ChunkState *pre[kBatchClasses];         // group states (group_cap)
    int64_t *doc_kept;
    const uint8_t *src_arena;
    uint8_t *dst_arena;
    uint8_t *fin;
    int64_t dst_cap;
    int64_t fin_cap;
    int64_t group_cap;
    int32_t num_docs;
    int32_t round;
    int32_t gt;                             // tiles per group
    int32_t pad;
};

struct BatchRoundShape {
    int present[kBatchClasses + 1];         // class has documents this round
    int class_docs[kBatchClasses + 1];      // exact, known from the pairs
    int work_grid[kBatchClasses];           // estimated work items per class: one workgroup each
    int cus;                                // compute units of the device
};

// One whole round of a device-driven batch on stream s (capturable).
hipError_t launch_batch_round(const BatchRound &a, const BatchRoundShape &shape, hipStream_t s);

hipError_t launch_fused(const ReplaceParams &p, const FusedTable *t, int which, int64_t first_group,
                        int64_t groups, hipStream_t s);   // which: 0 count, 1 spine, 2 emit

// ---- launchers (cr_kernels.hip) --------------------------------------------
// One chunk of groups [first_group, first_group + groups): count, spine, emit.
// gt: tiles per group (1, 2 or 4; the bitmap emit only kGroupTiles); first_group and
// groups count groups of that size.
hipError_t launch_count(const ReplaceParams &p, int mode, int table_size, int gt, int64_t first_group,
                        int64_t groups, hipStream_t s);
// The spine of a whole pass (one chunk) as three launches: per block of groups its
// summary, one wave scanning the block summaries (block states in chunks[], the
// pass total in stats), then every block's groups in parallel.
hipError_t launch_spine_blocks(const ReplaceParams &p, int mode, int table_size, int64_t groups, hipStream_t s);
int64_t spine_blocks_needed(int64_t groups);   // chunks[] entries and block_tabs blocks it uses
hipError_t launch_spine(const ReplaceParams &p, int mode, int table_size, int64_t first_group,
                        int64_t groups, int64_t chunk, int last_chunk, hipStream_t s);
hipError_t launch_emit(const ReplaceParams &p, int mode, int table_size, int gt, int64_t first_group,
                       int64_t groups, hipStream_t s);
// The lean border-free pass over 16 KiB groups (kGroupTiles tiles): count, then the
// border-free spine, then emit.
hipError_t launch_bf_count(const ReplaceParams &p, int64_t first_group, int64_t groups, hipStream_t s);
hipError_t launch_bf_emit(const ReplaceParams &p, int64_t first_group, int64_t groups, hipStream_t s);
// The whole border-free pass in one persistent launch (count ahead, scanner, emit).
// sweep: sweep_words(num_groups) 64-bit words holding no word tagged with p.epoch;
// the kept total lands in p.stats (tagged), or sweep[2 * groups] = epoch << 48 | 1
// when a wait gave up (then nothing of the pass is valid: run the three launches).
hipError_t launch_bf_sweep(const ReplaceParams &p, uint64_t *sweep, int ahead, int spin_limit, bool keep,
                           hipStream_t s);
// the held-group single launch of a 2-state carry-table pass (p.num_groups groups of
// 16 KiB); same sweep words and give-up protocol as launch_bf_sweep
hipError_t launch_tbl_sweep(const ReplaceParams &p, uint64_t *sweep, int spin_limit, hipStream_t s);
// the held-group single launch of a fused multi-pattern pass (16 KiB groups)
hipError_t launch_fused_sweep(const ReplaceParams &p, const FusedTable *t, uint64_t *sweep, int spin_limit,
                              hipStream_t s);
int64_t sweep_words(int64_t groups);
hipError_t launch_bitmap_mark(const ReplaceParams &p, hipStream_t s);
hipError_t launch_bitmap_greedy(const ReplaceParams &p, hipStream_t s);
// Parallel bitmap selection (candidates, then per-tile tables composed level by level,
// then every tile resolved): bytes of scratch it needs for a pass, 0 when the serial
// greedy must run (a self-overlapping pattern over 256 bytes).  It also marks the
// candidates, and leaves the kept bitmap, the tile records and the stats as the
// serial greedy does.
size_t bitmap_select_bytes(int64_t num_tiles, int L, bool border_free);
hipError_t launch_bitmap_select(const ReplaceParams &p, bool border_free, uint8_t *scratch, hipStream_t s);
hipError_t launch_empty_pattern(const ReplaceParams &p, hipStream_t s);
hipError_t launch_fill_u64(uint64_t *ptr, uint64_t value, size_t count, hipStream_t s);
hipError_t launch_query(const ReplaceParams &p, int mode, int64_t limit,
                        int64_t *out2, hipStream_t s);
hipError_t launch_last_kept(const ReplaceParams &p, int64_t *out1, hipStream_t s);
hipError_t launch_presence_index(const uint8_t *src, int64_t n, int64_t chunk,
                                 uint64_t *masks, int64_t num_chunks, hipStream_t s);

} // namespace cr
