/*
 * cuda_replace.h -- C ABI of the MI355X (gfx950) byte-pattern replace library
 * (libcuda_replace.so).
 *
 * The first twelve entry points keep the names, signatures, ownership rules and
 * return codes of the reference library so its ctypes wrapper binds this .so
 * unchanged (reference: cuda_replace_wrapper.py:48-121 binds them).  The cr_*
 * entry points are extensions with no counterpart in the reference.
 *
 * Semantics: leftmost, non-overlapping, binary-safe replacement identical to
 * Python's bytes.replace (cuda_replace.cu:2-3), including the empty pattern
 * (the replacement is inserted at all n+1 gaps; the reference leaves the input
 * unchanged there).  Offsets and lengths are 64-bit throughout.
 *
 * Return codes: 0 ok; -1 null handle or argument; -2 session allocation failed
 * (open); -3 closed handle; -4 a result does not fit the ABI's int; -5 beyond
 * the limits of cr_replace_device_noalloc; any
 * positive value is a hipError_t (hipErrorInvalidValue=1, hipErrorOutOfMemory=2
 * and hipErrorUnknown=999 equal the CUDA codes the reference returns).
 *
 * Threading: every handle call takes that handle's mutex, so calls on one handle
 * serialize and distinct handles run concurrently, each on its own HIP stream.
 * Use after close is undefined (cuda_replace.cu:50-55).  A handle remembers the
 * device that was current at open and runs there.
 */
#ifndef CUDA_REPLACE_H
#define CUDA_REPLACE_H

#include <stddef.h>
#include <stdint.h>

#if defined(__GNUC__) || defined(__clang__)
#define CR_API __attribute__((visibility("default")))
#else
#define CR_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* One-shot: open, apply, result, close.  *out_host is a pinned buffer the
 * caller frees with cuda_free_host (NULL when *out_len == 0).
 * Replaces cuda_replace.cu:2579-2601. */
CR_API int cuda_replace_unified(const uint8_t *src, size_t src_len,
                         const uint8_t *pat, int pat_len,
                         const uint8_t *rep, int rep_len,
                         uint8_t **out_host, size_t *out_len);

/* Frees a buffer returned by cuda_replace_unified / cuda_replace_result;
 * NULL is ignored.  Replaces cuda_replace.cu:2559-2563. */
CR_API void cuda_free_host(void *p);

/* Opens a session holding a device-resident copy of src.
 * Replaces cuda_replace.cu:2256-2293. */
CR_API int cuda_replace_open(void **ph, const uint8_t *src, size_t len);

/* Replaces the session's buffer with src (grows capacity if needed).
 * Replaces cuda_replace.cu:2295-2318. */
CR_API int cuda_replace_reset(void *h, const uint8_t *src, size_t len);

/* Builds a per-chunk 256-bit byte-presence index (chunk_size 0 -> 1 MiB).  As in
 * the reference, no apply consults it.  Replaces cuda_replace.cu:2320-2354. */
CR_API int cuda_replace_build_index(void *h, size_t chunk_size);

/* One replace pass over the session buffer; *new_len gets the new length.
 * Replaces cuda_replace.cu:2356-2401. */
CR_API int cuda_replace_apply(void *h, const uint8_t *pat, int pat_len,
                       const uint8_t *rep, int rep_len, size_t *new_len);

/* Apply with a carry: prev_last_rel is the last kept start of the previous
 * chunk relative to this buffer (candidates i with i - prev_last_rel < pat_len
 * are not kept; any value <= -pat_len means no carry).  *out_last_rel gets the
 * last kept start of this pass or -1.  Replaces cuda_replace.cu:2403-2450. */
CR_API int cuda_replace_apply_seeded(void *h, const uint8_t *pat, int pat_len,
                              const uint8_t *rep, int rep_len,
                              long long prev_last_rel, long long *out_last_rel,
                              size_t *new_len);

/* Kept starts of the most recent apply strictly below limit (clamped to that
 * apply's source length), in that source's coordinates: their number and the
 * last one (or -1).  Returns -4 if the count exceeds INT_MAX (see
 * cr_query_prefix64).  Replaces cuda_replace.cu:2452-2463. */
CR_API int cuda_replace_query_prefix(void *h, size_t limit,
                              int *kept_before_limit, int *last_kept_before_limit);

/* count sequential seedless applies.  NULL pats or pat_lens skips every pair;
 * NULL reps or rep_lens means empty replacements.  When cr_batch_fusable holds
 * for the pairs, they run as one fused pass over the buffer (same result); a
 * later prefix query then replays the pairs one by one to answer for the last
 * pair.  Replaces cuda_replace.cu:2465-2527. */
CR_API int cuda_replace_apply_batch(void *h, const uint8_t **pats, const int *pat_lens,
                             const uint8_t **reps, const int *rep_lens,
                             int count, size_t *new_len);

/* Pinned host copy of the session buffer (free with cuda_free_host).
 * Replaces cuda_replace.cu:2528-2557. */
CR_API int cuda_replace_result(void *h, uint8_t **out_host, size_t *outlen_host);

/* Closes a session; NULL is ignored.  Replaces cuda_replace.cu:2565-2577. */
CR_API void cuda_replace_close(void *h);

/* cuda_replace_apply_batch with the fusion choice exposed: allow_fuse=0 always
 * applies pair by pair; *fused (optional) receives 1 if the pairs ran as one
 * fused pass.  No counterpart in the reference. */
CR_API int cr_apply_batch_ex(void *h, const uint8_t *const *pats, const int *pat_lens,
                             const uint8_t *const *reps, const int *rep_lens, int count,
                             size_t *new_len, int allow_fuse, int *fused);

/* 1 if applying the pairs in order equals one fused pass that replaces every
 * occurrence of every pattern: 2..256 pairs; patterns of 2..64 bytes sharing
 * their first byte, each border-free (no proper prefix equals a suffix),
 * distinct, none inside another, no proper suffix of one equal to a prefix of
 * another; replacements of 1..64 bytes with no byte that occurs in any pattern;
 * patterns and replacements together at most 16384 bytes.  0 otherwise, -2 on
 * allocation failure.  No counterpart in the reference. */
CR_API int cr_batch_fusable(const uint8_t *const *pats, const int *pat_lens,
                            const uint8_t *const *reps, const int *rep_lens, int count);

/* Autotuner.  Carry-table classes (2/4/8/16 states): the tiles per workgroup group
 * (1, 2 or 4 tiles of 8 KiB).  Border-free patterns: the form of the pass -- one
 * persistent launch (64 source bytes per thread, counts ahead of the emit) that
 * reads each group again for its emit, or one that holds it in LDS from its count
 * to its emit, the same kernels as three launches, or the 32-byte-per-thread count
 * and emit with 1, 2 or 4 tiles per group.  The first pass of a class in an input size bucket
 * (under 4 MiB, 4-32 MiB, 32-256 MiB, 256 MiB-2 GiB, larger) on a device times
 * each candidate after its normal pass and keeps the fastest; inputs under 4 MiB
 * keep the default unless pinned.  CR_GROUP_TILES=1|2|4 fixes the tables' choice,
 * CR_BF_VARIANT=0..5 the border-free form (0 one launch, 1 three launches, 2/3/4
 * 32-byte threads with 1/2/4 tiles, 5 one launch holding groups in LDS),
 * CR_AUTOTUNE=0 disables measuring.  No
 * counterpart in the reference.
 *
 * cr_autotune_set pins the choice for the class of pattern pat and the bucket of n
 * on the current device; 0 unpins.  group_tiles 1, 2 or 4 is a group size (for a
 * border-free pattern: the 32-byte form with that group size); for a border-free
 * pattern 8 pins the three-launch 64-byte form, 16 the single launch and 32 the
 * single launch that holds groups in LDS.  Returns
 * 0, or -1 for a pattern outside the tuned classes (empty, or the bitmap path) or a
 * bad value. */
CR_API int cr_autotune_set(const uint8_t *pat, int pat_len, size_t n, int group_tiles);
/* Forgets every measured or pinned choice. */
CR_API void cr_autotune_reset(void);
/* JSON array of the current choices and their measured milliseconds per candidate,
 * NUL-terminated into buf (truncated to cap); returns the full length. */
CR_API int cr_autotune_report(char *buf, size_t cap);

/* Library version string.  Replaces cuda_replace.cu:2251-2254. */
CR_API const char *cuda_replace_version(void);

/* ---- extensions (no counterpart in the reference) ------------------------ */

/* Current session length without a device-to-host copy of the buffer (the
 * reference wrapper's Session.length() copies the whole buffer). */
CR_API int cr_session_length(void *h, size_t *len);

/* 64-bit form of cuda_replace_query_prefix. */
CR_API int cr_query_prefix64(void *h, size_t limit, long long *kept_before_limit,
                      long long *last_kept_before_limit);

/* Replace between caller-owned device buffers on the session's stream and
 * scratch: d_src (n bytes, 16-byte aligned) -> d_dst (dst_cap bytes, 16-byte
 * aligned, not overlapping d_src).  seed as in cuda_replace_apply_seeded
 * (pass a value <= -pat_len for none).  *out_len gets the output length; if
 * dst_cap is too small nothing is written, *out_len still gets the needed
 * length and hipErrorInvalidValue is returned.  *kernel_ms (optional) gets the
 * device time of the pass measured with HIP events on the session stream.
 * The pass uses the session's scratch, so the record of the session's last
 * apply is dropped: a prefix query answers 0 / -1 until the next apply; the
 * session's own buffer is not touched. */
CR_API int cr_replace_device(void *h, const uint8_t *d_src, size_t n, uint8_t *d_dst,
                      size_t dst_cap, const uint8_t *pat, int pat_len,
                      const uint8_t *rep, int rep_len, long long seed,
                      size_t *out_len, long long *out_last, float *kernel_ms);

/* No counterpart in the reference.  For callers that cannot allocate (tight
 * serving loops, harnesses that forbid hipMalloc): the pass of cr_replace_device
 * with no session and no allocation at all.  Its scratch is one static device
 * array of about 20 MiB per GPU (held from library load until the process ends);
 * it runs on `stream` (a hipStream_t; NULL is the default stream) and returns
 * when the output is written.  Limits: n <= 4 GiB, pat_len and rep_len <= 256,
 * and a border-free or carry-table pattern (a self-overlapping pattern longer
 * than 16 bytes needs the bitmap path); anything beyond returns -5 and touches
 * nothing.  Every call on one GPU shares the array, so calls that overlap on one
 * GPU (several threads or streams at once) race.  Arguments and the too-small
 * dst_cap rule as cr_replace_device. */
CR_API int cr_replace_device_noalloc(const uint8_t *d_src, size_t n, uint8_t *d_dst,
                      size_t dst_cap, const uint8_t *pat, int pat_len,
                      const uint8_t *rep, int rep_len, long long seed,
                      void *stream, size_t *out_len);

/* Which suppression path a pattern takes: 1 border-free single pass, 2 carry
 * tables (self-overlapping, length <= 16), 3 bitmap + serial greedy, 0 empty. */
CR_API int cr_pattern_mode(const uint8_t *pat, int pat_len);

/* Streams a host buffer through the GPU: for each pair in order, the current data
 * passes in chunks of chunk_bytes source bytes (0 -> 256 MiB) with an (L-1)-byte
 * overlap, the seed carried from chunk to chunk and each chunk's output cut at the
 * end of a kept match straddling its end, so the result equals bytes.replace of the
 * whole input (replaces the protocol of cuda_replace_wrapper.py:305-366, which cuts
 * at the chunk end and loses bytes there).  The next chunk is staged in pinned
 * memory and uploaded while the current one computes, and each chunk's output is
 * downloaded while the next computes.  *out_host is a pinned buffer freed with
 * cuda_free_host (NULL when *out_len == 0).  stats (optional, 4 doubles): wall ms,
 * bytes streamed in, bytes streamed out (summed over pairs), chunks processed.
 * Uses the session's stream and device; leaves the session's buffer untouched. */
CR_API int cr_stream_replace(void *h, const uint8_t *src, size_t n, const uint8_t *const *pats,
                             const int *pat_lens, const uint8_t *const *reps, const int *rep_lens,
                             int count, size_t chunk_bytes, uint8_t **out_host, size_t *out_len,
                             double *stats);

/* ---- batched documents (no counterpart in the reference; its callers loop over
 * cuda_replace_unified per document and pair, cuda_replace.cu:2579-2601) ----- */

/* Creates a batch engine on the current device: its own stream, plus device
 * arenas and pinned staging that stay resident between runs. */
CR_API int cr_batch_create(void **pb);

/* Runs num_docs independent documents.  Document i (docs[i], doc_lens[i] bytes)
 * gets pair_counts[i] (pattern, replacement) pairs applied in order; the pairs are
 * listed document after document in pats / pat_lens / reps / rep_lens.  Each
 * result equals applying that document's pairs one after another with
 * bytes.replace, including empty patterns.  Results are packed into an
 * engine-owned pinned buffer *out_base, valid until the next run or destroy;
 * document i's result is out_base[out_offsets[i] .. out_offsets[i] + out_lens[i]).
 * stats (optional, 4 doubles): device ms of the rounds (host round trips included,
 * uploads of the documents and the final download excluded), wall ms of the call,
 * algorithmic bytes (input + output of every applied pair), number of rounds.
 * Calls on one engine serialize. */
CR_API int cr_batch_run(void *b, int num_docs, const uint8_t *const *docs, const size_t *doc_lens,
                        const int *pair_counts, const uint8_t *const *pats, const int *pat_lens,
                        const uint8_t *const *reps, const int *rep_lens, const uint8_t **out_base,
                        size_t *out_offsets, size_t *out_lens, double *stats);

/* How the engine runs rounds.  mode 0 (default; CR_BATCH_DEVICE=0 in the
 * environment makes 1 the default): device-driven rounds whenever no pair needs
 * the bitmap path -- every round is planned, counted, placed and emitted on the
 * GPU by persistent kernels over device-built work lists; a run shape that
 * repeats is captured as a HIP graph and replayed while the shape and the
 * engine's buffers stay the same (CR_BATCH_GRAPH=0 launches the kernels
 * directly); the tiles per group are measured on the first run of an input size
 * bucket (CR_GROUP_TILES fixes them).  A round that does not fit the engine's
 * buffers commits nothing; the host grows the buffers and resumes at that round.
 * mode 1: host-driven rounds only (one readback of kept counts per round; also
 * used for pairs that need the bitmap path).  Returns 0 or -1.  No counterpart in
 * the reference. */
CR_API int cr_batch_set_mode(void *b, int mode);

/* The last run of the engine: info[0] 1 host-driven / 2 device-driven, [1] tiles
 * per group, [2] the HIP graph was replayed without a new capture, [3] the group
 * size was measured in this run, [4] rounds, [5] ran as a HIP graph, [6] times a
 * round was resumed after the buffers grew.  n entries (at most 7) are written.
 * Returns 0 or -1. */
CR_API int cr_batch_info(void *b, double *info, int n);

/* Destroys a batch engine; NULL is ignored. */
CR_API void cr_batch_destroy(void *b);

/* ---- continuous-batching server (no counterpart in the reference) ---------- */

/* Starts a server on the current device with `workers` batch engines (0 -> 4),
 * an HBM budget for admitted requests (0 -> 200 GB) and a per-batch cap (0 -> 4 GiB).
 * Each request holds 3 x its input + 1 MiB of budget while its batch runs. */
CR_API int cr_server_create(void **ps, size_t hbm_budget_bytes, int workers, size_t max_batch_bytes);

/* Queues one request: doc and its ordered (pattern, replacement) pairs, with
 * bytes.replace semantics per pair.  The caller keeps doc and the pattern and
 * replacement bytes alive until cr_server_wait returns for *ticket. */
CR_API int cr_server_submit(void *srv, const uint8_t *doc, size_t len, int npairs,
                            const uint8_t *const *pats, const int *pat_lens,
                            const uint8_t *const *reps, const int *rep_lens, void **ticket);

/* Blocks until the request is done; *out / *out_len get its result (owned by the
 * ticket).  times (optional, 3 doubles): ms from submit to its batch's start, ms
 * from submit to completion, documents in its batch. */
CR_API int cr_server_wait(void *ticket, const uint8_t **out, size_t *out_len, double *times);

/* Frees a ticket and its result (waits for completion first). */
CR_API void cr_server_release(void *ticket);

/* Completes every queued request, then stops the workers and frees the server. */
CR_API void cr_server_destroy(void *srv);

#ifdef __cplusplus
}
#endif

#endif /* CUDA_REPLACE_H */
