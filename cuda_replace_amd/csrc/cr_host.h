// cr_host.h -- host-side helpers shared by the runtime files.
#pragma once
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#include <stdint.h>

#include <algorithm>
#include <functional>
#include <thread>
#include <vector>

namespace cr {

// memcpy split over up to 16 threads for large copies (CR_MEMCPY_THREADS overrides):
// pageable host memory into pinned staging runs at a few GB/s per thread.  The
// streaming pipeline's input stage is this copy: 8 threads held it to 16-19 GB/s on
// the MI355X host, 16 threads to 36 GB/s, where the PCIe copies take over.
inline int host_copy_threads()
{
    const unsigned hw = std::thread::hardware_concurrency();
    const char *env = getenv("CR_MEMCPY_THREADS");
    const unsigned cap = env && atoi(env) > 0 ? (unsigned)atoi(env) : 16u;
    return (int)std::min<unsigned>(hw ? hw : 1, cap);
}

inline void parallel_memcpy(void *dst, const void *src, size_t n)
{
    const int nt = n < (16u << 20) ? 1 : host_copy_threads();
    if (nt <= 1) {
        if (n)
            memcpy(dst, src, n);
        return;
    }
    std::vector<std::thread> ts;
    for (int t = 0; t < nt; ++t) {
        const size_t a = n * t / nt, b = n * (t + 1) / nt;
        ts.emplace_back([=] { memcpy((char *)dst + a, (const char *)src + a, b - a); });
    }
    for (auto &t : ts)
        t.join();
}

// A request of a live run: one document and its ordered pair list (caller-owned
// buffers, valid until its finish callback has returned).
struct LiveReq {
    const uint8_t *doc;
    size_t len;
    int npairs;
    const uint8_t *const *pats;
    const int *pat_lens;
    const uint8_t *const *reps;
    const int *rep_lens;
    void *tag;
};

// Continuous batching on one batch engine (cr_batch_create): a live set of documents,
// each at its own pair, advances one host-driven round at a time.  Before every
// round admit(out, idle) appends the requests that join (idle: nothing is live; it
// may block then, and returning false while idle ends the run); they are uploaded
// behind the live documents and take their first pair in that round.  As a document
// applies its last pair, finish(tag, rc, out, out_len) gets its result (out valid
// during the call).  Returns 0, or the error of a round (every live document is
// finished with it first).
int batch_live(void *eng, const std::function<bool(std::vector<LiveReq> &, bool)> &admit,
               const std::function<void(void *, int, const uint8_t *, size_t)> &finish);

} // namespace cr

