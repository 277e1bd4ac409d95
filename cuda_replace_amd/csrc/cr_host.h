// cr_host.h -- host-side helpers shared by the runtime files.
#pragma once
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
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

} // namespace cr
