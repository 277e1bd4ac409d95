// cr_server.hip -- continuous-batching request server on top of the batch engine.
//
// Requests (one document and its ordered pair list) enter a FIFO.  Each worker
// thread owns a batch engine (its own HIP stream, resident arenas and pinned
// staging); four workers by default.  A worker runs a live set of documents, each
// at its own pair, one round at a time: before every round it admits queued
// requests in arrival order while the server's HBM budget and the worker's live-set
// cap allow, uploads them behind the live documents and starts them on their first
// pair in that same round; a document that has applied its last pair returns at
// once with its own copy of the result.  So an arrival waits at most for the round
// in flight, not for a whole batch, and the workers' rounds overlap on the device
// (one uploading or downloading while another computes).  A request larger than the
// whole budget runs alone once nothing else holds budget.  CR_SERVER_LIVE=0 runs
// the earlier form instead: a free worker takes the queued requests as one batch and
// returns them together.
//
// No counterpart in the reference (a library without a serving layer).
#include "../../include/cuda_replace.h"
#include "cr_host.h"

#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <new>
#include <stdlib.h>
#include <string.h>
#include <thread>
#include <vector>

namespace {

using Clock = std::chrono::steady_clock;

struct Ticket {
    // request (caller-owned buffers, valid until cr_server_wait returns)
    const uint8_t *doc = nullptr;
    size_t len = 0;
    std::vector<const uint8_t *> pats, reps;
    std::vector<int> pat_lens, rep_lens;
    size_t reserve = 0;   // HBM budget held while in a batch
    // result
    std::mutex mtx;
    std::condition_variable cv;
    bool done = false;
    int rc = 0;
    uint8_t *out = nullptr;
    size_t out_len = 0;
    Clock::time_point t_submit, t_start, t_done;
    int batch_docs = 0;
};

struct Server {
    int device = 0;
    bool live = true;
    size_t budget = 0;
    size_t max_batch = 0;
    std::mutex mtx;
    std::condition_variable cv;
    std::deque<Ticket *> queue;
    size_t in_use = 0;
    bool stop = false;
    std::vector<void *> engines;
    std::vector<std::thread> workers;
};

void complete(Ticket *t, int rc)
{
    std::lock_guard<std::mutex> lk(t->mtx);
    t->rc = rc;
    t->t_done = Clock::now();
    t->done = true;
    t->cv.notify_all();
}

void worker_loop(Server *S, int w)
{
    (void)hipSetDevice(S->device);
    void *eng = S->engines[(size_t)w];
    std::vector<Ticket *> batch;
    std::vector<const uint8_t *> docs, pats, reps;
    std::vector<size_t> lens, offs, olens;
    std::vector<int> counts, pat_lens, rep_lens;
    for (;;) {
        size_t held = 0;
        {
            std::unique_lock<std::mutex> lk(S->mtx);
            for (;;) {
                if (S->stop && S->queue.empty())
                    return;
                // admit in arrival order while budget and batch cap allow; a request
                // larger than the budget goes alone when nothing else holds budget
                while (!S->queue.empty()) {
                    Ticket *t = S->queue.front();
                    const bool fits = S->in_use + held + t->reserve <= S->budget &&
                                      (batch.empty() || held + t->reserve <= S->max_batch);
                    const bool alone = batch.empty() && S->in_use == 0;
                    if (!fits && !alone)
                        break;
                    S->queue.pop_front();
                    batch.push_back(t);
                    held += t->reserve;
                    if (!fits)
                        break;   // oversized request runs by itself
                }
                if (!batch.empty())
                    break;
                S->cv.wait(lk);
            }
            S->in_use += held;
        }
        const auto t_start = Clock::now();
        docs.clear();
        lens.clear();
        counts.clear();
        pats.clear();
        reps.clear();
        pat_lens.clear();
        rep_lens.clear();
        for (Ticket *t : batch) {
            t->t_start = t_start;
            t->batch_docs = (int)batch.size();
            docs.push_back(t->doc);
            lens.push_back(t->len);
            counts.push_back((int)t->pats.size());
            pats.insert(pats.end(), t->pats.begin(), t->pats.end());
            reps.insert(reps.end(), t->reps.begin(), t->reps.end());
            pat_lens.insert(pat_lens.end(), t->pat_lens.begin(), t->pat_lens.end());
            rep_lens.insert(rep_lens.end(), t->rep_lens.begin(), t->rep_lens.end());
        }
        offs.assign(batch.size(), 0);
        olens.assign(batch.size(), 0);
        const uint8_t *base = nullptr;
        int rc = cr_batch_run(eng, (int)batch.size(), docs.data(), lens.data(), counts.data(), pats.data(),
                              pat_lens.data(), reps.data(), rep_lens.data(), &base, offs.data(), olens.data(),
                              nullptr);
        for (size_t i = 0; i < batch.size(); ++i) {
            Ticket *t = batch[i];
            int trc = rc;
            if (rc == 0 && olens[i] > 0) {
                t->out = (uint8_t *)malloc(olens[i]);
                if (t->out) {
                    cr::parallel_memcpy(t->out, base + offs[i], olens[i]);
                    t->out_len = olens[i];
                } else {
                    trc = -2;
                }
            }
            complete(t, trc);
        }
        batch.clear();
        {
            std::lock_guard<std::mutex> lk(S->mtx);
            S->in_use -= held;
        }
        S->cv.notify_all();
    }
}

// Admission test for a queued request: it fits the budget and the worker's live-set
// cap, or it may run alone (larger than either, with nothing else holding budget).
bool admissible(const Server *S, const Ticket *t, size_t held, bool worker_idle, bool *alone)
{
    const bool fits = S->in_use + t->reserve <= S->budget && (worker_idle || held + t->reserve <= S->max_batch);
    *alone = !fits && worker_idle && S->in_use == 0;
    return fits || *alone;
}

void live_worker_loop(Server *S, int w)
{
    (void)hipSetDevice(S->device);
    void *eng = S->engines[(size_t)w];
    size_t held = 0;    // budget held by this worker's live requests
    size_t nlive = 0;
    bool solo = false;  // an oversized request runs alone
    auto admit = [&](std::vector<cr::LiveReq> &out, bool idle) -> bool {
        std::unique_lock<std::mutex> lk(S->mtx);
        for (;;) {
            while (!S->queue.empty() && !solo) {
                Ticket *t = S->queue.front();
                bool alone = false;
                if (!admissible(S, t, held, nlive == 0 && out.empty(), &alone))
                    break;
                S->queue.pop_front();
                S->in_use += t->reserve;
                held += t->reserve;
                ++nlive;
                solo = alone;
                t->t_start = Clock::now();
                t->batch_docs = (int)nlive;
                out.push_back(cr::LiveReq{t->doc, t->len, (int)t->pats.size(), t->pats.data(), t->pat_lens.data(),
                                          t->reps.data(), t->rep_lens.data(), t});
            }
            if (!out.empty() || !idle)
                return true;
            if (S->stop && S->queue.empty())
                return false;
            S->cv.wait(lk);
        }
    };
    auto finish = [&](void *tag, int rc, const uint8_t *res, size_t len) {
        Ticket *t = static_cast<Ticket *>(tag);
        int trc = rc;
        if (rc == 0 && len > 0) {
            t->out = (uint8_t *)malloc(len);
            if (t->out) {
                cr::parallel_memcpy(t->out, res, len);
                t->out_len = len;
            } else {
                trc = -2;
            }
        }
        {
            std::lock_guard<std::mutex> lk(S->mtx);
            S->in_use -= t->reserve;
            held -= t->reserve;
            --nlive;
            if (nlive == 0)
                solo = false;
        }
        S->cv.notify_all();
        complete(t, trc);
    };
    for (;;) {
        if (cr::batch_live(eng, admit, finish) == 0)
            return;   // stopped with nothing queued or live
    }
}

void shutdown(Server *S)
{
    {
        std::lock_guard<std::mutex> lk(S->mtx);
        S->stop = true;
    }
    S->cv.notify_all();
    for (auto &th : S->workers)
        if (th.joinable())
            th.join();
    for (void *e : S->engines)
        cr_batch_destroy(e);
}

} // namespace

extern "C" {

CR_API int cr_server_create(void **ps, size_t hbm_budget_bytes, int workers, size_t max_batch_bytes)
{
    if (!ps)
        return -1;
    *ps = nullptr;
    Server *S = new (std::nothrow) Server();
    if (!S)
        return -2;
    hipError_t e = hipGetDevice(&S->device);
    if (e != hipSuccess) {
        delete S;
        return (int)e;
    }
    S->budget = hbm_budget_bytes ? hbm_budget_bytes : (size_t)200 << 30;
    S->max_batch = max_batch_bytes ? max_batch_bytes : (size_t)4 << 30;
    const int nw = workers > 0 ? workers : 4;
    for (int w = 0; w < nw; ++w) {
        void *eng = nullptr;
        int rc = cr_batch_create(&eng);
        if (rc != 0) {
            shutdown(S);
            delete S;
            return rc;
        }
        S->engines.push_back(eng);
    }
    const char *lv = getenv("CR_SERVER_LIVE");
    S->live = !(lv && atoi(lv) == 0);
    for (int w = 0; w < nw; ++w) {
        if (S->live)
            S->workers.emplace_back(live_worker_loop, S, w);
        else
            S->workers.emplace_back(worker_loop, S, w);
    }
    *ps = S;
    return 0;
}

CR_API int cr_server_submit(void *srv, const uint8_t *doc, size_t len, int npairs, const uint8_t *const *pats,
                            const int *pat_lens, const uint8_t *const *reps, const int *rep_lens, void **ticket)
{
    if (!srv || !ticket || npairs < 0 || (len > 0 && !doc) || (npairs > 0 && (!pats || !pat_lens)))
        return -1;
    *ticket = nullptr;
    Server *S = static_cast<Server *>(srv);
    Ticket *t = new (std::nothrow) Ticket();
    if (!t)
        return -2;
    t->doc = doc;
    t->len = len;
    for (int i = 0; i < npairs; ++i) {
        const int rl = (reps && rep_lens) ? rep_lens[i] : 0;
        if (pat_lens[i] < 0 || rl < 0 || (pat_lens[i] > 0 && !pats[i]) || (rl > 0 && !reps[i])) {
            delete t;
            return -1;
        }
        t->pats.push_back(pats[i]);
        t->pat_lens.push_back(pat_lens[i]);
        t->reps.push_back(rl > 0 ? reps[i] : nullptr);
        t->rep_lens.push_back(rl);
    }
    t->reserve = 3 * len + ((size_t)1 << 20);   // source + two output arenas, roughly
    t->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(S->mtx);
        if (S->stop) {
            delete t;
            return -3;
        }
        S->queue.push_back(t);
    }
    S->cv.notify_one();
    *ticket = t;
    return 0;
}

// Blocks until the request is done.  *out (malloc'd, owned by the ticket) and
// *out_len receive the result; times (optional, 3 doubles): ms from submit to its
// admission, ms from submit to completion, and the documents live on its worker
// when it was admitted, itself included (the batch size in the batch-at-a-time form).
CR_API int cr_server_wait(void *ticket, const uint8_t **out, size_t *out_len, double *times)
{
    if (!ticket)
        return -1;
    Ticket *t = static_cast<Ticket *>(ticket);
    std::unique_lock<std::mutex> lk(t->mtx);
    t->cv.wait(lk, [&] { return t->done; });
    if (out)
        *out = t->out;
    if (out_len)
        *out_len = t->out_len;
    if (times) {
        times[0] = std::chrono::duration<double, std::milli>(t->t_start - t->t_submit).count();
        times[1] = std::chrono::duration<double, std::milli>(t->t_done - t->t_submit).count();
        times[2] = t->batch_docs;
    }
    return t->rc;
}

// Frees a ticket (after cr_server_wait returned) and its result.
CR_API void cr_server_release(void *ticket)
{
    if (!ticket)
        return;
    Ticket *t = static_cast<Ticket *>(ticket);
    {
        std::unique_lock<std::mutex> lk(t->mtx);
        t->cv.wait(lk, [&] { return t->done; });
    }
    free(t->out);
    delete t;
}

// Finishes every queued request, then stops the workers.
CR_API void cr_server_destroy(void *srv)
{
    if (!srv)
        return;
    Server *S = static_cast<Server *>(srv);
    shutdown(S);
    delete S;
}

} // extern "C"
