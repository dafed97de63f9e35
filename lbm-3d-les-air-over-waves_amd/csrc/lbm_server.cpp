// lbm_server.cpp -- native stepping runtime of the ensemble server.
//
// Many independent cases share one GPU.  Each case is a handle with its own
// HIP stream; worker threads round-robin over the cases they own and, for
// every case whose stream has drained, enqueue the next chunk of steps
// (lbm_step is asynchronous), so the kernels of all active cases overlap on
// the device and no case waits for another to finish.  A case is finished
// when its last chunk has completed on the device (lbm_query reports its
// stream idle); its ticket then moves to the finished list that
// lbm_server_poll drains.  Cases are spread over the workers at add time
// (fewest active first), so host launch work is spread over several threads.
// The server never allocates device memory; admission by footprint,
// initialisation and downloads are the caller's.
//
// With batching on (lbm_server_batch), the fused-path cases go to one batch
// thread instead: every chunk it advances all of its cases that share a
// storage type and collision model with one lbm_step_batch, so each step of
// all of them is a single launch whose grid holds every case's tiles (a
// merged launch grid); a case leaves the batch when its steps are done and
// is reported once its stream has drained.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <set>
#include <stdexcept>
#include <thread>
#include <vector>

#include <hip/hip_runtime.h>

#include "../../include/lbm.h"
#include "lbm_internal.h"

namespace {

struct Job {
    lbm_handle_t* h;
    int32_t nsteps;
    int32_t done;
    int64_t ticket;
    hipEvent_t last = nullptr;  // batch thread: recorded after the job's last chunk
};

struct Worker {
    std::mutex m;
    std::condition_variable cv;
    std::vector<Job> incoming;
    std::vector<Job> active;
    std::atomic<int> n_active{0};
    std::thread thread;
};

}  // namespace

struct lbm_server {
    int32_t chunk = 16;
    std::vector<Worker*> workers;
    Worker* batcher = nullptr;  // the batch thread (batching on)
    std::atomic<bool> stop{false};
    std::mutex fin_m;
    std::deque<std::pair<int64_t, int32_t>> finished;
    std::set<lbm_handle_t*> inflight;  // handles owned by the server (guarded by fin_m)

    void finish(lbm_handle_t* h, int64_t ticket, int32_t status) {
        std::lock_guard<std::mutex> g(fin_m);
        inflight.erase(h);
        finished.emplace_back(ticket, status);
    }

    // the batch thread: one lbm_step_batch per group of cases that can share
    // a launch, every round; cases whose steps are all enqueued wait on
    // their streams before they are reported
    void run_batch(Worker* w) {
        // the batches run on a stream of their own, one chunk after another;
        // a batched handle's own stream stays free of work, and the thread
        // reports a case once the event after its last chunk has completed
        hipStream_t bs = nullptr;
        lbm::BatchBuf buf;
        if (hipStreamCreateWithFlags(&bs, hipStreamNonBlocking) != hipSuccess) bs = nullptr;
        std::vector<Job> draining;
        struct Cleanup {
            hipStream_t& s;
            lbm::BatchBuf& b;
            ~Cleanup() {
                if (s) (void)hipStreamSynchronize(s);
                lbm::batch_buf_free(b);
                if (s) (void)hipStreamDestroy(s);
            }
        } cleanup{bs, buf};
        while (true) {
            {
                std::unique_lock<std::mutex> lk(w->m);
                if (w->active.empty() && w->incoming.empty() && draining.empty()) {
                    if (stop.load()) return;
                    w->cv.wait_for(lk, std::chrono::milliseconds(2));
                    continue;
                }
                for (auto& j : w->incoming) w->active.push_back(j);
                w->incoming.clear();
            }
            bool progressed = false;
            if (!stop.load() && !w->active.empty()) {
                // groups by batch key, in order of arrival
                std::vector<int> keys;
                for (auto& j : w->active) {
                    const int k = lbm::batch_key(j.h);
                    if (std::find(keys.begin(), keys.end(), k) == keys.end()) keys.push_back(k);
                }
                for (int key : keys) {
                    std::vector<lbm_handle_t*> hs;
                    std::vector<int32_t> ns;
                    std::vector<size_t> idx;
                    for (size_t a = 0; a < w->active.size(); a++) {
                        Job& j = w->active[a];
                        if (lbm::batch_key(j.h) != key) continue;
                        hs.push_back(j.h);
                        ns.push_back(std::min(chunk, j.nsteps - j.done));
                        idx.push_back(a);
                    }
                    const int rc = (key < 0 || !bs) ? LBM_E_ARG : lbm::step_batch(hs.data(), ns.data(), (int)hs.size(), bs, buf, false);
                    for (size_t m = 0; m < idx.size(); m++) {
                        Job& j = w->active[idx[m]];
                        if (rc == LBM_OK) j.done += ns[m];
                        else j.done = -1 - rc;  // failed: reported with the error code below
                    }
                    progressed = true;
                }
                // cases with every step enqueued (or failed) leave the batch
                for (size_t a = 0; a < w->active.size();) {
                    Job& j = w->active[a];
                    if (j.done < 0 || j.done >= j.nsteps) {
                        if (bs && hipEventCreateWithFlags(&j.last, hipEventDisableTiming) == hipSuccess &&
                            hipEventRecord(j.last, bs) != hipSuccess) {
                            (void)hipEventDestroy(j.last);
                            j.last = nullptr;
                        }
                        if (!j.last && bs) (void)hipStreamSynchronize(bs);
                        draining.push_back(j);
                        w->active.erase(w->active.begin() + (long)a);
                    } else {
                        a++;
                    }
                }
            }
            for (size_t d = 0; d < draining.size();) {
                Job& j = draining[d];
                int32_t idle = 0;
                // the device must be done with the case before it is
                // reported, failed or not
                int rc = LBM_OK;
                if (j.last) {
                    const hipError_t e = hipEventQuery(j.last);
                    if (e == hipErrorNotReady) {
                        d++;
                        continue;
                    }
                    (void)hipEventDestroy(j.last);
                    j.last = nullptr;
                    if (e != hipSuccess) rc = LBM_E_HIP;
                }
                if (rc == LBM_OK) rc = lbm_query(j.h, &idle);
                if (rc == LBM_OK && !idle) {
                    d++;
                    continue;
                }
                if (rc == LBM_OK && j.done < 0) rc = -(j.done + 1);
                finish(j.h, j.ticket, rc);
                draining.erase(draining.begin() + (long)d);
                w->n_active.fetch_sub(1);
                progressed = true;
            }
            if (stop.load() && w->active.empty() && draining.empty()) {
                std::lock_guard<std::mutex> g(w->m);
                if (w->incoming.empty()) return;
            }
            if (stop.load() && !w->active.empty()) {  // stopping: abandon the cases not yet done
                if (bs) (void)hipStreamSynchronize(bs);
                for (auto& j : w->active) {
                    j.done = -1 - LBM_E_STATE;
                    draining.push_back(j);
                }
                w->active.clear();
            }
            if (!progressed) std::this_thread::sleep_for(std::chrono::microseconds(20));
        }
    }

    void run(Worker* w) {
        while (true) {
            {
                std::unique_lock<std::mutex> lk(w->m);
                if (w->active.empty() && w->incoming.empty()) {
                    if (stop.load()) return;
                    w->cv.wait_for(lk, std::chrono::milliseconds(2));
                    continue;
                }
                for (auto& j : w->incoming) w->active.push_back(j);
                w->incoming.clear();
            }
            bool progressed = false;
            for (size_t a = 0; a < w->active.size();) {
                Job& j = w->active[a];
                int32_t idle = 0;
                int rc = lbm_query(j.h, &idle);
                if (rc == LBM_OK && !idle) {
                    a++;
                    continue;
                }
                if (rc == LBM_OK && j.done < j.nsteps && !stop.load()) {
                    const int32_t n = std::min(chunk, j.nsteps - j.done);
                    rc = lbm_step(j.h, n);
                    if (rc == LBM_OK) {
                        j.done += n;
                        progressed = true;
                        a++;
                        continue;
                    }
                }
                if (rc == LBM_OK && j.done < j.nsteps) {  // stopping: abandon
                    a++;
                    continue;
                }
                finish(j.h, j.ticket, rc);
                w->active.erase(w->active.begin() + (long)a);
                w->n_active.fetch_sub(1);
                progressed = true;
            }
            if (stop.load()) {
                bool all_idle = true;
                for (auto& j : w->active) {
                    int32_t idle = 1;
                    if (lbm_query(j.h, &idle) == LBM_OK && !idle) all_idle = false;
                }
                if (all_idle) {
                    std::lock_guard<std::mutex> g(w->m);
                    if (w->incoming.empty()) return;
                }
            }
            if (!progressed) std::this_thread::sleep_for(std::chrono::microseconds(20));
        }
    }
};

extern "C" {

int lbm_server_create(int32_t chunk_steps, int32_t n_workers, lbm_server_t** out) {
    if (!out || chunk_steps < 1 || n_workers < 1 || n_workers > 64) return LBM_E_ARG;
    lbm_server* s = nullptr;
    try {
        s = new lbm_server();
        s->chunk = chunk_steps;
        for (int i = 0; i < n_workers; i++) s->workers.push_back(new Worker());
        for (Worker* w : s->workers) w->thread = std::thread([s, w]() { s->run(w); });
    } catch (const std::exception&) {
        if (s) {
            s->stop = true;
            for (Worker* w : s->workers) {
                if (w->thread.joinable()) w->thread.join();
                delete w;
            }
            delete s;
        }
        return LBM_E_NOMEM;
    }
    *out = s;
    return LBM_OK;
}

int lbm_server_add(lbm_server_t* s, lbm_handle_t* h, int32_t nsteps, int64_t ticket) {
    if (!s || !h || nsteps < 0) return LBM_E_ARG;
    if (s->stop.load()) return LBM_E_STATE;
    {
        // a handle is stepped by one job at a time: two jobs on one handle
        // would race on its step counter, ping-pong state and stream
        std::lock_guard<std::mutex> g(s->fin_m);
        if (!s->inflight.insert(h).second) return LBM_E_STATE;
    }
    // the workers step their cases on concurrent streams with the fused
    // step's three-launch schedule, in which no block waits for another (the
    // library would otherwise chain the waiting launches of the cases, one
    // at a time: measured slower here); the merged grid of the batching
    // thread is the one waiting launch.  Nothing a worker runs waits for the
    // device: no choice of schedule by timing, no flood-fill host reads.
    (void)lbm_set_option(h, "resident", 0);
    (void)lbm_set_option(h, "wavefront", 0);
    (void)lbm_set_option(h, "flood_sync", 0);
    Worker* best = s->workers[0];
    for (Worker* w : s->workers)
        if (w->n_active.load() < best->n_active.load()) best = w;
    if (s->batcher && lbm::batch_key(h) >= 0) best = s->batcher;
    {
        std::lock_guard<std::mutex> g(best->m);
        best->incoming.push_back(Job{h, nsteps, 0, ticket});
        best->n_active.fetch_add(1);
    }
    best->cv.notify_one();
    return LBM_OK;
}

int lbm_server_poll(lbm_server_t* s, int64_t* tickets, int32_t* status, int32_t max, int32_t* n_done) {
    if (!s || !n_done || max < 0 || (max > 0 && (!tickets || !status))) return LBM_E_ARG;
    std::lock_guard<std::mutex> g(s->fin_m);
    int32_t n = 0;
    while (n < max && !s->finished.empty()) {
        tickets[n] = s->finished.front().first;
        status[n] = s->finished.front().second;
        s->finished.pop_front();
        n++;
    }
    *n_done = n;
    return LBM_OK;
}

int lbm_server_active(lbm_server_t* s, int32_t* n_active) {
    if (!s || !n_active) return LBM_E_ARG;
    int32_t n = 0;
    for (Worker* w : s->workers) n += w->n_active.load();
    if (s->batcher) n += s->batcher->n_active.load();
    *n_active = n;
    return LBM_OK;
}

int lbm_server_batch(lbm_server_t* s, int32_t on) {
    if (!s) return LBM_E_ARG;
    if (!on) return s->batcher ? LBM_E_STATE : LBM_OK;  // a running batch thread stays
    if (s->batcher) return LBM_OK;
    {
        std::lock_guard<std::mutex> g(s->fin_m);
        if (!s->inflight.empty()) return LBM_E_STATE;  // before the first case only
    }
    try {
        Worker* w = new Worker();
        w->thread = std::thread([s, w]() { s->run_batch(w); });
        s->batcher = w;
    } catch (const std::exception&) {
        return LBM_E_NOMEM;
    }
    return LBM_OK;
}

int lbm_server_destroy(lbm_server_t* s) {
    if (!s) return LBM_E_ARG;
    s->stop = true;
    for (Worker* w : s->workers) {
        w->cv.notify_all();
        if (w->thread.joinable()) w->thread.join();
        delete w;
    }
    if (s->batcher) {
        s->batcher->cv.notify_all();
        if (s->batcher->thread.joinable()) s->batcher->thread.join();
        delete s->batcher;
    }
    delete s;
    return LBM_OK;
}

}  // extern "C"
