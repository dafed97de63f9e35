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

#include "../../include/lbm.h"

namespace {

struct Job {
    lbm_handle_t* h;
    int32_t nsteps;
    int32_t done;
    int64_t ticket;
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
    std::atomic<bool> stop{false};
    std::mutex fin_m;
    std::deque<std::pair<int64_t, int32_t>> finished;
    std::set<lbm_handle_t*> inflight;  // handles owned by the server (guarded by fin_m)

    void finish(lbm_handle_t* h, int64_t ticket, int32_t status) {
        std::lock_guard<std::mutex> g(fin_m);
        inflight.erase(h);
        finished.emplace_back(ticket, status);
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
    Worker* best = s->workers[0];
    for (Worker* w : s->workers)
        if (w->n_active.load() < best->n_active.load()) best = w;
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
    *n_active = n;
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
    delete s;
    return LBM_OK;
}

}  // extern "C"
