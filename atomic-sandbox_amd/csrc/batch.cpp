// Batched SCF: many independent all-electron molecules on one GPU. The reference evaluates such sets one molecule at a
// time (crates/providers/src/sampling/mod.rs:1635, one GpuKs solve per sample); here host worker threads each own one
// solver (and so one HIP stream) and pull molecules from a shared queue, so the molecules' kernels overlap on the
// device and one molecule's host Rayleigh-Ritz steps overlap with the others' kernels. Each molecule's SCF is the
// single-solve path unchanged (lira_solve, then lira_forces), cold, with its result bitwise that of a lone solve.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <numeric>
#include <thread>
#include <vector>

#include "lira.h"

namespace {

typedef std::chrono::steady_clock Clock;

double since(Clock::time_point t0) { return std::chrono::duration<double>(Clock::now() - t0).count(); }

bool job_ok(const lira_batch_job &j) {
    if (j.n < 3 || j.n > LIRA_MAX_GRID || !(j.h > 0.0)) return false;
    if (j.system.n_nuclei < 1 || !j.system.nuclei || j.system.n_up < 0 || j.system.n_dn < 0) return false;
    return true;
}

void note(lira_batch_result &r, const char *msg) {
    std::strncpy(r.message, msg ? msg : "", sizeof(r.message) - 1);
    r.message[sizeof(r.message) - 1] = '\0';
}

}  // namespace

extern "C" lira_status lira_solve_batch(const lira_batch_job *jobs, int n_jobs, int precision,
                                        const lira_settings *settings, int workers, lira_batch_result *results,
                                        double *forces) {
    if (n_jobs < 0 || workers < 1 || workers > LIRA_BATCH_MAX_WORKERS) return LIRA_INVALID;
    if (precision != LIRA_FP64 && precision != LIRA_MIXED) return LIRA_INVALID;
    if (n_jobs == 0) return LIRA_OK;
    if (!jobs || !results) return LIRA_INVALID;
    for (int i = 0; i < n_jobs; i++)
        if (!job_ok(jobs[i])) return LIRA_INVALID;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return LIRA_DEVICE;

    std::vector<size_t> foff(n_jobs + 1, 0);
    for (int i = 0; i < n_jobs; i++) foff[i + 1] = foff[i] + 3 * (size_t)jobs[i].system.n_nuclei;
    // largest grids first (longest-processing-time order), molecules of one grid together so that a worker's solver
    // is rebuilt only when the grid changes
    std::vector<int> order(n_jobs);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        if (jobs[a].n != jobs[b].n) return jobs[a].n > jobs[b].n;
        return jobs[a].h < jobs[b].h;
    });
    for (int i = 0; i < n_jobs; i++) {
        std::memset(&results[i], 0, sizeof(lira_batch_result));
        results[i].status = LIRA_ABORTED;  // until a worker takes it
        results[i].worker = -1;
    }

    std::atomic<int> next{0};
    const Clock::time_point t0 = Clock::now();
    auto work = [&](int w) {
        if (hipSetDevice(dev) != hipSuccess) return;  // its molecules go to the other workers
        lira_solver *s = nullptr;
        int sn = -1;
        double sh = 0.0;
        for (;;) {
            const int k = next.fetch_add(1);
            if (k >= n_jobs) break;
            const int j = order[k];
            const lira_batch_job &job = jobs[j];
            lira_batch_result &r = results[j];
            r.worker = w;
            r.start = since(t0);
            lira_status st = LIRA_OK;
            if (!s || sn != job.n || sh != job.h) {
                if (s) lira_solver_destroy(s);
                s = nullptr;
                sn = -1;
                st = lira_solver_create(job.n, job.h, precision, &s);
                if (st == LIRA_OK) {
                    sn = job.n;
                    sh = job.h;
                    // eager launches: a stream capture on this thread would meet the other workers' allocations
                    st = lira_set_graphs(s, 0);
                }
                if (st != LIRA_OK) note(r, lira_last_error(s));
            }
            if (st == LIRA_OK) {
                lira_settings set;
                if (settings)
                    set = settings[j];
                else
                    lira_default_settings(job.h, &set);
                st = lira_solve(s, &job.system, &set, 0, nullptr, nullptr, &r.stats);
                if (st == LIRA_OK && forces)
                    st = lira_forces(s, job.system.nuclei, job.system.n_nuclei, set.sigma, forces + foff[j]);
                if (st != LIRA_OK) note(r, lira_last_error(s));
            }
            r.status = st;
            r.end = since(t0);
            if (st == LIRA_DEVICE && s) {  // a device fault leaves the solver unusable: rebuild for the next one
                lira_solver_destroy(s);
                s = nullptr;
                sn = -1;
            }
        }
        if (s) lira_solver_destroy(s);
    };
    const int nw = std::min(workers, n_jobs);
    std::vector<std::thread> pool;
    pool.reserve(nw);
    for (int w = 0; w < nw; w++) {
        try {
            pool.emplace_back(work, w);
        } catch (...) {  // no more threads: the ones started share the queue
            break;
        }
    }
    if (pool.empty()) work(0);  // not even one: the calling thread works through the queue
    for (auto &t : pool) t.join();
    // a molecule no worker took (every worker failed to select the device) is a device failure of the batch
    for (int i = 0; i < n_jobs; i++)
        if (results[i].worker < 0) return LIRA_DEVICE;
    return LIRA_OK;
}
