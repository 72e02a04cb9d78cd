// Host policies the reference's Kohn-Sham provider applies around GpuSolver (crates/providers/src/ks.rs,
// crates/providers/src/pseudo.rs), kept free of HIP so they can be exercised without a device.
#pragma once

#include <algorithm>

namespace lira {

// The f32 storage-rounding floor of the GPU SCF's delta_rho per valence electron (crates/providers/src/pseudo.rs:423-427)
constexpr double GPU_DRHO_FLOOR_PER_ELECTRON = 5e-8;

// delta_rho floor of an N-electron GPU solve (pseudo.rs:430-432)
inline double gpu_tol_floor(long n_valence) { return GPU_DRHO_FLOOR_PER_ELECTRON * (double)std::max(n_valence, 0L); }

// max(tol, 2 floor(N)): the SCF tolerance the GPU can reach (pseudo.rs:437-440); applied under an ionic setup only
inline double effective_gpu_tol(double tol, long n_valence) { return std::max(tol, 2.0 * gpu_tol_floor(n_valence)); }

// R15 (crates/providers/src/ks.rs:585-601): solve(w) runs one SCF, warm iff w, and returns its status. The first
// attempt runs with `warm`; a warm attempt that ended not converged or stalled is followed by exactly one cold attempt,
// whose outcome is returned. Every other outcome (success, a cold failure, any other error) is the first attempt's.
// *retried says whether the cold attempt ran, *first the first attempt's status.
template <typename Status, typename Solve>
Status retry_cold_on_stall(bool warm, Status not_converged, Status stalled, Solve &&solve, bool *retried,
                           Status *first) {
    const Status s = solve(warm);
    *first = s;
    *retried = false;
    if (warm && (s == not_converged || s == stalled)) {
        *retried = true;
        return solve(false);
    }
    return s;
}

// The "GPU busy" policy (crates/providers/src/ks.rs:317-350): attempt() runs; a device failure (a readback the shared
// device did not deliver) is retried after a pause, at most DEVICE_RETRIES times, pause(ms) returning false when the
// evaluation was cancelled meanwhile (then `cancelled` is returned). Every other outcome is the attempt's own.
// *failed receives the number of failed attempts before the returned one.
constexpr int DEVICE_RETRIES = 2;
constexpr int DEVICE_BACKOFF_MS[DEVICE_RETRIES] = {100, 400};
template <typename Status, typename Attempt, typename Pause>
Status retry_on_device_fault(Status device, Status cancelled, Attempt &&attempt, Pause &&pause, int *failed) {
    *failed = 0;
    for (;;) {
        const Status s = attempt();
        if (s != device || *failed == DEVICE_RETRIES) return s;
        if (!pause(DEVICE_BACKOFF_MS[*failed])) return cancelled;
        ++*failed;
    }
}

// witness_cap_for (crates/providers/src/ks.rs:400-428): the largest w in (counted, max_cap] whose footprint(w) fits
// budget, never below `counted` (the start the solve allocates). footprint must grow with w.
template <typename Footprint>
int witness_cap_for_budget(Footprint &&footprint, unsigned long long budget, int counted, int max_cap) {
    int cap = counted;
    for (int w = counted + 1; w <= max_cap && footprint(w) <= budget; w++) cap = w;
    return cap;
}

}  // namespace lira
