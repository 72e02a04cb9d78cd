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

}  // namespace lira
