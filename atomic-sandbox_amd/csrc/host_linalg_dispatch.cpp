// The lira:: entry points of the host linear algebra. The Rayleigh-Ritz step and the symmetric eigensolver run in the
// AVX2 + FMA build of host_linalg.cpp when the host CPU has both (LIRA_HOST_ISA=base forces the baseline build), the
// small solves (LU, Cholesky) in the baseline build. The two builds differ only in floating-point contraction and
// vector width: the same algorithm, results equal to rounding.
#include <cstdlib>
#include <cstring>

#include "lira_internal.hpp"

namespace lira {
namespace linalg_base {
void sym_eigen(int m, std::vector<double> &a, std::vector<double> &evals, std::vector<double> &evecs);
bool lu_solve(int m, std::vector<double> a, std::vector<double> &b);
bool cholesky(int m, std::vector<double> &a);
void chol_solve(int m, const std::vector<double> &l, std::vector<double> &b, int nrhs);
double rayleigh_ritz(int m, const std::vector<double> &A, const std::vector<double> &B, int nk, double drop,
                     std::vector<double> &theta, std::vector<double> &C);
}  // namespace linalg_base
namespace linalg_avx2 {
void sym_eigen(int m, std::vector<double> &a, std::vector<double> &evals, std::vector<double> &evecs);
double rayleigh_ritz(int m, const std::vector<double> &A, const std::vector<double> &B, int nk, double drop,
                     std::vector<double> &theta, std::vector<double> &C);
}  // namespace linalg_avx2

static bool use_avx2() {
    static const bool on = [] {
        const char *e = std::getenv("LIRA_HOST_ISA");
        if (e && std::strcmp(e, "base") == 0) return false;
        __builtin_cpu_init();
        return __builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma");
    }();
    return on;
}

void sym_eigen(int m, std::vector<double> &a, std::vector<double> &evals, std::vector<double> &evecs) {
    if (use_avx2()) linalg_avx2::sym_eigen(m, a, evals, evecs);
    else linalg_base::sym_eigen(m, a, evals, evecs);
}

double rayleigh_ritz(int m, const std::vector<double> &A, const std::vector<double> &B, int nk, double drop,
                     std::vector<double> &theta, std::vector<double> &C) {
    return use_avx2() ? linalg_avx2::rayleigh_ritz(m, A, B, nk, drop, theta, C)
                      : linalg_base::rayleigh_ritz(m, A, B, nk, drop, theta, C);
}

bool lu_solve(int m, std::vector<double> a, std::vector<double> &b) { return linalg_base::lu_solve(m, std::move(a), b); }
bool cholesky(int m, std::vector<double> &a) { return linalg_base::cholesky(m, a); }
void chol_solve(int m, const std::vector<double> &l, std::vector<double> &b, int nrhs) {
    linalg_base::chol_solve(m, l, b, nrhs);
}

}  // namespace lira
