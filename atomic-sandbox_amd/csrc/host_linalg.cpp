// Host f64 dense linear algebra of the solver's small problems: the symmetric eigensolver, LU and Cholesky, and the
// generalised Rayleigh–Ritz step of LOBPCG (m <= a few hundred). These run on the CPU in f64 in every precision mode.
#include <algorithm>
#include <limits>
#include <cmath>
#include <numeric>

#include "lira_internal.hpp"

// This file is compiled twice: as lira::linalg_base for any x86-64 host, and as lira::linalg_avx2 with
// LIRA_LINALG_AVX2 defined; host_linalg_dispatch.cpp defines the lira:: entry points and picks the build the host CPU
// runs. The AVX2 + FMA target is given to the functions defined below only (a pragma, not -mavx2): the standard-library
// code the two objects instantiate (std::vector members and the like) is emitted as baseline code in both, so whichever
// copy the linker keeps of such a shared inline function runs on any x86-64 host, whatever the link order.
#ifndef LIRA_LINALG_NS
#define LIRA_LINALG_NS linalg_base
#endif

namespace lira {
namespace LIRA_LINALG_NS {
#ifdef LIRA_LINALG_AVX2
#pragma clang attribute push(__attribute__((target("avx2,fma"))), apply_to = function)
#endif

// Cyclic Jacobi (Golub & Van Loan §8.5; the reference's fallback eigensolver, crates/qm/src/eigen.rs:175-211):
// rotations until the off-diagonal norm is below 1e-15 ||A||; pairs returned ascending. a is m x m row-major.
static void jacobi_eigen(int m, std::vector<double> &a, std::vector<double> &evals, std::vector<double> &evecs) {
    std::vector<double> v((size_t)m * m, 0.0);
    for (int i = 0; i < m; i++) v[(size_t)i * m + i] = 1.0;
    double norm = 0.0;
    for (double x : a) norm += x * x;
    norm = std::sqrt(norm);
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0.0;
        for (int i = 0; i < m; i++)
            for (int j = 0; j < m; j++)
                if (i != j) off += a[(size_t)i * m + j] * a[(size_t)i * m + j];
        if (!(std::sqrt(off) > 1e-15 * norm)) break;
        for (int p = 0; p < m; p++)
            for (int q = p + 1; q < m; q++) {
                const double apq = a[(size_t)p * m + q];
                if (apq == 0.0) continue;
                const double theta = (a[(size_t)q * m + q] - a[(size_t)p * m + p]) / (2.0 * apq);
                double t = (theta > 0 ? 1.0 : (theta < 0 ? -1.0 : 0.0)) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                if (theta == 0.0) t = 1.0;
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < m; k++) {
                    const double akp = a[(size_t)k * m + p], akq = a[(size_t)k * m + q];
                    a[(size_t)k * m + p] = c * akp - s * akq;
                    a[(size_t)k * m + q] = s * akp + c * akq;
                }
                for (int k = 0; k < m; k++) {
                    const double apk = a[(size_t)p * m + k], aqk = a[(size_t)q * m + k];
                    a[(size_t)p * m + k] = c * apk - s * aqk;
                    a[(size_t)q * m + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < m; k++) {
                    const double vkp = v[(size_t)k * m + p], vkq = v[(size_t)k * m + q];
                    v[(size_t)k * m + p] = c * vkp - s * vkq;
                    v[(size_t)k * m + q] = s * vkp + c * vkq;
                }
            }
    }
    std::vector<int> order(m);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int x, int y) { return a[(size_t)x * m + x] < a[(size_t)y * m + y]; });
    evals.assign(m, 0.0);
    evecs.assign((size_t)m * m, 0.0);
    for (int c = 0; c < m; c++) {
        evals[c] = a[(size_t)order[c] * m + order[c]];
        for (int r = 0; r < m; r++) evecs[(size_t)r * m + c] = v[(size_t)r * m + order[c]];
    }
}

// sqrt(f^2 + g^2) for the QL rotations: the subspace matrices here are scaled to O(1) (unit-diagonal B, energies in
// A), far from the overflow std::hypot guards against, and hypot's careful path is the bulk of a rotation's latency
static inline double norm2(double f, double g) {
    const double af = std::fabs(f), ag = std::fabs(g);
    if (af > 1e150 || ag > 1e150 || (af < 1e-150 && ag < 1e-150)) return std::hypot(f, g);
    return std::sqrt(f * f + g * g);
}

// dot product of n contiguous values with four interleaved partial sums (a fixed order, so deterministic; the
// compiler maps the four sums onto vector lanes)
static inline double dot4(const double *__restrict__ x, const double *__restrict__ y, int n) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int j = 0;
    for (; j + 4 <= n; j += 4) {
        s0 += x[j] * y[j];
        s1 += x[j + 1] * y[j + 1];
        s2 += x[j + 2] * y[j + 2];
        s3 += x[j + 3] * y[j + 3];
    }
    for (; j < n; j++) s0 += x[j] * y[j];
    return (s0 + s1) + (s2 + s3);
}

// Householder reduction to tridiagonal form followed by the implicit-shift QL iteration (Golub & Van Loan §8.3),
// accumulating the orthogonal transformations. Q is kept column-major (qc[c m + r] = Q[r][c]), so the Householder
// update and every QL rotation run over contiguous memory. Returns false if the QL iteration does not converge.
static bool tridiagonal_ql_eigen(int m, std::vector<double> a, std::vector<double> &evals, std::vector<double> &evecs) {
    std::vector<double> qcv((size_t)m * m, 0.0), vv_(m), pv(m), wv(m), qvv(m);
    double *__restrict__ qc = qcv.data();
    double *__restrict__ A = a.data();
    double *__restrict__ v = vv_.data();
    double *__restrict__ p = pv.data();
    double *__restrict__ w = wv.data();
    double *__restrict__ qv = qvv.data();
    for (int i = 0; i < m; i++) qc[(size_t)i * m + i] = 1.0;
    for (int k = 0; k + 2 < m; k++) {
        const int len = m - k - 1;
        double xn = 0.0;
        for (int i = 0; i < len; i++) {
            v[i] = A[(size_t)(k + 1 + i) * m + k];
            xn += v[i] * v[i];
        }
        if (xn == 0.0) continue;
        double alpha = std::sqrt(xn);
        if (v[0] > 0) alpha = -alpha;
        v[0] -= alpha;  // v = x - alpha e1, H x = alpha e1
        const double vv = dot4(v, v, len);
        if (vv == 0.0) continue;
        const double beta = 2.0 / vv;
        // trailing block B <- H B H: p = beta B v, K = beta/2 v.p, w = p - K v, B -= v w^T + w v^T
        for (int i = 0; i < len; i++) p[i] = beta * dot4(&A[(size_t)(k + 1 + i) * m + k + 1], v, len);
        const double kk = 0.5 * beta * dot4(v, p, len);
        for (int i = 0; i < len; i++) w[i] = p[i] - kk * v[i];
        for (int i = 0; i < len; i++) {
            double *__restrict__ row = &A[(size_t)(k + 1 + i) * m + k + 1];
            const double vi = v[i], wi = w[i];
            for (int j = 0; j < len; j++) row[j] -= vi * w[j] + wi * v[j];
        }
        A[(size_t)(k + 1) * m + k] = A[(size_t)k * m + k + 1] = alpha;
        for (int i = 1; i < len; i++) A[(size_t)(k + 1 + i) * m + k] = A[(size_t)k * m + k + 1 + i] = 0.0;
        // Q <- Q H on columns k+1..: qv = beta Q[:, k+1:] v, then Q[:, k+1+j] -= qv v_j
        for (int r = 0; r < m; r++) qv[r] = 0.0;
        for (int j = 0; j < len; j++) {
            const double *__restrict__ col = &qc[(size_t)(k + 1 + j) * m];
            const double vj = v[j];
            for (int r = 0; r < m; r++) qv[r] += col[r] * vj;
        }
        for (int r = 0; r < m; r++) qv[r] *= beta;
        for (int j = 0; j < len; j++) {
            double *__restrict__ col = &qc[(size_t)(k + 1 + j) * m];
            const double vj = v[j];
            for (int r = 0; r < m; r++) col[r] -= qv[r] * vj;
        }
    }
    std::vector<double> d(m), e(m, 0.0);
    for (int i = 0; i < m; i++) d[i] = A[(size_t)i * m + i];
    for (int i = 0; i + 1 < m; i++) e[i] = A[(size_t)(i + 1) * m + i];
    // implicit QL on (d, e): e[i] couples rows i and i + 1
    const double eps = std::numeric_limits<double>::epsilon();
    for (int l = 0; l < m; l++) {
        int iter = 0, mm;
        do {
            for (mm = l; mm < m - 1; mm++) {
                const double dd = std::fabs(d[mm]) + std::fabs(d[mm + 1]);
                if (std::fabs(e[mm]) <= eps * dd) break;
            }
            if (mm != l) {
                if (iter++ == 60) return false;
                double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
                double r = norm2(g, 1.0);
                g = d[mm] - d[l] + e[l] / (g + std::copysign(r, g));
                double s = 1.0, c = 1.0, pp = 0.0;
                int i;
                for (i = mm - 1; i >= l; i--) {
                    double f = s * e[i];
                    const double b = c * e[i];
                    r = norm2(f, g);
                    e[i + 1] = r;
                    if (r == 0.0) {
                        d[i + 1] -= pp;
                        e[mm] = 0.0;
                        break;
                    }
                    const double inv = 1.0 / r;
                    s = f * inv;
                    c = g * inv;
                    g = d[i + 1] - pp;
                    r = (d[i] - g) * s + 2.0 * c * b;
                    pp = s * r;
                    d[i + 1] = g + pp;
                    g = c * r - b;
                    // rotate columns i and i + 1 of Q
                    double *__restrict__ q0 = &qc[(size_t)i * m];
                    double *__restrict__ q1 = &qc[(size_t)(i + 1) * m];
                    for (int k = 0; k < m; k++) {
                        const double fk = q1[k];
                        q1[k] = s * q0[k] + c * fk;
                        q0[k] = c * q0[k] - s * fk;
                    }
                }
                if (r == 0.0 && i >= l) continue;
                d[l] -= pp;
                e[l] = g;
                e[mm] = 0.0;
            }
        } while (mm != l);
    }
    std::vector<int> order(m);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return d[x] < d[y]; });
    evals.assign(m, 0.0);
    evecs.assign((size_t)m * m, 0.0);
    for (int c = 0; c < m; c++) {
        evals[c] = d[order[c]];
        const double *col = &qc[(size_t)order[c] * m];
        for (int r = 0; r < m; r++) evecs[(size_t)r * m + c] = col[r];
    }
    return true;
}

// The checked symmetric eigen-decomposition (crates/qm/src/eigen.rs:141-170): Householder + implicit QL, then every
// pair's residual ||A v - lambda v|| against 1e-11 (1 + max |lambda|); a decomposition that fails the check (or a QL
// that does not converge) is replaced by cyclic Jacobi. Pairs ascending.
void sym_eigen(int m, std::vector<double> &a, std::vector<double> &evals, std::vector<double> &evecs) {
    if (m <= 0) {
        evals.clear();
        evecs.clear();
        return;
    }
    if (tridiagonal_ql_eigen(m, a, evals, evecs)) {
        double scale = 1.0;
        for (double x : evals) scale = std::max(scale, 1.0 + std::fabs(x));
        // R = A V - V diag(lambda), row by row with contiguous inner loops; every column's norm against the bound
        std::vector<double> res((size_t)m * m), col2(m, 0.0);
        for (int i = 0; i < m; i++) {
            double *__restrict__ ri = &res[(size_t)i * m];
            const double *__restrict__ vi = &evecs[(size_t)i * m];
            for (int c = 0; c < m; c++) ri[c] = -evals[c] * vi[c];
            for (int j = 0; j < m; j++) {
                const double aij = a[(size_t)i * m + j];
                const double *__restrict__ vj = &evecs[(size_t)j * m];
                for (int c = 0; c < m; c++) ri[c] += aij * vj[c];
            }
            for (int c = 0; c < m; c++) col2[c] += ri[c] * ri[c];
        }
        bool ok = true;
        for (int c = 0; c < m && ok; c++) ok = std::sqrt(col2[c]) <= 1e-11 * scale;
        if (ok) return;
    }
    jacobi_eigen(m, a, evals, evecs);
}

// LU with partial pivoting; false on a zero pivot (nalgebra's LU::solve returning None, crates/qm/src/scf.rs:696)
bool lu_solve(int m, std::vector<double> a, std::vector<double> &b) {
    for (int c = 0; c < m; c++) {
        int piv = c;
        double best = std::fabs(a[(size_t)c * m + c]);
        for (int r = c + 1; r < m; r++)
            if (std::fabs(a[(size_t)r * m + c]) > best) {
                best = std::fabs(a[(size_t)r * m + c]);
                piv = r;
            }
        if (best == 0.0) return false;
        if (piv != c) {
            for (int k = 0; k < m; k++) std::swap(a[(size_t)c * m + k], a[(size_t)piv * m + k]);
            std::swap(b[c], b[piv]);
        }
        for (int r = c + 1; r < m; r++) {
            const double f = a[(size_t)r * m + c] / a[(size_t)c * m + c];
            for (int k = c; k < m; k++) a[(size_t)r * m + k] -= f * a[(size_t)c * m + k];
            b[r] -= f * b[c];
        }
    }
    for (int r = m - 1; r >= 0; r--) {
        double s = b[r];
        for (int k = r + 1; k < m; k++) s -= a[(size_t)r * m + k] * b[k];
        b[r] = s / a[(size_t)r * m + r];
    }
    return true;
}

bool cholesky(int m, std::vector<double> &a) {
    for (int j = 0; j < m; j++) {
        double d = a[(size_t)j * m + j];
        for (int k = 0; k < j; k++) d -= a[(size_t)j * m + k] * a[(size_t)j * m + k];
        if (!(d > 0.0)) return false;
        d = std::sqrt(d);
        a[(size_t)j * m + j] = d;
        for (int i = j + 1; i < m; i++) {
            double s = a[(size_t)i * m + j];
            for (int k = 0; k < j; k++) s -= a[(size_t)i * m + k] * a[(size_t)j * m + k];
            a[(size_t)i * m + j] = s / d;
        }
        for (int k = j + 1; k < m; k++) a[(size_t)j * m + k] = 0.0;
    }
    return true;
}

// solves (L L^T) X = B for nrhs columns of B (row-major m x nrhs), in place
void chol_solve(int m, const std::vector<double> &l, std::vector<double> &b, int nrhs) {
    for (int c = 0; c < nrhs; c++) {
        for (int i = 0; i < m; i++) {
            double s = b[(size_t)i * nrhs + c];
            for (int k = 0; k < i; k++) s -= l[(size_t)i * m + k] * b[(size_t)k * nrhs + c];
            b[(size_t)i * nrhs + c] = s / l[(size_t)i * m + i];
        }
        for (int i = m - 1; i >= 0; i--) {
            double s = b[(size_t)i * nrhs + c];
            for (int k = i + 1; k < m; k++) s -= l[(size_t)k * m + i] * b[(size_t)k * nrhs + c];
            b[(size_t)i * nrhs + c] = s / l[(size_t)i * m + i];
        }
    }
}

// Generalised Rayleigh–Ritz A c = theta B c on a raw (not orthonormal) subspace (crates/qm-gpu/src/eigen.rs:515-542):
// D-scaling to unit diagonal, the eigen-decomposition of B~ with directions below drop * sigma_max removed,
// T = U Sigma^{-1/2}, the standard problem T^T A~ T, and C = D T Y. Returns alpha = max_i ||Sigma^{-1/2} y_i||, the
// coefficient amplification of the step (1 for an orthonormal subspace).
double rayleigh_ritz(int m, const std::vector<double> &A, const std::vector<double> &B, int nk, double drop,
                     std::vector<double> &theta, std::vector<double> &C) {
    for (int i = 0; i < m; i++)
        if (!(B[(size_t)i * m + i] > 0.0))
            throw NumericalError("lobpcg: column " + std::to_string(i) + " of the subspace has a non-positive norm");
    std::vector<double> d(m);
    for (int i = 0; i < m; i++) d[i] = 1.0 / std::sqrt(B[(size_t)i * m + i]);
    std::vector<double> an((size_t)m * m), bn((size_t)m * m);
    for (int i = 0; i < m; i++)
        for (int j = 0; j < m; j++) {
            an[(size_t)i * m + j] = A[(size_t)i * m + j] * d[i] * d[j];
            bn[(size_t)i * m + j] = B[(size_t)i * m + j] * d[i] * d[j];
        }
    std::vector<double> eb, ub;
    sym_eigen(m, bn, eb, ub);
    const double smax = *std::max_element(eb.begin(), eb.end());
    std::vector<int> keep;
    for (int i = 0; i < m; i++)
        if (eb[i] > drop * smax) keep.push_back(i);
    const int nkeep = (int)keep.size();
    if (nkeep < nk)
        throw NumericalError("lobpcg: the subspace lost rank (" + std::to_string(nkeep) + " of " + std::to_string(m) +
                             " directions kept, " + std::to_string(nk) + " needed)");
    std::vector<double> t((size_t)m * nkeep);
    for (int i = 0; i < m; i++)
        for (int c = 0; c < nkeep; c++) t[(size_t)i * nkeep + c] = ub[(size_t)i * m + keep[c]] / std::sqrt(eb[keep[c]]);
    // ah = T^T an T
    std::vector<double> at((size_t)m * nkeep, 0.0);
    for (int i = 0; i < m; i++)
        for (int k = 0; k < m; k++) {
            const double aik = an[(size_t)i * m + k];
            if (aik == 0.0) continue;
            for (int c = 0; c < nkeep; c++) at[(size_t)i * nkeep + c] += aik * t[(size_t)k * nkeep + c];
        }
    std::vector<double> ah((size_t)nkeep * nkeep, 0.0);
    for (int i = 0; i < m; i++)
        for (int r = 0; r < nkeep; r++) {
            const double tir = t[(size_t)i * nkeep + r];
            const double *__restrict__ ati = &at[(size_t)i * nkeep];
            double *__restrict__ ahr = &ah[(size_t)r * nkeep];
            for (int c = 0; c < nkeep; c++) ahr[c] += tir * ati[c];
        }
    for (int r = 0; r < nkeep; r++)
        for (int c = r + 1; c < nkeep; c++) {
            const double s = 0.5 * (ah[(size_t)r * nkeep + c] + ah[(size_t)c * nkeep + r]);
            ah[(size_t)r * nkeep + c] = ah[(size_t)c * nkeep + r] = s;
        }
    std::vector<double> ea, ya;
    sym_eigen(nkeep, ah, ea, ya);
    theta.assign(ea.begin(), ea.begin() + nk);
    double alpha = 0.0;
    for (int c = 0; c < nk; c++) {
        double s = 0.0;
        for (int dd = 0; dd < nkeep; dd++) s += ya[(size_t)dd * nkeep + c] * ya[(size_t)dd * nkeep + c] / eb[keep[dd]];
        alpha = std::max(alpha, std::sqrt(s));
    }
    C.assign((size_t)m * nk, 0.0);
    for (int i = 0; i < m; i++) {
        double *__restrict__ ci = &C[(size_t)i * nk];
        for (int k = 0; k < nkeep; k++) {
            const double tik = t[(size_t)i * nkeep + k];
            const double *__restrict__ yk = &ya[(size_t)k * nkeep];
            for (int c = 0; c < nk; c++) ci[c] += tik * yk[c];
        }
        for (int c = 0; c < nk; c++) ci[c] *= d[i];
    }
    return alpha;
}

#ifdef LIRA_LINALG_AVX2
#pragma clang attribute pop
#endif
}  // namespace LIRA_LINALG_NS
}  // namespace lira
