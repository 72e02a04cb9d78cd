// The LOBPCG Rayleigh–Ritz step on the device, in f64, for subspaces of up to RR_DEV_MAX columns: one workgroup of
// 256 threads with the eigensolver's working matrices in LDS (crates/qm-gpu/src/eigen.rs:515-542 and 669-681, the
// reference's host step, restated for the GPU). From the Gram of S = [X, W, P] against [S, R, HW, HP] and the
// residual norms it builds A and B (A's X columns through R), applies the iteration's stop test, and when the step
// goes on: the generalised Rayleigh–Ritz with D-scaling, the eigen-decomposition of B~ with directions below
// drop * max dropped, T = U Sigma^-1/2, the eigen-decomposition of T^T A~ T, C = D T Y, the amplification guard with
// the step redone without P when it fails, the locked head's projection Q = [-Byy^-1 Byz; I], and the P update
// coefficients C - [Bxx^-1 Bxs C; 0] by Cholesky. It writes the lincomb's padded coefficient block, the new Ritz values
// and a record the host reads one iteration later. A step that stops writes the identity on X and zeros elsewhere, so
// the lincomb and H X that follow it leave X and HX bit for bit as they were: iterations queued past the stop change
// nothing.
//
// The symmetric eigen-decompositions are cyclic Jacobi in the round-robin parallel order: every round rotates n/2
// disjoint pairs (p, q) at once, the matrix and the eigenvectors double-buffered in LDS (two barriers per round), until
// the off-diagonal norm is below 1e-15 of the matrix norm.
#include <cmath>

#include "lira.h"
#include "lira_internal.hpp"

namespace lira {

namespace {

constexpr int RT = 256;
constexpr int SLOT = RR_DEV_MAX * RR_DEV_MAX;

struct RrSmem {
    double a[2][SLOT];  // the Jacobi matrix, double-buffered
    double v[2][SLOT];  // its eigenvectors (columns), double-buffered
    double al[RR_DEV_MAX], be[RR_DEV_MAX];  // per index: column i of the round's rotation = al e_i + be e_partner
    int pa[RR_DEV_MAX];
    double w[RR_DEV_MAX], eb[RR_DEV_MAX], d[RR_DEV_MAX], th[RR_DEV_MAX], theta_in[RR_DEV_MAX];
    int keep[RR_DEV_MAX], perm[RR_DEV_MAX];
    double red[RT / 64];
    double scal[8];
    int iscal[8];
};

__device__ inline double block_sum(double x, RrSmem &sm) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sm.red[threadIdx.x >> 6] = x;
    __syncthreads();
    return (sm.red[0] + sm.red[1]) + (sm.red[2] + sm.red[3]);
}
__device__ inline double block_max(double x, RrSmem &sm) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x = fmax(x, __shfl_xor(x, off, 64));
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sm.red[threadIdx.x >> 6] = x;
    __syncthreads();
    return fmax(fmax(sm.red[0], sm.red[1]), fmax(sm.red[2], sm.red[3]));
}

// Symmetric eigen-decomposition of the n x n matrix in sm.a[0] (row-major): on return the eigenvalues are in sm.w
// (unsorted) and the eigenvectors the columns of sm.v[returned index]. Every thread calls it.
__device__ int jacobi_eigen(RrSmem &sm, int n) {
    const int tid = threadIdx.x;
    int cur = 0;
    for (int e = tid; e < n * n; e += RT) sm.v[0][e] = (e / n == e % n) ? 1.0 : 0.0;
    double f = 0.0;
    for (int e = tid; e < n * n; e += RT) f += sm.a[0][e] * sm.a[0][e];
    f = block_sum(f, sm);
    const int np = n + (n & 1);
    int sweep = 0;
    for (; sweep < 40 && n > 1; sweep++) {
        double off = 0.0;
        for (int e = tid; e < n * n; e += RT)
            if (e / n != e % n) off += sm.a[cur][e] * sm.a[cur][e];
        off = block_sum(off, sm);
        if (!(off > 1e-30 * f)) break;
        for (int round = 0; round < np - 1; round++) {
            if (tid < np / 2) {
                int p, q;
                if (tid == 0) {
                    p = np - 1;
                    q = round;
                } else {
                    p = (round + tid) % (np - 1);
                    q = (round - tid + np - 1) % (np - 1);
                }
                const double *A = sm.a[cur];
                if (p < n && q < n) {
                    const double apq = A[p * n + q];
                    double c = 1.0, s = 0.0;
                    if (apq != 0.0) {
                        const double th = (A[q * n + q] - A[p * n + p]) / (2.0 * apq);
                        const double t = (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
                        c = 1.0 / sqrt(t * t + 1.0);
                        s = t * c;
                    }
                    sm.al[p] = c;
                    sm.be[p] = -s;
                    sm.pa[p] = q;
                    sm.al[q] = c;
                    sm.be[q] = s;
                    sm.pa[q] = p;
                } else {
                    const int i = p < n ? p : q;  // the one real index of a pair with the padding index
                    if (i < n) {
                        sm.al[i] = 1.0;
                        sm.be[i] = 0.0;
                        sm.pa[i] = i;
                    }
                }
            }
            __syncthreads();
            const double *A = sm.a[cur], *V = sm.v[cur];
            double *An = sm.a[1 - cur], *Vn = sm.v[1 - cur];
            for (int e = tid; e < n * n; e += RT) {
                const int i = e / n, j = e - i * n;
                const int pj = sm.pa[j];
                const double aj = sm.al[j], bj = sm.be[j];
                Vn[e] = aj * V[i * n + j] + bj * V[i * n + pj];
                if (i > j) continue;
                const int pi = sm.pa[i];
                const double ai = sm.al[i], bi = sm.be[i];
                const double val = ai * (aj * A[i * n + j] + bj * A[i * n + pj]) +
                                   bi * (aj * A[pi * n + j] + bj * A[pi * n + pj]);
                An[i * n + j] = val;
                An[j * n + i] = val;
            }
            __syncthreads();
            cur = 1 - cur;
        }
    }
    for (int i = tid; i < n; i += RT) sm.w[i] = sm.a[cur][i * n + i];
    if (tid == 0) sm.iscal[6] += sweep;
    __syncthreads();
    return cur;
}

// Cholesky of the n x n SPD matrix a (global, row-major, lda) in place (lower triangle, L L^T); false on a
// non-positive pivot. Every thread calls it.
__device__ bool cholesky_dev(double *a, int n, int lda, RrSmem &sm) {
    const int tid = threadIdx.x;
    for (int j = 0; j < n; j++) {
        if (tid == 0) {
            double s = a[j * lda + j];
            for (int k = 0; k < j; k++) s -= a[j * lda + k] * a[j * lda + k];
            sm.iscal[0] = s > 0.0;
            a[j * lda + j] = s > 0.0 ? sqrt(s) : 0.0;
        }
        __syncthreads();
        if (!sm.iscal[0]) return false;
        const double ljj = a[j * lda + j];
        for (int i = j + 1 + tid; i < n; i += RT) {
            double s = a[i * lda + j];
            for (int k = 0; k < j; k++) s -= a[i * lda + k] * a[j * lda + k];
            a[i * lda + j] = s / ljj;
        }
        __syncthreads();
    }
    return true;
}

// b (n x nrhs, row-major, ldb) <- (L L^T)^-1 b, one thread per right-hand side
__device__ void chol_solve_dev(const double *l, int n, int lda, double *b, int nrhs, int ldb) {
    for (int c = threadIdx.x; c < nrhs; c += RT) {
        for (int i = 0; i < n; i++) {
            double s = b[i * ldb + c];
            for (int k = 0; k < i; k++) s -= l[i * lda + k] * b[k * ldb + c];
            b[i * ldb + c] = s / l[i * lda + i];
        }
        for (int i = n - 1; i >= 0; i--) {
            double s = b[i * ldb + c];
            for (int k = i + 1; k < n; k++) s -= l[k * lda + i] * b[k * ldb + c];
            b[i * ldb + c] = s / l[i * lda + i];
        }
    }
    __syncthreads();
}

// The scratch matrices of one step in global memory (every one RR_DEV_MAX^2 doubles).
struct RrWork {
    double *A, *B, *Ap, *Bp, *K, *T, *AT, *ATT, *Cr, *C, *M, *L;
};
__device__ inline RrWork work_of(double *w) {
    RrWork r;
    double *p = w;
    r.A = p, p += SLOT;
    r.B = p, p += SLOT;
    r.Ap = p, p += SLOT;
    r.Bp = p, p += SLOT;
    r.K = p, p += SLOT;
    r.T = p, p += SLOT;
    r.AT = p, p += SLOT;
    r.ATT = p, p += SLOT;
    r.Cr = p, p += SLOT;
    r.C = p, p += SLOT;
    r.M = p, p += 2 * SLOT;
    r.L = p;
    return r;
}

// Generalised Rayleigh–Ritz on the r x r pencil (Ap, Bp) (global, lda r): the nk lowest Ritz values into sm.th and the
// coefficients C (r x nk, row-major) into Cr; returns the amplification, or -1 when a column has a non-positive norm,
// -2 when fewer than nk directions survive the drop. Every thread calls it.
__device__ double rr_plain(RrSmem &sm, const RrWork &wk, const double *Ap, const double *Bp, int r, int nk,
                           double drop) {
    const int tid = threadIdx.x;
    if (tid == 0) {
        sm.iscal[1] = 1;
        for (int i = 0; i < r; i++)
            if (!(Bp[i * r + i] > 0.0)) sm.iscal[1] = 0;
    }
    __syncthreads();
    if (!sm.iscal[1]) return -1.0;
    for (int i = tid; i < r; i += RT) sm.d[i] = 1.0 / sqrt(Bp[i * r + i]);
    __syncthreads();
    // B~ = D B D into the eigensolver, A~ = D A D kept in AT
    for (int e = tid; e < r * r; e += RT) {
        const int i = e / r, j = e - i * r;
        sm.a[0][e] = Bp[e] * sm.d[i] * sm.d[j];
        wk.AT[e] = Ap[e] * sm.d[i] * sm.d[j];
    }
    __syncthreads();
    const int vb = jacobi_eigen(sm, r);
    for (int i = tid; i < r; i += RT) sm.eb[i] = sm.w[i];
    double smax = -INFINITY;
    for (int i = tid; i < r; i += RT) smax = fmax(smax, sm.w[i]);
    smax = block_max(smax, sm);
    if (tid == 0) {
        int nkeep = 0;
        for (int i = 0; i < r; i++)
            if (sm.eb[i] > drop * smax) sm.keep[nkeep++] = i;
        sm.iscal[2] = nkeep;
    }
    __syncthreads();
    const int nkeep = sm.iscal[2];
    if (nkeep < nk) return -2.0;
    // T = U(:, keep) Sigma^-1/2 (r x nkeep)
    for (int e = tid; e < r * nkeep; e += RT) {
        const int i = e / nkeep, c = e - i * nkeep;
        const int kc = sm.keep[c];
        wk.T[e] = sm.v[vb][i * r + kc] / sqrt(sm.eb[kc]);
    }
    __syncthreads();
    // A~ T (r x nkeep), then T^T (A~ T) (nkeep x nkeep) into the eigensolver, symmetrised
    for (int e = tid; e < r * nkeep; e += RT) {
        const int i = e / nkeep, c = e - i * nkeep;
        double s = 0.0;
        for (int k = 0; k < r; k++) s += wk.AT[i * r + k] * wk.T[k * nkeep + c];
        wk.ATT[e] = s;
    }
    __syncthreads();
    for (int e = tid; e < nkeep * nkeep; e += RT) {
        const int a = e / nkeep, b = e - a * nkeep;
        double s = 0.0;
        for (int i = 0; i < r; i++) s += wk.T[i * nkeep + a] * wk.ATT[i * nkeep + b];
        sm.a[1][e] = s;
    }
    __syncthreads();
    for (int e = tid; e < nkeep * nkeep; e += RT) {
        const int a = e / nkeep, b = e - a * nkeep;
        sm.a[0][e] = a == b ? sm.a[1][e] : 0.5 * (sm.a[1][a * nkeep + b] + sm.a[1][b * nkeep + a]);
    }
    __syncthreads();
    const int vy = jacobi_eigen(sm, nkeep);
    // ascending order: perm[rank] = index
    for (int i = tid; i < nkeep; i += RT) {
        int rank = 0;
        for (int j = 0; j < nkeep; j++) rank += sm.w[j] < sm.w[i] || (sm.w[j] == sm.w[i] && j < i);
        sm.perm[rank] = i;
    }
    __syncthreads();
    for (int c = tid; c < nk; c += RT) sm.th[c] = sm.w[sm.perm[c]];
    // alpha = max_c ||Sigma^-1/2 y_c||
    double al = 0.0;
    for (int c = tid; c < nk; c += RT) {
        const int yc = sm.perm[c];
        double s = 0.0;
        for (int dd = 0; dd < nkeep; dd++) {
            const double y = sm.v[vy][dd * nkeep + yc];
            s += y * y / sm.eb[sm.keep[dd]];
        }
        al = fmax(al, sqrt(s));
    }
    al = block_max(al, sm);
    // C = D T Y(:, perm[:nk])
    for (int e = tid; e < r * nk; e += RT) {
        const int i = e / nk, c = e - i * nk;
        const int yc = sm.perm[c];
        double s = 0.0;
        for (int k = 0; k < nkeep; k++) s += wk.T[i * nkeep + k] * sm.v[vy][k * nkeep + yc];
        wk.Cr[e] = s * sm.d[i];
    }
    __syncthreads();
    return al;
}

// rr_plain with the first `locked` columns held (eigen.rs:669-681): the pencil on Q = [-Byy^-1 Byz; I], C = Q C'.
// A, B: m x m (global, lda m); the result (m x na) in wk.C. Returns the amplification or a negative failure code (-3:
// Cholesky of Byy failed).
__device__ double rr_locked(RrSmem &sm, const RrWork &wk, const double *A, const double *B, int m, int locked, int na,
                            double drop) {
    const int tid = threadIdx.x;
    if (locked == 0) {
        const double al = rr_plain(sm, wk, A, B, m, na, drop);
        if (al < 0.0) return al;
        for (int e = tid; e < m * na; e += RT) wk.C[e] = wk.Cr[e];
        __syncthreads();
        return al;
    }
    const int r = m - locked;
    // Byy into L, Byz into K; K <- Byy^-1 Byz
    for (int e = tid; e < locked * locked; e += RT) wk.L[e] = B[(e / locked) * m + e % locked];
    for (int e = tid; e < locked * r; e += RT) wk.K[e] = B[(e / r) * m + locked + e % r];
    __syncthreads();
    if (!cholesky_dev(wk.L, locked, locked, sm)) return -3.0;
    chol_solve_dev(wk.L, locked, locked, wk.K, r, r);
    // Ap = Q^T A Q, Bp = Q^T B Q (symmetrised): MQ (m x r) in T, then the reduced matrix in AT, then symmetrise
    for (int pass = 0; pass < 2; pass++) {
        const double *M = pass == 0 ? A : B;
        double *out = pass == 0 ? wk.Ap : wk.Bp;
        for (int e = tid; e < m * r; e += RT) {
            const int i = e / r, j = e - i * r;
            double v = M[i * m + locked + j];
            for (int l = 0; l < locked; l++) v -= M[i * m + l] * wk.K[l * r + j];
            wk.T[e] = v;
        }
        __syncthreads();
        for (int e = tid; e < r * r; e += RT) {
            const int a = e / r, b = e - a * r;
            double v = wk.T[(locked + a) * r + b];
            for (int l = 0; l < locked; l++) v -= wk.K[l * r + a] * wk.T[l * r + b];
            wk.AT[e] = v;
        }
        __syncthreads();
        for (int e = tid; e < r * r; e += RT) {
            const int a = e / r, b = e - a * r;
            out[e] = a == b ? wk.AT[e] : 0.5 * (wk.AT[a * r + b] + wk.AT[b * r + a]);
        }
        __syncthreads();
    }
    const double al = rr_plain(sm, wk, wk.Ap, wk.Bp, r, na, drop);
    if (al < 0.0) return al;
    for (int e = tid; e < m * na; e += RT) {
        const int i = e / na, j = e - i * na;
        double v;
        if (i < locked) {
            v = 0.0;
            for (int b = 0; b < r; b++) v -= wk.K[i * r + b] * wk.Cr[b * na + j];
        } else {
            v = wk.Cr[(i - locked) * na + j];
        }
        wk.C[e] = v;
    }
    __syncthreads();
    return al;
}

__global__ __launch_bounds__(RT) void k_rr_step(RrStepArgs a) {
    __shared__ RrSmem sm;
    const int tid = threadIdx.x;
    const RrWork wk = work_of(a.work);
    const int m = a.m, k = a.k, q = a.q, nm = a.nm;
    double *rec = a.record;
    for (int i = tid; i < k; i += RT) sm.theta_in[i] = a.theta[i];
    __syncthreads();
    // A and B from the Gram's lower triangle (A's X columns restored from R: A_ij = S_i^T R_j + theta_j B_ij)
    bool finite = true;
    for (int e = tid; e < m * m; e += RT) {
        const int i = e / m, j = e - i * m;
        if (j > i) continue;
        const double bij = a.G[(long)i * q + j];
        double aij = a.G[(long)i * q + m + j];
        if (j < k) aij += sm.theta_in[j] * bij;
        wk.B[i * m + j] = wk.B[j * m + i] = bij;
        wk.A[i * m + j] = wk.A[j * m + i] = aij;
        if (!isfinite(aij)) finite = false;
    }
    __syncthreads();
    // a non-finite entry anywhere in A fails the step (the host's check over every entry)
    const double nonfinite = block_max(finite ? 0.0 : 1.0, sm);
    // the stop test (one thread; k <= RR_DEV_MAX)
    if (tid == 0) {
        double head_res = 0.0, tail_res = 0.0, lock_res = 0.0;
        bool bad = nonfinite > 0.0;
        for (int i = 0; i < k; i++) {
            const double xx = wk.B[i * m + i];
            const double lam = wk.A[i * m + i] / xx;
            const double res = sqrt(fmax(a.G[(long)m * q + i], 0.0) / xx);
            rec[RR_REC_HEAD + i] = lam;
            rec[RR_REC_HEAD + k + i] = res;
            if (!isfinite(lam) || !isfinite(res)) bad = true;
            if (i >= a.checked) continue;
            if (i < a.head) head_res = fmax(head_res, res);
            else {
                tail_res = fmax(tail_res, res);
                lock_res = fmax(lock_res, res * a.tail_scale);
            }
        }
        const double split_res = fmax(head_res, tail_res * a.tail_scale);
        const double stop_res = a.locked > 0 ? lock_res : split_res;
        const bool stop = stop_res < a.tol || (a.iter == 0 && stop_res < a.guess_tol) || a.iter >= a.max_iter;
        rec[0] = stop ? 1.0 : 0.0;
        rec[1] = bad ? 1.0 : 0.0;  // error code
        rec[2] = head_res;
        rec[3] = tail_res;
        rec[4] = 0.0;  // the step dropped P
        rec[5] = 0.0;  // its amplification
        rec[6] = 0.0;  // Jacobi sweeps of the step's eigen-decompositions
        rec[7] = 0.0;
        sm.iscal[6] = 0;
        sm.iscal[4] = stop;
        sm.iscal[5] = bad;
        const double budget_res = a.locked > 0 ? tail_res : stop_res;
        sm.scal[0] = a.is_f32 ? fmax(a.alpha_max, a.alpha_fraction * budget_res / fmax(a.floor, 1e-300)) : 1e300;
    }
    __syncthreads();
    const bool error = sm.iscal[5];
    const bool stop = sm.iscal[4];
    const int mp = a.mp, ldc = a.ldc, q2 = 2 * nm;
    if (stop || error) {
        // X' = X (identity on the moving columns), P' = 0
        for (int e = tid; e < mp * ldc; e += RT) {
            const int i = e / ldc, j = e - i * ldc;
            a.C[e] = (j < nm && i == a.locked + j) ? 1.0 : 0.0;
        }
        return;
    }
    // the Rayleigh–Ritz step, without P when it loses rank or its amplification exceeds the budget
    int mm = m;
    double al = rr_locked(sm, wk, wk.A, wk.B, m, a.locked, nm, a.drop);
    const bool failed = al < 0.0;
    if ((failed || al > sm.scal[0]) && a.np > 0) {
        mm = k + a.na;
        // the leading mm x mm block of A and B, repacked at lda mm
        for (int e = tid; e < mm * mm; e += RT) {
            const int i = e / mm, j = e - i * mm;
            wk.M[e] = wk.A[i * m + j];
            wk.M[SLOT + e] = wk.B[i * m + j];
        }
        __syncthreads();
        for (int e = tid; e < mm * mm; e += RT) {
            wk.A[e] = wk.M[e];
            wk.B[e] = wk.M[SLOT + e];
        }
        __syncthreads();
        al = rr_locked(sm, wk, wk.A, wk.B, mm, a.locked, nm, a.drop);
        if (tid == 0) rec[4] = 1.0;
    }
    if (al < 0.0) {
        if (tid == 0) rec[1] = al == -2.0 ? 2.0 : al == -1.0 ? 4.0 : 3.0;
        __syncthreads();
        for (int e = tid; e < mp * ldc; e += RT) {
            const int i = e / ldc, j = e - i * ldc;
            a.C[e] = (j < nm && i == a.locked + j) ? 1.0 : 0.0;
        }
        return;
    }
    if (tid == 0) {
        rec[5] = al;
        rec[6] = sm.iscal[6];
    }
    // P' coefficients: C - [Bxx^-1 Bxs C; 0], Bxx the k x k head of B (lda mm)
    for (int e = tid; e < k * k; e += RT) wk.L[e] = wk.B[(e / k) * mm + e % k];
    for (int e = tid; e < k * nm; e += RT) {
        const int i = e / nm, j = e - i * nm;
        double s = 0.0;
        for (int x = 0; x < mm; x++) s += wk.B[i * mm + x] * wk.C[x * nm + j];
        wk.M[e] = s;
    }
    __syncthreads();
    if (!cholesky_dev(wk.L, k, k, sm)) {
        if (tid == 0) rec[1] = 3.0;
        __syncthreads();
        for (int e = tid; e < mp * ldc; e += RT) {
            const int i = e / ldc, j = e - i * ldc;
            a.C[e] = (j < nm && i == a.locked + j) ? 1.0 : 0.0;
        }
        return;
    }
    chol_solve_dev(wk.L, k, k, wk.M, nm, nm);
    // the padded coefficient block: rows < mm hold [C | C_p]; rows mm .. mp-1 and columns 2 nm .. ldc-1 are zero
    for (int e = tid; e < mp * ldc; e += RT) {
        const int i = e / ldc, j = e - i * ldc;
        double v = 0.0;
        if (i < mm && j < q2) {
            const int jj = j < nm ? j : j - nm;
            v = wk.C[i * nm + jj];
            if (j >= nm && i < k) v -= wk.M[i * nm + jj];
        }
        a.C[e] = v;
    }
    for (int c = tid; c < nm; c += RT) a.theta[a.locked + c] = sm.th[c];
}

}  // namespace

size_t rr_step_work_doubles() { return (size_t)14 * SLOT; }

void rr_step(const RrStepArgs &args, hipStream_t s) {
    hipLaunchKernelGGL(k_rr_step, dim3(1), dim3(RT), 0, s, args);
    LIRA_HIP(hipGetLastError());
}

}  // namespace lira
