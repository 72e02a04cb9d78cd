/*
 * lira_oracle.c — CPU f64 restatement of the reference's CPU Kohn–Sham solver (crates/qm).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path (see lira_oracle.h). Each function cites the
 * reference file:line it restates. Differences from the reference, all confined to paths that do not change a
 * converged answer:
 *   - the dense symmetric eigensolver is cyclic Jacobi throughout (the reference calls nalgebra 0.33.3's
 *     SymmetricEigen and falls back to the same cyclic Jacobi when a pair residual fails, qm/src/eigen.rs:150-211);
 *   - erf is glibc's (the reference uses libm 0.2.16's erf; both are correctly rounded to < 1 ulp);
 *   - under Fermi–Dirac every block column is held to eig_tol (the reference's `Tiers::AllTight` test path,
 *     qm/src/scf.rs:794-800) and the cold density is the sigma = 1 blob guess (the reference's Aufbau guess; its FD
 *     path starts from free-atom densities, which changes the step count, not the converged state).
 * Parity pins for every function are in tests/test_oracle.py.
 */
#include "lira_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* the OpenMP thread count of the grid loops (the reference's CPU solver is single-threaded: 1 is the faithful run) */
void or_set_threads(int threads) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
}
int or_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

static const double COEFF[5] = {-205.0 / 72.0, 8.0 / 5.0, -1.0 / 5.0, 8.0 / 315.0, -1.0 / 560.0};

static double now_s(void) {
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

static double *alloc0(long len) { return (double *)calloc((size_t)(len > 0 ? len : 1), sizeof(double)); }

static double origin_of(int n, double h) { return -((double)n - 1.0) * h / 2.0; }

/* qm/src/grid.rs:91-99 */
double or_neumaier(const double *x, long len) {
    double sum = 0.0, comp = 0.0;
    for (long i = 0; i < len; i++) {
        double t = sum + x[i];
        if (fabs(sum) >= fabs(x[i])) comp += (sum - t) + x[i];
        else comp += (x[i] - t) + sum;
        sum = t;
    }
    return sum + comp;
}

/* Neumaier over a*b without a temporary (grid.rs:36-39) */
static double neumaier_prod(const double *a, const double *b, long len) {
    double sum = 0.0, comp = 0.0;
    for (long i = 0; i < len; i++) {
        double x = a[i] * b[i];
        double t = sum + x;
        if (fabs(sum) >= fabs(x)) comp += (sum - t) + x;
        else comp += (x - t) + sum;
        sum = t;
    }
    return sum + comp;
}

double or_dot(int n, double h, const double *a, const double *b) {
    long len = (long)n * n * n;
    return neumaier_prod(a, b, len) * h * h * h;
}

double or_integrate(int n, double h, const double *a) { return or_neumaier(a, (long)n * n * n) * h * h * h; }

/* qm/src/stencil.rs:24-46: acc = 3 c0 f + sum_axes sum_m c_m (f_+ + f_-), Dirichlet (dropped terms). */
void or_laplacian(int n, double h, const double *f, double *out) {
    const double inv_h2 = 1.0 / (h * h);
    const long s[3] = {1, n, (long)n * n};
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) {
                long id = i + (long)n * (j + (long)n * k);
                double acc = 3.0 * COEFF[0] * f[id];
                int pos[3] = {i, j, k};
                for (int a = 0; a < 3; a++)
                    for (int m = 1; m <= 4; m++) {
                        double c = COEFF[m];
                        if (pos[a] + m < n) acc += c * f[id + m * s[a]];
                        if (pos[a] >= m) acc += c * f[id - m * s[a]];
                    }
                out[id] = acc * inv_h2;
            }
}

/* qm/src/stencil.rs:49-53 */
void or_kinetic(int n, double h, const double *f, double *out) {
    or_laplacian(n, h, f, out);
    long len = (long)n * n * n;
    for (long i = 0; i < len; i++) out[i] *= -0.5;
}

/* qm/src/eigen.rs:124-129 */
void or_hamiltonian(int n, double h, const double *v, const double *x, double *out) {
    or_kinetic(n, h, x, out);
    long len = (long)n * n * n;
    for (long i = 0; i < len; i++) out[i] += v[i] * x[i];
}

/* qm/src/eigen.rs:218-222 */
double or_jacobi_omega(void) {
    double symbol_pi = COEFF[0] + 2.0 * (-COEFF[1] + COEFF[2] - COEFF[3] + COEFF[4]);
    double rho = symbol_pi / COEFF[0];
    return 4.0 / (3.0 * rho);
}

/* qm/src/eigen.rs:227-240 (JACOBI_SWEEPS = 4) */
void or_jacobi_precond(int n, double h, const double *v, double lambda, const double *r, double *w) {
    long len = (long)n * n * n;
    double diag_t = -1.5 * COEFF[0] / (h * h);
    double omega = or_jacobi_omega();
    double *inv_d = alloc0(len), *hw = alloc0(len);
    for (long i = 0; i < len; i++) {
        double d = diag_t + v[i] - lambda;
        inv_d[i] = 1.0 / (d > 0.1 ? d : 0.1);
    }
    for (long i = 0; i < len; i++) w[i] = r[i] * inv_d[i];
    for (int sweep = 1; sweep < 4; sweep++) {
        or_hamiltonian(n, h, v, w, hw);
        for (long i = 0; i < len; i++) w[i] += omega * inv_d[i] * (r[i] - (hw[i] - lambda * w[i]));
    }
    free(inv_d);
    free(hw);
}

/* ---------------------------------------------------------------- Poisson ---- */

static int lm_index(int l, int m) { return l * l + (m + l); }

/* qm/src/poisson.rs:97-130: complex Racah-normalised C_l^m by the diagonal ladder and the three-term recursion */
void or_solid_harmonics(const double r[3], double out[OR_N_MOMENTS]) {
    double x = r[0], y = r[1], z = r[2];
    double r2 = x * x + y * y + z * z;
    double cr[OR_L_MAX + 1][OR_L_MAX + 1], ci[OR_L_MAX + 1][OR_L_MAX + 1];
    memset(cr, 0, sizeof cr);
    memset(ci, 0, sizeof ci);
    cr[0][0] = 1.0;
    for (int l = 1; l <= OR_L_MAX; l++) {
        double lf = (double)l;
        double f = sqrt((2.0 * lf - 1.0) / (2.0 * lf));
        double pr = cr[l - 1][l - 1], pi = ci[l - 1][l - 1];
        cr[l][l] = f * (x * pr - y * pi);
        ci[l][l] = f * (x * pi + y * pr);
        for (int m = 0; m < l; m++) {
            double mf = (double)m;
            double a = (2.0 * lf - 1.0) * z;
            double r1 = cr[l - 1][m], i1 = ci[l - 1][m];
            double r0 = 0.0, i0 = 0.0;
            if (l >= 2 && m <= l - 2) {
                r0 = cr[l - 2][m];
                i0 = ci[l - 2][m];
            }
            double t = (lf - 1.0) * (lf - 1.0) - mf * mf;
            double b = sqrt(t > 0.0 ? t : 0.0) * r2;
            double d = sqrt(lf * lf - mf * mf);
            cr[l][m] = (a * r1 - b * r0) / d;
            ci[l][m] = (a * i1 - b * i0) / d;
        }
    }
    double s2 = sqrt(2.0);
    for (int l = 0; l <= OR_L_MAX; l++) {
        out[lm_index(l, 0)] = cr[l][0];
        for (int m = 1; m <= l; m++) {
            out[lm_index(l, m)] = s2 * cr[l][m];
            out[lm_index(l, -m)] = s2 * ci[l][m];
        }
    }
}

static void grid_point(int n, double h, long idx, double p[3]) {
    double o = origin_of(n, h);
    long i = idx % n, j = (idx / n) % n, k = idx / ((long)n * n);
    p[0] = o + (double)i * h;
    p[1] = o + (double)j * h;
    p[2] = o + (double)k * h;
}

/* qm/src/poisson.rs:155-162: Q_lm = grid.dot(rho, R_lm) per moment, each one compensated sum */
void or_multipole_moments(int n, double h, const double *rho, double q[OR_N_MOMENTS]) {
    long len = (long)n * n * n;
    double *harm = alloc0(len * OR_N_MOMENTS);
    double hv[OR_N_MOMENTS], p[3];
    for (long i = 0; i < len; i++) {
        grid_point(n, h, i, p);
        or_solid_harmonics(p, hv);
        for (int lm = 0; lm < OR_N_MOMENTS; lm++) harm[lm * len + i] = hv[lm];
    }
    for (int lm = 0; lm < OR_N_MOMENTS; lm++) q[lm] = neumaier_prod(rho, harm + lm * len, len) * h * h * h;
    free(harm);
}

/* qm/src/poisson.rs:175-190 */
double or_multipole_potential(const double q[OR_N_MOMENTS], const double r[3]) {
    double d2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
    double hv[OR_N_MOMENTS];
    or_solid_harmonics(r, hv);
    double acc = 0.0, inv_pow = 1.0 / sqrt(d2);
    for (int l = 0; l <= OR_L_MAX; l++) {
        double s = 0.0;
        for (int m = -l; m <= l; m++) {
            int i = lm_index(l, m);
            s += q[i] * hv[i];
        }
        acc += s * inv_pow;
        inv_pow /= d2;
    }
    return acc;
}

/* qm/src/poisson.rs:174-197 + 204-214: b = 4 pi rho + L_bc with ghost values from the moments of rho */
void or_poisson_rhs(int n, double h, const double *rho, double *b) {
    long len = (long)n * n * n;
    double q[OR_N_MOMENTS];
    or_multipole_moments(n, h, rho, q);
    for (long i = 0; i < len; i++) b[i] = 4.0 * M_PI * rho[i];
    const double inv_h2 = 1.0 / (h * h);
    for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) {
                int pos[3] = {i, j, k};
                double acc = 0.0;
                int touched = 0;
                for (int a = 0; a < 3; a++) {
                    if (pos[a] >= 4 && pos[a] + 4 < n) continue;
                    touched = 1;
                }
                if (!touched) continue;
                long id = i + (long)n * (j + (long)n * k);
                for (int a = 0; a < 3; a++)
                    for (int m = 1; m <= 4; m++)
                        for (int sg = -1; sg <= 1; sg += 2) {
                            int qa = pos[a] + sg * m;
                            if (qa < 0 || qa >= n) {
                                double r[3];
                                grid_point(n, h, id, r);
                                r[a] += (double)(sg * m) * h;
                                acc += COEFF[m] * or_multipole_potential(q, r);
                            }
                        }
                b[id] += acc * inv_h2;
            }
}

/* qm/src/poisson.rs:200-255: Jacobi-PCG (equal to plain CG on the constant diagonal) on A V = b, A = -L_int */
int or_poisson_solve(int n, double h, const double *rho, double *v, double tol, int max_iter, double *residual) {
    long len = (long)n * n * n;
    double *b = alloc0(len), *r = alloc0(len), *z = alloc0(len), *p = alloc0(len), *ap = alloc0(len);
    or_poisson_rhs(n, h, rho, b);
    or_laplacian(n, h, v, ap);
    for (long i = 0; i < len; i++) r[i] = b[i] + ap[i]; /* r = b - A v, A v = -lap v */
    double inv_diag = -(h * h) / (3.0 * COEFF[0]);
    double b_norm = sqrt(or_dot(n, h, b, b));
    if (!(b_norm > 1e-300)) b_norm = 1e-300;
    for (long i = 0; i < len; i++) z[i] = inv_diag * r[i];
    memcpy(p, z, (size_t)len * sizeof(double));
    double rz = or_dot(n, h, r, z);
    int it = 0;
    double res = sqrt(or_dot(n, h, r, r)) / b_norm;
    while (res > tol && it < max_iter) {
        or_laplacian(n, h, p, ap);
        for (long i = 0; i < len; i++) ap[i] = -ap[i];
        double pap = or_dot(n, h, p, ap);
        double alpha = rz / pap;
        for (long i = 0; i < len; i++) {
            v[i] += alpha * p[i];
            r[i] -= alpha * ap[i];
        }
        for (long i = 0; i < len; i++) z[i] = inv_diag * r[i];
        double rz_new = or_dot(n, h, r, z);
        double beta = rz_new / rz;
        rz = rz_new;
        for (long i = 0; i < len; i++) p[i] = z[i] + beta * p[i];
        res = sqrt(or_dot(n, h, r, r)) / b_norm;
        it++;
    }
    if (residual) *residual = res;
    free(b);
    free(r);
    free(z);
    free(p);
    free(ap);
    return it;
}

/* ----------------------------------------------------------------- nuclei ---- */

/* qm/src/nuclear.rs:122-132 */
void or_deposit_charge(int n, double h, int nat, const double *pos, const int *z, double sigma, double *out) {
    long len = (long)n * n * n;
    double norm = 1.0 / pow(2.0 * M_PI * sigma * sigma, 1.5);
    double inv_2s2 = 1.0 / (2.0 * sigma * sigma);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < len; i++) {
        double p[3];
        grid_point(n, h, i, p);
        double acc = 0.0;
        for (int a = 0; a < nat; a++) {
            double dx = p[0] - pos[3 * a], dy = p[1] - pos[3 * a + 1], dz = p[2] - pos[3 * a + 2];
            acc += (double)z[a] * norm * exp(-(dx * dx + dy * dy + dz * dz) * inv_2s2);
        }
        out[i] = acc;
    }
}

/* qm/src/nuclear.rs:258-273 (ON_SITE = 1e-10) */
void or_nuclear_potential(int n, double h, int nat, const double *pos, const int *z, double sigma, double *out) {
    long len = (long)n * n * n;
    double s2 = sigma * sqrt(2.0);
    double on_site = sqrt(2.0 / M_PI) / sigma;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < len; i++) {
        double p[3];
        grid_point(n, h, i, p);
        double acc = 0.0;
        for (int a = 0; a < nat; a++) {
            double dx = p[0] - pos[3 * a], dy = p[1] - pos[3 * a + 1], dz = p[2] - pos[3 * a + 2];
            double d = sqrt(dx * dx + dy * dy + dz * dz);
            double zz = (double)z[a];
            acc += d < 1e-10 ? -zz * on_site : -zz * erf(d / s2) / d;
        }
        out[i] = acc;
    }
}

/* qm/src/nuclear.rs:277-289 (COINCIDENCE = 1e-8) */
int or_nuclear_repulsion(int nat, const double *pos, const int *z, double *e, int *pi, int *pj) {
    double acc = 0.0;
    for (int i = 0; i < nat; i++)
        for (int j = i + 1; j < nat; j++) {
            double dx = pos[3 * i] - pos[3 * j], dy = pos[3 * i + 1] - pos[3 * j + 1], dz = pos[3 * i + 2] - pos[3 * j + 2];
            double d = sqrt(dx * dx + dy * dy + dz * dz);
            if (d < 1e-8) {
                if (pi) *pi = i;
                if (pj) *pj = j;
                return OR_COINCIDENT;
            }
            acc += (double)z[i] * (double)z[j] / d;
        }
    *e = acc;
    return OR_OK;
}

/* -------------------------------------------------------------------- XC ---- */

/* PW92 Table I (p = 1): PARA, FERRO, MALPHA (qm/src/xc.rs:88-115) */
static const double PW[3][6] = {
    {0.031091, 0.21370, 7.5957, 3.5876, 1.6382, 0.49294},
    {0.015545, 0.20548, 14.1189, 6.1977, 3.3662, 0.62517},
    {0.016887, 0.11125, 10.357, 3.6231, 0.88026, 0.49671},
};

/* qm/src/xc.rs:120-128 */
void or_pw92_g(double rs, int which, double *g, double *dg) {
    const double *p = PW[which];
    double a = p[0], a1 = p[1], b1 = p[2], b2 = p[3], b3 = p[4], b4 = p[5];
    double srs = sqrt(rs);
    double q0 = -2.0 * a * (1.0 + a1 * rs);
    double q1 = 2.0 * a * (b1 * srs + b2 * rs + b3 * rs * srs + b4 * rs * rs);
    double q1p = a * (b1 / srs + 2.0 * b2 + 3.0 * b3 * srs + 4.0 * b4 * rs);
    double ln = log(1.0 + 1.0 / q1);
    *g = q0 * ln;
    *dg = -2.0 * a * a1 * ln - q0 * q1p / (q1 * q1 + q1);
}

/* qm/src/xc.rs:131-136 */
static void f_zeta(double z, double *f, double *fp) {
    double denom = pow(2.0, 4.0 / 3.0) - 2.0;
    *f = (pow(1.0 + z, 4.0 / 3.0) + pow(1.0 - z, 4.0 / 3.0) - 2.0) / denom;
    *fp = (4.0 / 3.0) * (pow(1.0 + z, 1.0 / 3.0) - pow(1.0 - z, 1.0 / 3.0)) / denom;
}

/* qm/src/xc.rs:145-165 (N_MIN = 1e-20, FPP0 = 1.709921) */
static void correlation(double nu, double nd, double *e, double *vu, double *vd) {
    const double FPP0 = 1.709921;
    double n = nu + nd;
    if (n < 1e-20) {
        *e = *vu = *vd = 0.0;
        return;
    }
    double rs = pow(3.0 / (4.0 * M_PI * n), 1.0 / 3.0);
    double z = (nu - nd) / n;
    if (z < -1.0) z = -1.0;
    if (z > 1.0) z = 1.0;
    double ec0, dec0, ec1, dec1, mac, dmac;
    or_pw92_g(rs, 0, &ec0, &dec0);
    or_pw92_g(rs, 1, &ec1, &dec1);
    or_pw92_g(rs, 2, &mac, &dmac);
    double ac = -mac, dac = -dmac;
    double f, fp;
    f_zeta(z, &f, &fp);
    double z3 = z * z * z, z4 = z3 * z;
    double ec = ec0 + ac * f / FPP0 * (1.0 - z4) + (ec1 - ec0) * f * z4;
    double dec_drs = dec0 + dac * f / FPP0 * (1.0 - z4) + (dec1 - dec0) * f * z4;
    double dec_dz = ac / FPP0 * (fp * (1.0 - z4) - 4.0 * z3 * f) + (ec1 - ec0) * (fp * z4 + 4.0 * z3 * f);
    double common = ec - rs / 3.0 * dec_drs;
    *e = n * ec;
    *vu = common - (z - 1.0) * dec_dz;
    *vd = common - (z + 1.0) * dec_dz;
}

/* qm/src/xc.rs:168-176 */
static void exchange(double nu, double nd, double *e, double *vu, double *vd) {
    double c = (3.0 / 4.0) * pow(3.0 / M_PI, 1.0 / 3.0) * pow(2.0, 1.0 / 3.0);
    *e = -c * (pow(nu, 4.0 / 3.0) + pow(nd, 4.0 / 3.0));
    *vu = -pow(6.0 * nu / M_PI, 1.0 / 3.0);
    *vd = -pow(6.0 * nd / M_PI, 1.0 / 3.0);
}

/* qm/src/xc.rs:179-187 */
void or_exc_point(double nu, double nd, double *e, double *vu, double *vd) {
    if (nu < 0.0) nu = 0.0;
    if (nd < 0.0) nd = 0.0;
    if (nu + nd < 1e-20) {
        *e = *vu = *vd = 0.0;
        return;
    }
    double ex, vxu, vxd, ec, vcu, vcd;
    exchange(nu, nd, &ex, &vxu, &vxd);
    correlation(nu, nd, &ec, &vcu, &vcd);
    *e = ex + ec;
    *vu = vxu + vcu;
    *vd = vxd + vcd;
}

/* qm/src/xc.rs:196-210 */
void or_lsda(long len, const double *nu, const double *nd, double *e, double *vu, double *vd) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < len; i++) or_exc_point(nu[i], nd[i], &e[i], &vu[i], &vd[i]);
}

/* --------------------------------------------------------- dense linear algebra ---- */

/* Cyclic Jacobi (qm/src/eigen.rs:175-211): rotations until the off-diagonal norm < 1e-15 ||A|| (100 sweeps cap);
 * then the pairs are sorted ascending (eigen.rs:345-347). */
void or_sym_eigen(int m, double *a, double *evals, double *evecs) {
    double *v = alloc0((long)m * m);
    for (int i = 0; i < m; i++) v[i * m + i] = 1.0;
    double norm = 0.0;
    for (long i = 0; i < (long)m * m; i++) norm += a[i] * a[i];
    norm = sqrt(norm);
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0.0;
        for (int i = 0; i < m; i++)
            for (int j = 0; j < m; j++)
                if (i != j) off += a[i * m + j] * a[i * m + j];
        off = sqrt(off);
        if (!(off > 1e-15 * norm)) break;
        for (int p = 0; p < m; p++)
            for (int q = p + 1; q < m; q++) {
                double apq = a[p * m + q];
                if (apq == 0.0) continue;
                double theta = (a[q * m + q] - a[p * m + p]) / (2.0 * apq);
                double t = (theta > 0 ? 1.0 : (theta < 0 ? -1.0 : 0.0)) / (fabs(theta) + sqrt(theta * theta + 1.0));
                if (theta == 0.0) t = 1.0;
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < m; k++) {
                    double akp = a[k * m + p], akq = a[k * m + q];
                    a[k * m + p] = c * akp - s * akq;
                    a[k * m + q] = s * akp + c * akq;
                }
                for (int k = 0; k < m; k++) {
                    double apk = a[p * m + k], aqk = a[q * m + k];
                    a[p * m + k] = c * apk - s * aqk;
                    a[q * m + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < m; k++) {
                    double vkp = v[k * m + p], vkq = v[k * m + q];
                    v[k * m + p] = c * vkp - s * vkq;
                    v[k * m + q] = s * vkp + c * vkq;
                }
            }
    }
    /* sort ascending (stable selection on the diagonal) */
    int *order = (int *)malloc(sizeof(int) * (size_t)m);
    for (int i = 0; i < m; i++) order[i] = i;
    for (int i = 1; i < m; i++) {
        int key = order[i], j = i - 1;
        while (j >= 0 && a[order[j] * m + order[j]] > a[key * m + key]) {
            order[j + 1] = order[j];
            j--;
        }
        order[j + 1] = key;
    }
    for (int c = 0; c < m; c++) {
        evals[c] = a[order[c] * m + order[c]];
        for (int r = 0; r < m; r++) evecs[r * m + c] = v[r * m + order[c]];
    }
    free(order);
    free(v);
}

/* LU with partial pivoting; returns 0 on a zero pivot (nalgebra's LU::solve -> None) */
static int lu_solve(int m, double *a, double *b) {
    for (int c = 0; c < m; c++) {
        int piv = c;
        double best = fabs(a[c * m + c]);
        for (int r = c + 1; r < m; r++)
            if (fabs(a[r * m + c]) > best) {
                best = fabs(a[r * m + c]);
                piv = r;
            }
        if (best == 0.0) return 0;
        if (piv != c) {
            for (int k = 0; k < m; k++) {
                double t = a[c * m + k];
                a[c * m + k] = a[piv * m + k];
                a[piv * m + k] = t;
            }
            double t = b[c];
            b[c] = b[piv];
            b[piv] = t;
        }
        for (int r = c + 1; r < m; r++) {
            double f = a[r * m + c] / a[c * m + c];
            for (int k = c; k < m; k++) a[r * m + k] -= f * a[c * m + k];
            b[r] -= f * b[c];
        }
    }
    for (int r = m - 1; r >= 0; r--) {
        double s = b[r];
        for (int k = r + 1; k < m; k++) s -= a[r * m + k] * b[k];
        b[r] = s / a[r * m + r];
    }
    return 1;
}

/* ----------------------------------------------------------------- LOBPCG ---- */

/* qm/src/eigen.rs:243-254: MGS in grid.dot; drops norm < 1e-12. s: cnt vectors (band-major), compacted in place.
 * Returns the kept count. */
static int orthonormalise(int n, double h, double *s, int cnt) {
    long len = (long)n * n * n;
    int kept = 0;
    for (int c = 0; c < cnt; c++) {
        double *v = s + (long)c * len;
        for (int q = 0; q < kept; q++) {
            double *qv = s + (long)q * len;
            double cf = or_dot(n, h, qv, v);
            for (long i = 0; i < len; i++) v[i] -= cf * qv[i];
        }
        double nrm = sqrt(or_dot(n, h, v, v));
        if (nrm > 1e-12) {
            double *dst = s + (long)kept * len;
            for (long i = 0; i < len; i++) dst[i] = v[i] / nrm;
            kept++;
        }
    }
    return kept;
}

/* LOBPCG iterations run by this process (the SCF reports its own share) */
static long or_lobpcg_iterations_total = 0;

int or_lobpcg(int n, double h, const double *v, int k, double *x, double *evals, double precond_lambda, double tol,
              int max_iter, double *max_residual, double *residuals) {
    return or_lobpcg_nl(n, h, v, NULL, k, x, evals, precond_lambda, tol, max_iter, max_residual, residuals);
}

/* qm/src/eigen.rs:290-369 (lobpcg_impl with head = k); H = PseudoHamiltonian when nl is given (nonlocal.rs:165-174),
 * the preconditioner stays the local Jacobi one (scf.rs:775-791) */
int or_lobpcg_nl(int n, double h, const double *v, const or_nl *nl, int k, double *x, double *evals,
                 double precond_lambda, double tol, int max_iter, double *max_residual, double *residuals) {
    long len = (long)n * n * n;
    double *xs = alloc0(k * len), *p = alloc0(k * len), *hx = alloc0(k * len), *w = alloc0(k * len);
    double *r = alloc0(len), *s = alloc0(3 * k * len), *hs = alloc0(3 * k * len);
    double *nx = alloc0(k * len), *np = alloc0(k * len);
    memcpy(xs, x, (size_t)(k * len) * sizeof(double));
    int kept = orthonormalise(n, h, xs, k);
    if (kept != k) { /* linearly dependent guess: reference asserts */
        free(xs); free(p); free(hx); free(w); free(r); free(s); free(hs); free(nx); free(np);
        return -1;
    }
    int have_p = 0, iterations = 0;
    double max_res = INFINITY;
    for (;;) {
        for (int i = 0; i < k; i++) {
            or_pseudo_hamiltonian(n, h, v, nl, xs + i * len, hx + i * len);
            evals[i] = or_dot(n, h, xs + i * len, hx + i * len);
        }
        max_res = 0.0;
        for (int i = 0; i < k; i++) {
            for (long j = 0; j < len; j++) r[j] = hx[i * len + j] - evals[i] * xs[i * len + j];
            double ri = sqrt(or_dot(n, h, r, r));
            if (ri > max_res || isnan(ri)) max_res = ri;
            if (residuals) residuals[i] = ri;
            if (isnan(precond_lambda)) memcpy(w + i * len, r, (size_t)len * sizeof(double));
            else or_jacobi_precond(n, h, v, precond_lambda, r, w + i * len);
        }
        if (max_res < tol || iterations >= max_iter) break;
        iterations++;
        int cnt = 0;
        memcpy(s, xs, (size_t)(k * len) * sizeof(double));
        cnt += k;
        memcpy(s + cnt * len, w, (size_t)(k * len) * sizeof(double));
        cnt += k;
        if (have_p) {
            memcpy(s + cnt * len, p, (size_t)(k * len) * sizeof(double));
            cnt += k;
        }
        int m = orthonormalise(n, h, s, cnt);
        for (int i = 0; i < m; i++) or_pseudo_hamiltonian(n, h, v, nl, s + i * len, hs + i * len);
        double *a = alloc0((long)m * m), *ev = alloc0(m), *vec = alloc0((long)m * m);
        for (int i = 0; i < m; i++)
            for (int j = 0; j <= i; j++) {
                double val = or_dot(n, h, s + i * len, hs + j * len);
                a[i * m + j] = val;
                a[j * m + i] = val;
            }
        or_sym_eigen(m, a, ev, vec);
        memset(nx, 0, (size_t)(k * len) * sizeof(double));
        memset(np, 0, (size_t)(k * len) * sizeof(double));
        for (int col = 0; col < k; col++)
            for (int si = 0; si < m; si++) {
                double coef = vec[si * m + col];
                if (coef == 0.0) continue;
                double *sv = s + si * len, *dx = nx + col * len, *dp = np + col * len;
                for (long j = 0; j < len; j++) dx[j] += coef * sv[j];
                if (si >= k)
                    for (long j = 0; j < len; j++) dp[j] += coef * sv[j];
            }
        memcpy(xs, nx, (size_t)(k * len) * sizeof(double));
        orthonormalise(n, h, xs, k);
        memcpy(p, np, (size_t)(k * len) * sizeof(double));
        have_p = 1;
        free(a);
        free(ev);
        free(vec);
    }
    memcpy(x, xs, (size_t)(k * len) * sizeof(double));
    if (max_residual) *max_residual = max_res;
    free(xs); free(p); free(hx); free(w); free(r); free(s); free(hs); free(nx); free(np);
    or_lobpcg_iterations_total += iterations;
    return iterations;
}

/* ------------------------------------------------------------ occupations ---- */

/* qm/src/occupations.rs:152-160 */
static double occupation(double x) {
    if (x >= 0.0) {
        double e = exp(-x);
        return e / (1.0 + e);
    }
    return 1.0 / (1.0 + exp(x));
}

static double fd_total(int m, const double *ev, double mu, double kt) {
    double sum = 0.0, comp = 0.0;
    for (int i = 0; i < m; i++) {
        double x = occupation((ev[i] - mu) / kt);
        double t = sum + x;
        if (fabs(sum) >= fabs(x)) comp += (sum - t) + x;
        else comp += (x - t) + sum;
        sum = t;
    }
    return sum + comp;
}

/* qm/src/occupations.rs:166-185 */
static double fd_bisect(int m, const double *ev, double nf, double kt, double lo, double hi, int ge) {
    for (int hv = 0; hv < 2048; hv++) {
        double mid = 0.5 * lo + 0.5 * hi;
        if (mid <= lo || mid >= hi) break;
        double d = fd_total(m, ev, mid, kt) - nf;
        int to_hi = ge ? (d >= 0.0) : (d > 0.0);
        if (to_hi) hi = mid;
        else lo = mid;
    }
    return ge ? hi : lo;
}

/* qm/src/occupations.rs:218-224 */
double or_fd_entropy(double x) {
    double a = fabs(x);
    if (isinf(a)) return 0.0;
    return log1p(exp(-a)) + a / (1.0 + exp(a));
}

/* qm/src/occupations.rs:189-213 */
int or_fermi_dirac(int m, const double *ev, int n, double kt, double *f, double *mu, double *entropy) {
    if (n == 0) {
        for (int i = 0; i < m; i++) f[i] = 0.0;
        *mu = NAN;
        *entropy = 0.0;
        return OR_OK;
    }
    if (m < n + 1) return OR_NUMERICAL;
    for (int i = 0; i < m; i++)
        if (!isfinite(ev[i])) return OR_NUMERICAL;
    if (!(isfinite(kt) && kt > 0.0)) return OR_NUMERICAL;
    double nf = (double)n, lo = INFINITY, hi = -INFINITY;
    for (int i = 0; i < m; i++) {
        if (ev[i] < lo) lo = ev[i];
        if (ev[i] > hi) hi = ev[i];
    }
    lo -= 50.0 * kt;
    hi += 50.0 * kt;
    if (!(fd_total(m, ev, lo, kt) < nf && fd_total(m, ev, hi, kt) > nf)) return OR_NUMERICAL;
    double mm = fd_bisect(m, ev, nf, kt, lo, hi, 1), mp = fd_bisect(m, ev, nf, kt, lo, hi, 0);
    double u = 0.5 * mm + 0.5 * mp;
    double *s = alloc0(m);
    for (int i = 0; i < m; i++) {
        f[i] = occupation((ev[i] - u) / kt);
        s[i] = or_fd_entropy((ev[i] - u) / kt);
    }
    *mu = u;
    *entropy = or_neumaier(s, m);
    free(s);
    return OR_OK;
}

/* ascending order of evals, stable (qm/src/scf.rs:902-906) */
static void ascending(int m, const double *ev, int *order) {
    for (int i = 0; i < m; i++) order[i] = i;
    for (int i = 1; i < m; i++) {
        int key = order[i], j = i - 1;
        while (j >= 0 && ev[order[j]] > ev[key]) {
            order[j + 1] = order[j];
            j--;
        }
        order[j + 1] = key;
    }
}

/* qm/src/occupations.rs:244-257 (WITNESS_GAP 0.1, WITNESS_TAIL_KT 36) */
static int witness_bound_reached(int m, const double *ev, const double *res, int n, double mu, double kt) {
    if (n == 0) return 1;
    if (m <= n) return 0;
    int *o = (int *)malloc(sizeof(int) * (size_t)m);
    ascending(m, ev, o);
    double eps_n = ev[o[n - 1]];
    int top = o[m - 1];
    double lower = ev[top] - res[top];
    free(o);
    return lower >= eps_n + 0.1 && lower - mu >= 36.0 * kt;
}

/* qm/src/occupations.rs:322-340 */
static int witness_keep(int m, const double *ev, const double *res, int n, double mu, double kt) {
    if (n == 0 || m <= n) return m;
    int *o = (int *)malloc(sizeof(int) * (size_t)m);
    ascending(m, ev, o);
    double eps_n = ev[o[n - 1]];
    int found = -1;
    for (int t = n; t < m; t++) {
        double lower = ev[o[t]] - res[o[t]];
        if (lower >= eps_n + 0.1 && lower - mu >= 36.0 * kt) {
            found = t;
            break;
        }
    }
    int keep = m;
    if (found >= 0) {
        keep = found + 1;
        if (found + 1 < m) {
            double lower = ev[o[found + 1]] - res[o[found + 1]];
            if (lower >= eps_n + 0.1 && lower - mu >= 36.0 * kt) keep = found + 2;
        }
    }
    free(o);
    return keep;
}

/* ------------------------------------------------------------------- SCF ---- */

/* qm/src/scf.rs:626-634: LCG stream, u = (seed >> 11) / 2^53, value (u + 0.5) * sum_I exp(-|r - R_I|^2 / 4) */
void or_initial_orbitals(int n, double h, int nat, const double *pos, int k, unsigned long long seed0, double *out) {
    long len = (long)n * n * n;
    unsigned long long seed = seed0;
    for (int b = 0; b < k; b++)
        for (long i = 0; i < len; i++) {
            seed = seed * 6364136223846793005ULL + 1442695040888963407ULL;
            double u = (double)(seed >> 11) / 9007199254740992.0;
            double p[3];
            grid_point(n, h, i, p);
            double env = 0.0;
            for (int a = 0; a < nat; a++) {
                double dx = p[0] - pos[3 * a], dy = p[1] - pos[3 * a + 1], dz = p[2] - pos[3 * a + 2];
                env += exp(-(dx * dx + dy * dy + dz * dz) / 4.0);
            }
            out[(long)b * len + i] = (u + 0.5) * env;
        }
}

typedef struct {
    double *rho_up, *rho_dn, *r_up, *r_dn;
} diis_entry;

/* qm/src/scf.rs:681-700 */
static int diis_coefficients(int n, double h, const diis_entry *hist, int m, double *c) {
    int mm = m + 1;
    double *b = alloc0((long)mm * mm), *rhs = alloc0(mm);
    for (int i = 0; i < m; i++)
        for (int j = 0; j <= i; j++) {
            double v = or_dot(n, h, hist[i].r_up, hist[j].r_up) + or_dot(n, h, hist[i].r_dn, hist[j].r_dn);
            b[i * mm + j] = v;
            b[j * mm + i] = v;
        }
    double scale = 0.0;
    for (int i = 0; i < m; i++)
        if (b[i * mm + i] > scale) scale = b[i * mm + i];
    int ok = 0;
    if (scale > 0.0) {
        for (int i = 0; i < m; i++) {
            for (int j = 0; j < m; j++) b[i * mm + j] /= scale;
            b[i * mm + i] += 1e-12;
        }
        for (int i = 0; i < m; i++) {
            b[i * mm + m] = 1.0;
            b[m * mm + i] = 1.0;
        }
        rhs[m] = 1.0;
        if (lu_solve(mm, b, rhs)) {
            ok = 1;
            for (int i = 0; i < m; i++) {
                c[i] = rhs[i];
                if (!isfinite(c[i])) ok = 0;
            }
        }
    }
    free(b);
    free(rhs);
    return ok;
}

static void free_entry(diis_entry *e) {
    free(e->rho_up);
    free(e->rho_dn);
    free(e->r_up);
    free(e->r_dn);
}

static double max_on_faces(int n, const double *f) {
    double m = 0.0;
    for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++)
                if (i == 0 || j == 0 || k == 0 || i == n - 1 || j == n - 1 || k == n - 1) {
                    double a = fabs(f[i + (long)n * (j + (long)n * k)]);
                    if (a > m) m = a;
                }
    return m;
}

/* one channel under Aufbau or Fermi–Dirac (qm/src/scf.rs:774-877, all columns tight). orb: *k vectors with room for
 * cap_total; grows *k. Returns status. */
static int solve_channel(int n, double h, const double *v_eff, const or_nl *nl, int nat, const double *pos,
                         double **orb, double *ev, double *res, int *k, int nel, const or_settings *s, int spin,
                         double *occ, double *mu, double *entropy, double *max_eig_res, int cap_total) {
    long len = (long)n * n * n;
    if (*k == 0) {
        *mu = NAN;
        *entropy = 0.0;
        return OR_OK;
    }
    for (;;) {
        double lambda = ev[0];
        double mr;
        int it = or_lobpcg_nl(n, h, v_eff, nl, *k, *orb, ev, lambda, s->eig_tol, 200, &mr, res);
        if (it < 0) return OR_NUMERICAL;
        if (mr > *max_eig_res || isnan(mr)) *max_eig_res = mr;
        if (s->occupations == 0) {
            for (int i = 0; i < *k; i++) occ[i] = 1.0;
            *mu = NAN;
            *entropy = 0.0;
            return OR_OK;
        }
        int st = or_fermi_dirac(*k, ev, nel, s->kt, occ, mu, entropy);
        if (st != OR_OK) return st;
        int extras = *k - nel;
        double guard_tol = s->eig_tol > 1e-3 ? s->eig_tol : 1e-3;
        int *o = (int *)malloc(sizeof(int) * (size_t)*k);
        ascending(*k, ev, o);
        double top_r = res[o[*k - 1]];
        free(o);
        int witness = witness_bound_reached(*k, ev, res, nel, *mu, s->kt);
        int cap = 32;
        if (witness && top_r > 1.5 * guard_tol && extras < cap && *k < cap_total) {
            or_initial_orbitals(n, h, nat, pos, 1, 1000ULL + 2ULL * (unsigned long long)*k + (unsigned long long)spin,
                                *orb + (long)*k * len);
            ev[*k] = ev[*k - 1];
            (*k)++;
            continue;
        }
        if (witness) {
            int keep = witness_keep(*k, ev, res, nel, *mu, s->kt);
            if (keep >= *k) return OR_OK;
            /* drop the levels above rank keep (in descending index order, as the reference) */
            int *ord = (int *)malloc(sizeof(int) * (size_t)*k);
            ascending(*k, ev, ord);
            int ndrop = *k - keep;
            int *drop = ord + keep;
            for (int a = 0; a < ndrop; a++)
                for (int b = a + 1; b < ndrop; b++)
                    if (drop[b] > drop[a]) {
                        int t = drop[a];
                        drop[a] = drop[b];
                        drop[b] = t;
                    }
            for (int a = 0; a < ndrop; a++) {
                int idx = drop[a];
                for (int c = idx; c < *k - 1; c++) {
                    memcpy(*orb + (long)c * len, *orb + (long)(c + 1) * len, (size_t)len * sizeof(double));
                    ev[c] = ev[c + 1];
                    res[c] = res[c + 1];
                }
                (*k)--;
            }
            free(ord);
            return or_fermi_dirac(*k, ev, nel, s->kt, occ, mu, entropy);
        }
        if (extras >= cap || *k >= cap_total) return OR_NUMERICAL;
        int target = 2 * extras;
        if (target < extras + 1) target = extras + 1;
        if (target > cap) target = cap;
        while (*k - nel < target && *k < cap_total) {
            or_initial_orbitals(n, h, nat, pos, 1, 1000ULL + 2ULL * (unsigned long long)*k + (unsigned long long)spin,
                                *orb + (long)*k * len);
            ev[*k] = ev[*k - 1];
            (*k)++;
        }
    }
}

int or_forces(int n, double h, const double *rho, int nat, const double *pos, const int *z, double sigma, double *out) {
    long len = (long)n * n * n;
    double *comp = alloc0(3 * len);
    double a = sigma * sqrt(2.0);
    for (int I = 0; I < nat; I++) {
        double zz = (double)z[I];
        memset(comp, 0, (size_t)(3 * len) * sizeof(double));
#pragma omp parallel for schedule(static)
        for (long p = 0; p < len; p++) {
            double r[3];
            grid_point(n, h, p, r);
            double dv[3] = {r[0] - pos[3 * I], r[1] - pos[3 * I + 1], r[2] - pos[3 * I + 2]};
            double d = sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
            if (d < 1e-8) continue;
            /* g'(d) = ((2/sqrt(pi)) u e^{-u^2} - erf(u)) / d^2, u = d / (sigma sqrt 2) (qm/src/forces.rs:176-181) */
            double u = d / a;
            double gp = ((2.0 / sqrt(M_PI)) * u * exp(-u * u) - erf(u)) / (d * d);
            double w = -zz * gp / d;
            for (int c = 0; c < 3; c++) comp[c * len + p] = rho[p] * w * dv[c];
        }
        for (int c = 0; c < 3; c++) out[3 * I + c] = or_integrate(n, h, comp + c * len);
        for (int J = 0; J < nat; J++) {
            if (J == I) continue;
            double dv[3] = {pos[3 * I] - pos[3 * J], pos[3 * I + 1] - pos[3 * J + 1], pos[3 * I + 2] - pos[3 * J + 2]};
            double d = sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
            if (d < 1e-8) {
                free(comp);
                return OR_COINCIDENT;
            }
            double sc = zz * (double)z[J] / (d * d * d);
            for (int c = 0; c < 3; c++) out[3 * I + c] += sc * dv[c];
        }
    }
    free(comp);
    return OR_OK;
}

/* qm/src/scf.rs:928-1098 */
int or_scf(int n, double h, int nat, const double *pos, const int *z, int n_up, int n_dn, const or_settings *s,
           or_scf_stats *st, int cap, double *evals_up, double *evals_dn, double *occ_up, double *occ_dn,
           double *rho_up_out, double *rho_dn_out, double *forces_out, double *orbitals_up_out,
           double *orbitals_dn_out) {
    return or_scf_ionic(n, h, nat, pos, z, n_up, n_dn, s, NULL, st, cap, evals_up, evals_dn, occ_up, occ_dn, rho_up_out,
                        rho_dn_out, forces_out, orbitals_up_out, orbitals_dn_out);
}

/* XC on (rho_up + pc/2, rho_dn + pc/2), plain LSDA without a partial core (scf.rs:757-769) */
static void xc_with_core(long len, const double *up, const double *dn, const double *pc, double *e, double *vu,
                         double *vd) {
    if (!pc) {
        or_lsda(len, up, dn, e, vu, vd);
        return;
    }
    double *u = alloc0(len), *d = alloc0(len);
    for (long i = 0; i < len; i++) {
        u[i] = up[i] + 0.5 * pc[i];
        d[i] = dn[i] + 0.5 * pc[i];
    }
    or_lsda(len, u, d, e, vu, vd);
    free(u);
    free(d);
}

/* qm/src/scf.rs:928-1098 with the ionic terms of scf.rs:720-755 (io NULL: all-electron, the same code path) */
int or_scf_ionic(int n, double h, int nat, const double *pos, const int *z, int n_up, int n_dn, const or_settings *s,
                 const or_ionic *io, or_scf_stats *st, int cap, double *evals_up, double *evals_dn, double *occ_up,
                 double *occ_dn, double *rho_up_out, double *rho_dn_out, double *forces_out, double *orbitals_up_out,
                 double *orbitals_dn_out) {
    double t_start = now_s();
    memset(st, 0, sizeof *st);
    const long lobpcg_before = or_lobpcg_iterations_total;
    long len = (long)n * n * n;
    double e_nn;
    int ci, cj;
    int rep = io ? or_ionic_repulsion(nat, pos, z, io, &e_nn, &ci, &cj) : or_nuclear_repulsion(nat, pos, z, &e_nn, &ci, &cj);
    if (rep != OR_OK) {
        st->status = OR_COINCIDENT;
        return OR_COINCIDENT;
    }
    double *v_nuc = alloc0(len);
    double *pc = NULL;
    or_nl *nl = NULL;
    if (io) {
        or_ionic_local_potential(n, h, nat, pos, z, io, v_nuc);
        pc = alloc0(len);
        if (!or_partial_core_field(n, h, nat, pos, z, io, pc)) {
            free(pc);
            pc = NULL;
        }
        nl = or_nl_build(n, h, nat, pos, z, io, 0);
        if (or_nl_spheres(nl) == 0) {
            or_nl_free(nl);
            nl = NULL;
        }
    } else {
        or_nuclear_potential(n, h, nat, pos, z, s->sigma, v_nuc);
    }
    double ptol = 1e-3 * s->tol < 1e-10 ? 1e-3 * s->tol : 1e-10; /* scf.rs:519 */
    int fd = s->occupations == 1;
    int cap_total = cap;
    int ku = n_up + (fd && n_up > 0 ? 1 : 0), kd = n_dn + (fd && n_dn > 0 ? 1 : 0);
    if (ku > cap_total || kd > cap_total) {
        free(v_nuc);
        free(pc);
        if (nl) or_nl_free(nl);
        st->status = OR_NUMERICAL;
        return OR_NUMERICAL;
    }
    double *orb_up = alloc0((long)cap_total * len), *orb_dn = alloc0((long)cap_total * len);
    double *ev_up = alloc0(cap_total), *ev_dn = alloc0(cap_total), *res_up = alloc0(cap_total), *res_dn = alloc0(cap_total);
    double *oc_up = alloc0(cap_total), *oc_dn = alloc0(cap_total);
    or_initial_orbitals(n, h, nat, pos, ku, 1, orb_up);
    or_initial_orbitals(n, h, nat, pos, kd, 2, orb_dn);
    for (int i = 0; i < cap_total; i++) ev_up[i] = ev_dn[i] = -1.0;
    /* cold guess: sigma = 1 blobs scaled to the electron count (scf.rs:951-964; nuclear.rs:204-208) */
    double *guess = alloc0(len);
    double z_tot = 0.0;
    if (io) {
        /* sigma = 1 blobs of each centre's charge q (Z_v pseudised, Z all-electron; nuclear.rs:221-226, 304-307) */
        const double norm = 1.0 / pow(2.0 * M_PI, 1.5);
        for (long i = 0; i < len; i++) {
            double p[3];
            grid_point(n, h, i, p);
            double acc = 0.0;
            for (int a = 0; a < nat; a++) {
                double dx = p[0] - pos[3 * a], dy = p[1] - pos[3 * a + 1], dz = p[2] - pos[3 * a + 2];
                acc += or_valence_charge(io, z[a]) * norm * exp(-(dx * dx + dy * dy + dz * dz) / 2.0);
            }
            guess[i] = acc;
        }
        for (int a = 0; a < nat; a++) z_tot += or_valence_charge(io, z[a]);
    } else {
        or_deposit_charge(n, h, nat, pos, z, 1.0, guess);
        for (int a = 0; a < nat; a++) z_tot += (double)z[a];
    }
    double total = (double)(n_up + n_dn);
    double scale = z_tot > 0.0 ? total / z_tot : 0.0;
    double fu = total > 0.0 ? (double)n_up / total : 0.0, fdn = total > 0.0 ? (double)n_dn / total : 0.0;
    double *rin_up = alloc0(len), *rin_dn = alloc0(len), *v_h = alloc0(len);
    for (long i = 0; i < len; i++) {
        rin_up[i] = guess[i] * scale * fu;
        rin_dn[i] = guess[i] * scale * fdn;
    }
    free(guess);
    double *rout_up = alloc0(len), *rout_dn = alloc0(len), *rho = alloc0(len);
    double *xe = alloc0(len), *xu = alloc0(len), *xd = alloc0(len), *vu = alloc0(len), *vd = alloc0(len);
    memcpy(rout_up, rin_up, (size_t)len * sizeof(double));
    memcpy(rout_dn, rin_dn, (size_t)len * sizeof(double));
    int hmax = s->diis_history > 1 ? s->diis_history : 1;
    diis_entry *hist = (diis_entry *)calloc((size_t)hmax + 1, sizeof(diis_entry));
    int hlen = 0;
    double *c = alloc0(hmax + 1);
    double delta_rho = INFINITY, mu_up = NAN, mu_dn = NAN, s_up = 0.0, s_dn = 0.0;
    double max_eig = 0.0;
    int iterations = 0, status = OR_OK, stalled = 0;
    /* stall watch (scf.rs:433-469) */
    double anchor = INFINITY, best = INFINITY;
    int anchor_at = 0;
    double level = 1e-3 * total;
    double t_poisson = 0.0, t_eigen = 0.0;
    while (iterations < s->max_iter) {
        iterations++;
        for (long i = 0; i < len; i++) rho[i] = rin_up[i] + rin_dn[i];
        double tp = now_s(), pres;
        int pit = or_poisson_solve(n, h, rho, v_h, ptol, 5000, &pres);
        t_poisson += now_s() - tp;
        if (pres > st->poisson_residual) st->poisson_residual = pres;
        if (pit > st->poisson_iterations) st->poisson_iterations = pit;
        st->cg_iterations += pit > 0 ? pit : 0;
        xc_with_core(len, rin_up, rin_dn, pc, xe, xu, xd);
        for (long i = 0; i < len; i++) {
            vu[i] = v_nuc[i] + v_h[i] + xu[i];
            vd[i] = v_nuc[i] + v_h[i] + xd[i];
        }
        double te = now_s();
        status = solve_channel(n, h, vu, nl, nat, pos, &orb_up, ev_up, res_up, &ku, n_up, s, 0, oc_up, &mu_up, &s_up,
                               &max_eig, cap_total);
        if (status == OR_OK)
            status = solve_channel(n, h, vd, nl, nat, pos, &orb_dn, ev_dn, res_dn, &kd, n_dn, s, 1, oc_dn, &mu_dn,
                                   &s_dn, &max_eig, cap_total);
        t_eigen += now_s() - te;
        if (status != OR_OK) break;
        memset(rout_up, 0, (size_t)len * sizeof(double));
        memset(rout_dn, 0, (size_t)len * sizeof(double));
        for (int b = 0; b < ku; b++) {
            double f = oc_up[b];
            if (f == 0.0) continue;
            const double *ps = orb_up + (long)b * len;
            for (long i = 0; i < len; i++) rout_up[i] += f * ps[i] * ps[i];
        }
        for (int b = 0; b < kd; b++) {
            double f = oc_dn[b];
            if (f == 0.0) continue;
            const double *ps = orb_dn + (long)b * len;
            for (long i = 0; i < len; i++) rout_dn[i] += f * ps[i] * ps[i];
        }
        for (long i = 0; i < len; i++) xe[i] = fabs(rout_up[i] - rin_up[i]) + fabs(rout_dn[i] - rin_dn[i]);
        delta_rho = or_integrate(n, h, xe);
        /* StallWatch::observe */
        if (delta_rho < best) best = delta_rho;
        if (isfinite(delta_rho) && (!isfinite(anchor) || delta_rho <= 0.5 * anchor)) {
            anchor = delta_rho;
            anchor_at = iterations;
        }
        int since = iterations - anchor_at;
        if (!isfinite(delta_rho)) break;
        if (delta_rho < s->tol) break;
        if (s->stall_window > 0 && since >= s->stall_window && best >= level) {
            stalled = 1;
            break;
        }
        /* DIIS push (scf.rs:1062-1070) */
        if (hlen == hmax + 1) {
            free_entry(&hist[0]);
            memmove(hist, hist + 1, sizeof(diis_entry) * (size_t)hmax);
            hlen--;
        }
        diis_entry e;
        e.rho_up = rin_up;
        e.rho_dn = rin_dn;
        e.r_up = alloc0(len);
        e.r_dn = alloc0(len);
        for (long i = 0; i < len; i++) {
            e.r_up[i] = rout_up[i] - rin_up[i];
            e.r_dn[i] = rout_dn[i] - rin_dn[i];
        }
        hist[hlen++] = e;
        if (hlen > hmax) {
            free_entry(&hist[0]);
            memmove(hist, hist + 1, sizeof(diis_entry) * (size_t)(hlen - 1));
            hlen--;
        }
        rin_up = alloc0(len);
        rin_dn = alloc0(len);
        /* diis_mix (scf.rs:704-718) */
        for (;;) {
            if (hlen == 1) {
                c[0] = 1.0;
                break;
            }
            if (diis_coefficients(n, h, hist, hlen, c)) break;
            free_entry(&hist[0]);
            memmove(hist, hist + 1, sizeof(diis_entry) * (size_t)(hlen - 1));
            hlen--;
        }
        for (long i = 0; i < len; i++) {
            double u = 0.0, d = 0.0;
            for (int q = 0; q < hlen; q++) {
                u += c[q] * (hist[q].rho_up[i] + s->mixing * hist[q].r_up[i]);
                d += c[q] * (hist[q].rho_dn[i] + s->mixing * hist[q].r_dn[i]);
            }
            rin_up[i] = u > 0.0 ? u : 0.0;
            rin_dn[i] = d > 0.0 ? d : 0.0;
        }
    }
    st->iterations = iterations;
    st->delta_rho = delta_rho;
    st->max_eigen_residual = max_eig;
    if (status == OR_OK) {
        if (stalled) status = OR_STALLED;
        else if (!(delta_rho < s->tol)) status = OR_NOT_CONVERGED;
    }
    if (status == OR_OK) {
        /* energy from the output density and orbitals (scf.rs:1048-1081) */
        for (long i = 0; i < len; i++) rho[i] = rout_up[i] + rout_dn[i];
        double pres;
        double tp = now_s();
        int pit = or_poisson_solve(n, h, rho, v_h, ptol, 5000, &pres);
        t_poisson += now_s() - tp;
        if (pres > st->poisson_residual) st->poisson_residual = pres;
        if (pit > st->poisson_iterations) st->poisson_iterations = pit;
        double kin = 0.0;
        for (int sp = 0; sp < 2; sp++) {
            int kk = sp == 0 ? ku : kd;
            double *orb = sp == 0 ? orb_up : orb_dn, *oc = sp == 0 ? oc_up : oc_dn;
            for (int b = 0; b < kk; b++) {
                double f = fd ? oc[b] : 1.0;
                if (fd && f == 0.0) continue;
                or_kinetic(n, h, orb + (long)b * len, xu);
                double t = or_dot(n, h, orb + (long)b * len, xu);
                kin += fd ? f * t : t;
            }
        }
        xc_with_core(len, rout_up, rout_dn, pc, xe, xu, xd);
        double external_ = or_dot(n, h, rho, v_nuc);
        if (nl) /* + E_NL, f-weighted under Fermi–Dirac (scf.rs:1066-1072) */
            external_ += or_nl_energy(nl, ku, orb_up, fd ? oc_up : NULL) + or_nl_energy(nl, kd, orb_dn, fd ? oc_dn : NULL);
        double hartree = 0.5 * or_dot(n, h, rho, v_h);
        double xc_e = or_integrate(n, h, xe);
        double internal = kin + external_ + hartree + xc_e + e_nn;
        double entropy = 0.0;
        if (fd) {
            double two[2] = {s_up, s_dn};
            entropy = or_neumaier(two, 2);
        }
        double ts = (fd ? s->kt : 0.0) * entropy;
        st->kinetic = kin;
        st->external_ = external_;
        st->hartree = hartree;
        st->xc = xc_e;
        st->nuclear = e_nn;
        st->internal = internal;
        st->ts = ts;
        st->total = internal - ts;
        st->entropy = entropy;
        st->mu[0] = mu_up;
        st->mu[1] = mu_dn;
        st->boundary_density = max_on_faces(n, rho);
        st->electrons[0] = or_integrate(n, h, rout_up);
        st->electrons[1] = or_integrate(n, h, rout_dn);
        st->k_up = ku;
        st->k_dn = kd;
        int kc_u = ku < cap ? ku : cap, kc_d = kd < cap ? kd : cap;
        for (int i = 0; i < kc_u; i++) {
            evals_up[i] = ev_up[i];
            occ_up[i] = fd ? oc_up[i] : 1.0;
        }
        for (int i = 0; i < kc_d; i++) {
            evals_dn[i] = ev_dn[i];
            occ_dn[i] = fd ? oc_dn[i] : 1.0;
        }
        if (rho_up_out) memcpy(rho_up_out, rout_up, (size_t)len * sizeof(double));
        if (rho_dn_out) memcpy(rho_dn_out, rout_dn, (size_t)len * sizeof(double));
        if (orbitals_up_out) memcpy(orbitals_up_out, orb_up, (size_t)(kc_u * len) * sizeof(double));
        if (orbitals_dn_out) memcpy(orbitals_dn_out, orb_dn, (size_t)(kc_d * len) * sizeof(double));
        if (forces_out && !io) status = or_forces(n, h, rho, nat, pos, z, s->sigma, forces_out);
        if (forces_out && io) {
            /* forces.rs:249-257: the per-component Neumaier sum of the five terms, in order */
            int kt = ku + kd;
            double *orbs = alloc0((long)kt * len), *oc = alloc0(kt), *terms = alloc0(15L * nat);
            memcpy(orbs, orb_up, (size_t)ku * len * sizeof(double));
            memcpy(orbs + (long)ku * len, orb_dn, (size_t)kd * len * sizeof(double));
            for (int i = 0; i < ku; i++) oc[i] = fd ? oc_up[i] : 1.0;
            for (int i = 0; i < kd; i++) oc[ku + i] = fd ? oc_dn[i] : 1.0;
            status = or_pseudo_force_terms(n, h, rout_up, rout_dn, kt, orbs, oc, nat, pos, z, io, terms);
            for (int I = 0; I < nat; I++)
                for (int a = 0; a < 3; a++) {
                    double t5[5];
                    for (int t = 0; t < 5; t++) t5[t] = terms[(t * nat + I) * 3 + a];
                    forces_out[3 * I + a] = or_neumaier(t5, 5);
                }
            free(orbs);
            free(oc);
            free(terms);
        }
    }
    st->status = status;
    st->seconds_poisson = t_poisson;
    st->seconds_eigen = t_eigen;
    st->seconds_total = now_s() - t_start;
    st->lobpcg_iterations = (int)(or_lobpcg_iterations_total - lobpcg_before);
    for (int q = 0; q < hlen; q++) free_entry(&hist[q]);
    free(hist);
    free(c);
    free(v_nuc); free(orb_up); free(orb_dn); free(ev_up); free(ev_dn); free(res_up); free(res_dn);
    free(oc_up); free(oc_dn); free(rin_up); free(rin_dn); free(v_h); free(rout_up); free(rout_dn); free(rho);
    free(xe); free(xu); free(xd); free(vu); free(vd);
    free(pc);
    if (nl) or_nl_free(nl);
    return status;
}

/* ---------------------------------------------------------- pseudised ions ---- */

/* qm/src/radial/filter.rs:302-318: 4-point Lagrange on a uniform table of an even function (index -1 mirrors 1) */
static void uniform_lagrange(const double *f, int len, double du, double d, double *val, double *der) {
    double sv = d / du;
    long i0 = (long)floor(sv) - 1;
    if (i0 + 3 >= (long)len) {
        *val = 0.0;
        *der = 0.0;
        return;
    }
    double t = sv - (double)i0;
    double a = f[labs(i0)], b = f[labs(i0 + 1)], c = f[labs(i0 + 2)], e = f[labs(i0 + 3)];
    double l0 = -(t - 1.0) * (t - 2.0) * (t - 3.0) / 6.0, l1 = t * (t - 2.0) * (t - 3.0) / 2.0;
    double l2 = -t * (t - 1.0) * (t - 3.0) / 2.0, l3 = t * (t - 1.0) * (t - 2.0) / 6.0;
    double d0 = -((t - 2.0) * (t - 3.0) + (t - 1.0) * (t - 3.0) + (t - 1.0) * (t - 2.0)) / 6.0;
    double d1 = ((t - 2.0) * (t - 3.0) + t * (t - 3.0) + t * (t - 2.0)) / 2.0;
    double d2 = -((t - 1.0) * (t - 3.0) + t * (t - 3.0) + t * (t - 1.0)) / 2.0;
    double d3 = ((t - 1.0) * (t - 2.0) + t * (t - 2.0) + t * (t - 1.0)) / 6.0;
    *val = a * l0 + b * l1 + c * l2 + e * l3;
    *der = (a * d0 + b * d1 + c * d2 + e * d3) / du;
}

/* filter.rs:355 (read): 0 beyond the table's radius */
void or_table_read(const double *f, int len, double du, double radius, double d, double *val, double *der) {
    if (d > radius || !f) {
        *val = 0.0;
        *der = 0.0;
        return;
    }
    uniform_lagrange(f, len, du, d, val, der);
}

static const or_species *find_species(const or_ionic *io, int z) {
    if (!io) return NULL;
    for (int i = 0; i < io->n_species; i++)
        if (io->species[i].z == z) return &io->species[i];
    return NULL;
}

/* ionic.rs:179-182 */
double or_valence_charge(const or_ionic *io, int z) {
    const or_species *sp = find_species(io, z);
    return sp ? sp->z_v : (double)z;
}

static double dist2(const double a[3], const double b[3]) {
    return (a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]);
}

/* nuclear.rs:322-337 */
long or_points_within(int n, double h, const double c[3], double radius, long *out) {
    double o = origin_of(n, h);
    long lo[3], hi[3];
    for (int a = 0; a < 3; a++) {
        double l = floor((c[a] - radius - o) / h);
        double u = ceil((c[a] + radius - o) / h);
        if (l < 0.0) l = 0.0;
        if (u > (double)(n - 1)) u = (double)(n - 1);
        if (!(u >= l)) return 0;
        lo[a] = (long)l;
        hi[a] = (long)u;
    }
    long cnt = 0;
    for (long k = lo[2]; k <= hi[2]; k++)
        for (long j = lo[1]; j <= hi[1]; j++)
            for (long i = lo[0]; i <= hi[0]; i++) {
                long p = i + (long)n * (j + (long)n * k);
                double r[3];
                grid_point(n, h, p, r);
                if (dist2(r, c) <= radius * radius) {
                    if (out) out[cnt] = p;
                    cnt++;
                }
            }
    return cnt;
}

static long *points_alloc(int n, double h, const double c[3], double radius, long *cnt) {
    *cnt = or_points_within(n, h, c, radius, NULL);
    long *pts = (long *)malloc(sizeof(long) * (size_t)(*cnt > 0 ? *cnt : 1));
    or_points_within(n, h, c, radius, pts);
    return pts;
}

/* nuclear.rs:290-302 (smeared_potential over centres (q, sigma)) + 342-350 (the V_SR tables) */
void or_ionic_local_potential(int n, double h, int nat, const double *pos, const int *z, const or_ionic *io,
                              double *out) {
    long len = (long)n * n * n;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < len; i++) {
        double p[3];
        grid_point(n, h, i, p);
        double acc = 0.0;
        for (int a = 0; a < nat; a++) {
            const or_species *sp = find_species(io, z[a]);
            double q = sp ? sp->z_v : (double)z[a], sg = sp ? sp->sigma_lr : io->sigma_ae;
            double d = sqrt(dist2(p, pos + 3 * a));
            acc += d < 1e-10 ? -q * sqrt(2.0 / M_PI) / sg : -q * erf(d / (sg * sqrt(2.0))) / d;
        }
        out[i] = acc;
    }
    for (int a = 0; a < nat; a++) {
        const or_species *sp = find_species(io, z[a]);
        if (!sp) continue;
        long cnt;
        long *pts = points_alloc(n, h, pos + 3 * a, sp->r_sr, &cnt);
        for (long q = 0; q < cnt; q++) {
            double r[3], v, dv;
            grid_point(n, h, pts[q], r);
            or_table_read(sp->v_sr_u, sp->n_v_sr, sp->du, sp->r_sr, sqrt(dist2(r, pos + 3 * a)), &v, &dv);
            out[pts[q]] += v;
        }
        free(pts);
    }
}

/* nuclear.rs:354-368 */
int or_partial_core_field(int n, double h, int nat, const double *pos, const int *z, const or_ionic *io,
                          double *out) {
    long len = (long)n * n * n;
    memset(out, 0, (size_t)len * sizeof(double));
    int any = 0;
    for (int a = 0; a < nat; a++) {
        const or_species *sp = find_species(io, z[a]);
        if (!sp || sp->pc_radius <= 0.0) continue;
        any = 1;
        long cnt;
        long *pts = points_alloc(n, h, pos + 3 * a, sp->pc_radius, &cnt);
        for (long q = 0; q < cnt; q++) {
            double r[3], v, dv;
            grid_point(n, h, pts[q], r);
            or_table_read(sp->rho_pc_u, sp->n_pc, sp->du, sp->pc_radius, sqrt(dist2(r, pos + 3 * a)), &v, &dv);
            out[pts[q]] += v;
        }
        free(pts);
    }
    return any;
}

/* nuclear.rs:371-382 */
int or_ionic_repulsion(int nat, const double *pos, const int *z, const or_ionic *io, double *e, int *pi, int *pj) {
    double acc = 0.0;
    for (int i = 0; i < nat; i++)
        for (int j = i + 1; j < nat; j++) {
            double d = sqrt(dist2(pos + 3 * i, pos + 3 * j));
            if (d < 1e-8) {
                *pi = i;
                *pj = j;
                return OR_COINCIDENT;
            }
            acc += or_valence_charge(io, z[i]) * or_valence_charge(io, z[j]) / d;
        }
    *e = acc;
    return OR_OK;
}

/* nonlocal.rs:63-73 */
void or_solid_harmonic_gradients(const double r[3], double out[9][3]) {
    double x = r[0], y = r[1], zz = r[2], s3 = sqrt(3.0);
    double g[9][3] = {{0.0, 0.0, 0.0},         {0.0, 1.0, 0.0},           {0.0, 0.0, 1.0},
                      {1.0, 0.0, 0.0},         {s3 * y, s3 * x, 0.0},     {0.0, s3 * zz, s3 * y},
                      {-x, -y, 2.0 * zz},      {s3 * zz, 0.0, s3 * x},    {s3 * x, -s3 * y, 0.0}};
    memcpy(out, g, sizeof g);
}

typedef struct {
    int ion, l;
    long npts;
    long *points;
    double *beta; /* (2l+1) x npts */
    double *grad; /* (2l+1) x npts x 3, or NULL */
    double e_kb;
} or_sphere;

struct or_nl {
    int ns, rows;
    or_sphere *s;
    double dv;
    long len; /* n^3: the length of one orbital */
};

/* nonlocal.rs:94-128 */
or_nl *or_nl_build(int n, double h, int nat, const double *pos, const int *z, const or_ionic *io, int with_grad) {
    or_nl *nl = (or_nl *)calloc(1, sizeof(or_nl));
    nl->dv = h * h * h;
    nl->len = (long)n * n * n;
    int cap = 0;
    for (int a = 0; a < nat; a++) {
        const or_species *sp = find_species(io, z[a]);
        if (sp) cap += sp->n_proj;
    }
    nl->s = (or_sphere *)calloc((size_t)(cap > 0 ? cap : 1), sizeof(or_sphere));
    for (int a = 0; a < nat; a++) {
        const or_species *sp = find_species(io, z[a]);
        if (!sp) continue;
        for (int k = 0; k < sp->n_proj; k++) {
            const or_projector *pr = &sp->proj[k];
            or_sphere *sph = &nl->s[nl->ns++];
            sph->ion = a;
            sph->l = pr->l;
            sph->e_kb = pr->e_kb;
            sph->points = points_alloc(n, h, pos + 3 * a, pr->radius, &sph->npts);
            int nm = 2 * pr->l + 1;
            nl->rows += nm;
            sph->beta = alloc0((long)nm * sph->npts);
            sph->grad = with_grad ? alloc0(3L * nm * sph->npts) : NULL;
            double norm = sqrt((double)(2 * pr->l + 1) / (4.0 * M_PI));
            for (long q = 0; q < sph->npts; q++) {
                double r[3], d[3], sh[OR_N_MOMENTS];
                grid_point(n, h, sph->points[q], r);
                for (int c = 0; c < 3; c++) d[c] = r[c] - pos[3 * a + c];
                double dd = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                or_solid_harmonics(d, sh);
                double tv, dt;
                or_table_read(pr->t_u, pr->n_t, sp->du, pr->radius, dd, &tv, &dt);
                double tn = tv * norm;
                for (int m = -pr->l; m <= pr->l; m++)
                    sph->beta[(long)(m + pr->l) * sph->npts + q] = tn * sh[lm_index(pr->l, m)];
                if (with_grad) {
                    double gr[9][3];
                    or_solid_harmonic_gradients(d, gr);
                    for (int m = -pr->l; m <= pr->l; m++) {
                        int lm = lm_index(pr->l, m);
                        double radial = dd > 1e-12 ? dt * norm * sh[lm] / dd : 0.0;
                        for (int c = 0; c < 3; c++)
                            sph->grad[((long)(m + pr->l) * sph->npts + q) * 3 + c] = radial * d[c] + tn * gr[lm][c];
                    }
                }
            }
        }
    }
    return nl;
}

void or_nl_free(or_nl *nl) {
    if (!nl) return;
    for (int i = 0; i < nl->ns; i++) {
        free(nl->s[i].points);
        free(nl->s[i].beta);
        free(nl->s[i].grad);
    }
    free(nl->s);
    free(nl);
}

int or_nl_rows(const or_nl *nl) { return nl ? nl->rows : 0; }
int or_nl_spheres(const or_nl *nl) { return nl ? nl->ns : 0; }
long or_nl_sphere_points(const or_nl *nl, int s) { return nl->s[s].npts; }

/* nonlocal.rs:133-135: c[sphere][m + l] = h^3 sum_p beta_lm(r_p) psi(r_p) (Neumaier) */
void or_nl_project(const or_nl *nl, const double *psi, double *c) {
    int row = 0;
    for (int i = 0; i < nl->ns; i++) {
        const or_sphere *sph = &nl->s[i];
        double *tmp = alloc0(sph->npts);
        for (int m = 0; m < 2 * sph->l + 1; m++) {
            const double *b = sph->beta + (long)m * sph->npts;
            for (long q = 0; q < sph->npts; q++) tmp[q] = b[q] * psi[sph->points[q]];
            c[row++] = or_neumaier(tmp, sph->npts) * nl->dv;
        }
        free(tmp);
    }
}

/* nonlocal.rs:138-145 */
void or_nl_apply_add(const or_nl *nl, const double *psi, double *out) {
    double *c = alloc0(nl->rows);
    or_nl_project(nl, psi, c);
    int row = 0;
    for (int i = 0; i < nl->ns; i++) {
        const or_sphere *sph = &nl->s[i];
        for (int m = 0; m < 2 * sph->l + 1; m++) {
            double f = sph->e_kb * c[row++];
            const double *b = sph->beta + (long)m * sph->npts;
            for (long q = 0; q < sph->npts; q++) out[sph->points[q]] += f * b[q];
        }
    }
    free(c);
}

/* nonlocal.rs:148-172 (energy / energy_weighted): one Neumaier sum over every (orbital, sphere, m) term */
double or_nl_energy(const or_nl *nl, int k, const double *orbitals, const double *f) {
    if (!nl || nl->rows == 0 || k == 0) return 0.0;
    const long len = nl->len;
    double *terms = alloc0((long)k * nl->rows), *c = alloc0(nl->rows);
    long nt = 0;
    for (int b = 0; b < k; b++) {
        double fi = f ? f[b] : 1.0;
        if (f && fi == 0.0) continue;
        or_nl_project(nl, orbitals + (long)b * len, c);
        int row = 0;
        for (int i = 0; i < nl->ns; i++)
            for (int m = 0; m < 2 * nl->s[i].l + 1; m++, row++)
                terms[nt++] = f ? fi * nl->s[i].e_kb * c[row] * c[row] : nl->s[i].e_kb * c[row] * c[row];
    }
    double e = or_neumaier(terms, nt);
    free(terms);
    free(c);
    return e;
}

/* nonlocal.rs:165-174 */
void or_pseudo_hamiltonian(int n, double h, const double *v, const or_nl *nl, const double *x, double *out) {
    or_hamiltonian(n, h, v, x, out);
    if (nl && nl->ns > 0) or_nl_apply_add(nl, x, out);
}

/* forces.rs:232-246: sum_p w_p V'(d_p)(r_p - R)/d_p h^3 over the points (every point when pts is NULL), on-site
 * points (d < 1e-8) skipped; kind 0: -q g'(d, sigma) (the smeared erf), 1: a table's interpolant derivative */
typedef struct {
    int kind;
    double q, sigma;
    const double *tab;
    int len;
    double du, radius;
} radial_deriv;

static double radial_dvdd(const radial_deriv *rd, double d) {
    if (rd->kind == 0) {
        /* FilteredTable::dv_lr (filter.rs:347-352) and -Z g'(d) (forces.rs:177-181): one formula */
        if (d < 1e-10) return 0.0;
        double x = d / (rd->sigma * sqrt(2.0));
        return -rd->q * ((2.0 / sqrt(M_PI)) * x * exp(-x * x) - erf(x)) / (d * d);
    }
    double v, dv;
    or_table_read(rd->tab, rd->len, rd->du, rd->radius, d, &v, &dv);
    return dv;
}

static void radial_force(int n, double h, const long *pts, long npts, const double c[3], const double *w,
                         const radial_deriv *rd, double out[3]) {
    long len = (long)n * n * n;
    long cnt = pts ? npts : len;
    double *comp = alloc0(3 * cnt);
    long m = 0;
    for (long q = 0; q < cnt; q++) {
        long p = pts ? pts[q] : q;
        double r[3];
        grid_point(n, h, p, r);
        double d[3] = {r[0] - c[0], r[1] - c[1], r[2] - c[2]};
        double dd = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        if (dd < 1e-8) continue;
        double s = w[p] * radial_dvdd(rd, dd) / dd;
        for (int a = 0; a < 3; a++) comp[a * cnt + m] = s * d[a];
        m++;
    }
    for (int a = 0; a < 3; a++) out[a] = or_neumaier(comp + a * cnt, m) * h * h * h;
    free(comp);
}

/* forces.rs:259-350 */
int or_pseudo_force_terms(int n, double h, const double *rho_up, const double *rho_dn, int k, const double *orbitals,
                          const double *occ, int nat, const double *pos, const int *z, const or_ionic *io,
                          double *terms) {
    long len = (long)n * n * n;
    memset(terms, 0, sizeof(double) * 15 * (size_t)nat);
    double *t_ii = terms, *t_lr = terms + 3 * nat, *t_sr = terms + 6 * nat, *t_nl = terms + 9 * nat,
           *t_pc = terms + 12 * nat;
    /* ion-ion (forces.rs:335-350) */
    for (int i = 0; i < nat; i++)
        for (int j = 0; j < nat; j++) {
            if (i == j) continue;
            double dv[3] = {pos[3 * i] - pos[3 * j], pos[3 * i + 1] - pos[3 * j + 1], pos[3 * i + 2] - pos[3 * j + 2]};
            double d = sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
            if (d < 1e-8) return OR_COINCIDENT;
            double sc = or_valence_charge(io, z[i]) * or_valence_charge(io, z[j]) / (d * d * d);
            for (int a = 0; a < 3; a++) t_ii[3 * i + a] += sc * dv[a];
        }
    double *rho = alloc0(len);
    for (long p = 0; p < len; p++) rho[p] = rho_up[p] + rho_dn[p];
    /* local LR and SR (forces.rs:281-297) */
    for (int a = 0; a < nat; a++) {
        const or_species *sp = find_species(io, z[a]);
        radial_deriv lr = {0, sp ? sp->z_v : (double)z[a], sp ? sp->sigma_lr : io->sigma_ae, NULL, 0, 0.0, 0.0};
        radial_force(n, h, NULL, 0, pos + 3 * a, rho, &lr, t_lr + 3 * a);
        if (sp) {
            long cnt;
            long *pts = points_alloc(n, h, pos + 3 * a, sp->r_sr, &cnt);
            radial_deriv sr = {1, 0.0, 0.0, sp->v_sr_u, sp->n_v_sr, sp->du, sp->r_sr};
            radial_force(n, h, pts, cnt, pos + 3 * a, rho, &sr, t_sr + 3 * a);
            free(pts);
        }
    }
    /* nonlocal (forces.rs:299-318) */
    or_nl *nl = or_nl_build(n, h, nat, pos, z, io, 1);
    long maxn = 0;
    for (int i = 0; i < nl->ns; i++)
        if (nl->s[i].npts > maxn) maxn = nl->s[i].npts;
    double **per = (double **)calloc((size_t)nat * 3, sizeof(double *));
    long *per_n = (long *)calloc((size_t)nat * 3, sizeof(long));
    long per_cap = (long)k * 9 * nl->ns + 1;
    for (int i = 0; i < nat * 3; i++) per[i] = alloc0(per_cap);
    double *tmp = alloc0(maxn > 0 ? maxn : 1);
    for (int i = 0; i < nl->ns; i++) {
        const or_sphere *sph = &nl->s[i];
        for (int b = 0; b < k; b++) {
            double f = occ[b];
            if (f == 0.0) continue;
            const double *psi = orbitals + (long)b * len;
            for (int m = 0; m < 2 * sph->l + 1; m++) {
                const double *bv = sph->beta + (long)m * sph->npts;
                for (long q = 0; q < sph->npts; q++) tmp[q] = bv[q] * psi[sph->points[q]];
                double c = or_neumaier(tmp, sph->npts) * nl->dv;
                for (int a = 0; a < 3; a++) {
                    for (long q = 0; q < sph->npts; q++)
                        tmp[q] = sph->grad[((long)m * sph->npts + q) * 3 + a] * psi[sph->points[q]];
                    double gpsi = or_neumaier(tmp, sph->npts) * nl->dv;
                    int slot = sph->ion * 3 + a;
                    per[slot][per_n[slot]++] = 2.0 * f * sph->e_kb * c * gpsi;
                }
            }
        }
    }
    for (int i = 0; i < nat * 3; i++) {
        t_nl[i] = or_neumaier(per[i], per_n[i]);
        free(per[i]);
    }
    free(per);
    free(per_n);
    free(tmp);
    or_nl_free(nl);
    /* partial core (forces.rs:320-333) */
    double *pc = alloc0(len);
    if (or_partial_core_field(n, h, nat, pos, z, io, pc)) {
        double *e = alloc0(len), *vu = alloc0(len), *vd = alloc0(len);
        xc_with_core(len, rho_up, rho_dn, pc, e, vu, vd);
        for (long p = 0; p < len; p++) e[p] = 0.5 * (vu[p] + vd[p]);
        for (int a = 0; a < nat; a++) {
            const or_species *sp = find_species(io, z[a]);
            if (!sp || sp->pc_radius <= 0.0) continue;
            long cnt;
            long *pts = points_alloc(n, h, pos + 3 * a, sp->pc_radius, &cnt);
            radial_deriv pd = {1, 0.0, 0.0, sp->rho_pc_u, sp->n_pc, sp->du, sp->pc_radius};
            radial_force(n, h, pts, cnt, pos + 3 * a, e, &pd, t_pc + 3 * a);
            free(pts);
        }
        free(e);
        free(vu);
        free(vd);
    }
    free(pc);
    free(rho);
    return OR_OK;
}

/* ---------------------------------------------------------- trilinear sampling ---- */

/* qm/src/grid.rs:54-86: the cell and fractions, neighbours beyond an upper face read 0, blended along x, y, z */
double or_trilinear(int n, double h, const double *f, const double p[3]) {
    double o = origin_of(n, h), t[3];
    long c[3];
    for (int a = 0; a < 3; a++) {
        double x = (p[a] - o) / h;
        if (!(x >= 0.0 && x <= (double)(n - 1))) return 0.0;
        double i0 = floor(x);
        c[a] = (long)i0;
        t[a] = x - i0;
    }
    double v[2][2][2];
    for (int dk = 0; dk < 2; dk++)
        for (int dj = 0; dj < 2; dj++)
            for (int di = 0; di < 2; di++) {
                long a = c[0] + di, b = c[1] + dj, cc = c[2] + dk;
                v[dk][dj][di] = (a < n && b < n && cc < n) ? f[a + (long)n * (b + (long)n * cc)] : 0.0;
            }
    double y[2];
    for (int dk = 0; dk < 2; dk++) {
        double x0 = v[dk][0][0] * (1.0 - t[0]) + v[dk][0][1] * t[0];
        double x1 = v[dk][1][0] * (1.0 - t[0]) + v[dk][1][1] * t[0];
        y[dk] = x0 * (1.0 - t[1]) + x1 * t[1];
    }
    return y[0] * (1.0 - t[2]) + y[1] * t[2];
}

/* ---------------------------------------------------------- multigrid model ---- */

/* qm-gpu/src/multigrid.rs:210-223 */
void or_mg_apply_a2(int n, double h, double d_low, double d_high, const double *x, double *out) {
    for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) {
                long id = i + (long)n * (j + (long)n * k);
                double c = x[id];
                int idx[3] = {i, j, k};
                double s = 0.0;
                for (int a = 0; a < 3; a++)
                    for (int sg = -1; sg <= 1; sg += 2) {
                        int q[3] = {idx[0], idx[1], idx[2]};
                        q[a] += sg;
                        if (q[0] < 0 || q[1] < 0 || q[2] < 0) s += (1.0 - 1.0 / d_low) * c;
                        else if (q[0] >= n || q[1] >= n || q[2] >= n) s += (1.0 - 1.0 / d_high) * c;
                        else s += x[q[0] + (long)n * (q[1] + (long)n * q[2])];
                    }
                out[id] = -(s - 6.0 * c) / (h * h);
            }
}

/* qm-gpu/src/multigrid.rs:226-240 */
void or_mg_restrict(int n, const double *f, double *out) {
    int nc = n / 2;
    for (int k = 0; k < nc; k++)
        for (int j = 0; j < nc; j++)
            for (int i = 0; i < nc; i++) {
                double s = 0.0;
                for (int dz = -1; dz <= 1; dz++)
                    for (int dy = -1; dy <= 1; dy++)
                        for (int dx = -1; dx <= 1; dx++) {
                            int a = 2 * i + 1 + dx, b = 2 * j + 1 + dy, c = 2 * k + 1 + dz;
                            double w = (dx == 0 ? 0.5 : 0.25) * (dy == 0 ? 0.5 : 0.25) * (dz == 0 ? 0.5 : 0.25);
                            if (a < 0 || b < 0 || c < 0 || a >= n || b >= n || c >= n) continue;
                            s += w * f[a + (long)n * (b + (long)n * c)];
                        }
                out[i + (long)nc * (j + (long)nc * k)] = s;
            }
}

/* qm-gpu/src/multigrid.rs:244-258 */
void or_mg_prolong(int n, const double *cf, double *out) {
    int nc = n / 2;
    for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) {
                int t[3] = {i, j, k};
                int ci[3][2];
                double wi[3][2];
                int cnt[3];
                for (int a = 0; a < 3; a++) {
                    int cc = t[a] / 2;
                    if (t[a] % 2 == 1) {
                        ci[a][0] = cc;
                        wi[a][0] = 1.0;
                        cnt[a] = 1;
                    } else {
                        ci[a][0] = cc - 1;
                        wi[a][0] = 0.5;
                        ci[a][1] = cc;
                        wi[a][1] = 0.5;
                        cnt[a] = 2;
                    }
                }
                double s = 0.0;
                for (int a = 0; a < cnt[0]; a++)
                    for (int b = 0; b < cnt[1]; b++)
                        for (int c = 0; c < cnt[2]; c++) {
                            int x = ci[0][a], y = ci[1][b], zz = ci[2][c];
                            if (x < 0 || y < 0 || zz < 0 || x >= nc || y >= nc || zz >= nc) continue;
                            s += wi[0][a] * wi[1][b] * wi[2][c] * cf[x + (long)nc * (y + (long)nc * zz)];
                        }
                out[i + (long)n * (j + (long)n * k)] = s;
            }
}
