// gfx950 kernels of the Poisson solve: multipole moments and the boundary right-hand side (f64), the multigrid
// V-cycle level operators (7-point A2 with boundary-distance ghost factors, full weighting, trilinear prolongation)
// and the device-side CG updates whose scalars never leave the GPU.
#include "lira_internal.hpp"

#include <cmath>

namespace lira {

constexpr int LMAX = 6;
constexpr int NMOM = 49;

// Real regular solid harmonics R_lm (Racah normalisation), qm/src/poisson.rs:97-130: complex C_l^m by the diagonal
// ladder and the three-term recursion in l at fixed m; R_l0 = Re C, R_lm = sqrt2 Re C, R_l,-m = sqrt2 Im C.
__device__ inline void solid_harmonics(double x, double y, double z, double out[NMOM]) {
    const double r2 = x * x + y * y + z * z;
    double cr[LMAX + 1][LMAX + 1], ci[LMAX + 1][LMAX + 1];
#pragma unroll
    for (int l = 0; l <= LMAX; l++)
#pragma unroll
        for (int m = 0; m <= LMAX; m++) cr[l][m] = ci[l][m] = 0.0;
    cr[0][0] = 1.0;
#pragma unroll
    for (int l = 1; l <= LMAX; l++) {
        const double lf = (double)l;
        const double f = sqrt((2.0 * lf - 1.0) / (2.0 * lf));
        const double pr = cr[l - 1][l - 1], pi = ci[l - 1][l - 1];
        cr[l][l] = f * (x * pr - y * pi);
        ci[l][l] = f * (x * pi + y * pr);
#pragma unroll
        for (int m = 0; m < l; m++) {
            const double mf = (double)m;
            const double a = (2.0 * lf - 1.0) * z;
            const double r1 = cr[l - 1][m], i1 = ci[l - 1][m];
            double r0 = 0.0, i0 = 0.0;
            if (l >= 2 && m <= l - 2) {
                r0 = cr[l - 2][m];
                i0 = ci[l - 2][m];
            }
            const double tt = (lf - 1.0) * (lf - 1.0) - mf * mf;
            const double b = sqrt(tt > 0.0 ? tt : 0.0) * r2;
            const double d = sqrt(lf * lf - mf * mf);
            cr[l][m] = (a * r1 - b * r0) / d;
            ci[l][m] = (a * i1 - b * i0) / d;
        }
    }
    const double s2 = sqrt(2.0);
#pragma unroll
    for (int l = 0; l <= LMAX; l++) {
        out[l * l + l] = cr[l][0];
#pragma unroll
        for (int m = 1; m <= l; m++) {
            out[l * l + l + m] = s2 * cr[l][m];
            out[l * l + l - m] = s2 * ci[l][m];
        }
    }
}

__device__ inline double block_sum256p(double v, double *sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
#pragma unroll
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

template <typename T>
__global__ __launch_bounds__(256) void k_moments(const T *rho, int n, double h, double *partials) {
    __shared__ double sh[256];
    const long N = (long)n * n * n, nn = (long)n * n;
    const double o = -((double)n - 1.0) * h / 2.0;
    double acc[NMOM];
#pragma unroll
    for (int q = 0; q < NMOM; q++) acc[q] = 0.0;
    for (long p = blockIdx.x * 256L + threadIdx.x; p < N; p += (long)gridDim.x * 256) {
        const double rv = (double)rho[p];
        if (rv == 0.0) continue;
        const long k = p / nn, rem = p - k * nn, j = rem / n, i = rem - j * n;
        double hv[NMOM];
        solid_harmonics(o + (double)i * h, o + (double)j * h, o + (double)k * h, hv);
#pragma unroll
        for (int q = 0; q < NMOM; q++) acc[q] += rv * hv[q];
    }
    for (int q = 0; q < NMOM; q++) {
        const double r = block_sum256p(acc[q], sh);
        if (threadIdx.x == 0) partials[(long)q * RED_BLOCKS + blockIdx.x] = r;
    }
}

__global__ __launch_bounds__(256) void k_moments_finish(const double *partials, double *q, double scale) {
    __shared__ double sh[256];
    double acc = 0.0;
    for (int p = threadIdx.x; p < RED_BLOCKS; p += 256) acc += partials[(long)blockIdx.x * RED_BLOCKS + p];
    const double r = block_sum256p(acc, sh);
    if (threadIdx.x == 0) q[blockIdx.x] = r * scale;
}

template <typename T> void multipole_moments(const T *rho, int n, double h, double *partials, double *q, hipStream_t s) {
    hipLaunchKernelGGL(k_moments<T>, dim3(RED_BLOCKS), dim3(256), 0, s, rho, n, h, partials);
    LIRA_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_moments_finish, dim3(NMOM), dim3(256), 0, s, partials, q, h * h * h);
    LIRA_HIP(hipGetLastError());
}

// V_bc(r) = sum_l (sum_m Q_lm R_lm(r)) / |r|^{2l+1}  (qm/src/poisson.rs:175-190)
__device__ inline double multipole_potential(const double *q, double x, double y, double z) {
    const double d2 = x * x + y * y + z * z;
    double hv[NMOM];
    solid_harmonics(x, y, z, hv);
    double acc = 0.0, inv_pow = 1.0 / sqrt(d2);
#pragma unroll
    for (int l = 0; l <= LMAX; l++) {
        double sl = 0.0;
#pragma unroll
        for (int m = -l; m <= l; m++) sl += q[l * l + l + m] * hv[l * l + l + m];
        acc += sl * inv_pow;
        inv_pow /= d2;
    }
    return acc;
}

template <typename T>
__global__ __launch_bounds__(256) void k_poisson_rhs(double *b, const T *rho, const double *qg, int n, double h) {
    __shared__ double q[NMOM];
    if (threadIdx.x < NMOM) q[threadIdx.x] = qg[threadIdx.x];
    __syncthreads();
    const long N = (long)n * n * n, nn = (long)n * n;
    const long p = blockIdx.x * 256L + threadIdx.x;
    if (p >= N) return;
    double val = 4.0 * M_PI * (double)rho[p];
    const int k = (int)(p / nn), rem = (int)(p - (long)k * nn), j = rem / n, i = rem - j * n;
    const int pos[3] = {i, j, k};
    const bool near = i < 4 || j < 4 || k < 4 || i >= n - 4 || j >= n - 4 || k >= n - 4;
    if (near) {
        const double c[5] = {0.0, 8.0 / 5.0, -1.0 / 5.0, 8.0 / 315.0, -1.0 / 560.0};
        const double o = -((double)n - 1.0) * h / 2.0;
        double acc = 0.0;
        for (int a = 0; a < 3; a++)
            for (int m = 1; m <= 4; m++)
                for (int sg = -1; sg <= 1; sg += 2) {
                    const int qa = pos[a] + sg * m;
                    if (qa >= 0 && qa < n) continue;
                    double r[3] = {o + (double)i * h, o + (double)j * h, o + (double)k * h};
                    r[a] += (double)(sg * m) * h;
                    acc += c[m] * multipole_potential(q, r[0], r[1], r[2]);
                }
        val += acc / (h * h);
    }
    b[p] = val;
}

template <typename T> void poisson_rhs(double *b, const T *rho, const double *q, int n, double h, hipStream_t s) {
    const long N = (long)n * n * n;
    hipLaunchKernelGGL(k_poisson_rhs<T>, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, b, rho, q, n, h);
    LIRA_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------------------- multigrid ----
// Level operator A2 x = (diag x - nb)/H^2 with nb the six in-box neighbours and diag = 6 - (ghost factors of the
// touched faces) (qm-gpu/src/multigrid.rs:10-26): a missing neighbour at boundary distance d is (1 - 1/d) x0.
template <typename T>
__device__ inline void a2_parts(const T *x, int i, int j, int k, int n, T g_low, T g_high, T &nb, T &diag) {
    const long nn = (long)n * n;
    const long id = i + (long)n * j + nn * k;
    T s = T(0), g = T(0);
    if (i > 0) s += x[id - 1]; else g += g_low;
    if (i < n - 1) s += x[id + 1]; else g += g_high;
    if (j > 0) s += x[id - n]; else g += g_low;
    if (j < n - 1) s += x[id + n]; else g += g_high;
    if (k > 0) s += x[id - nn]; else g += g_low;
    if (k < n - 1) s += x[id + nn]; else g += g_high;
    nb = s;
    diag = T(6) - g;
}

// the level operators per point (shared by the per-level kernels and the fused coarse V-cycle, so both compute
// bitwise the same values)
template <typename T>
__device__ inline T mg_jacobi_pt(const T *x, const T *b, long p, int i, int j, int k, int n, T h2, T omega, T g_low,
                                 T g_high) {
    T nb, diag;
    a2_parts(x, i, j, k, n, g_low, g_high, nb, diag);
    const T xq = x[p];
    return xq + (omega / diag) * (h2 * b[p] + nb - diag * xq);
}
template <typename T>
__device__ inline T mg_residual_pt(const T *x, const T *b, long p, int i, int j, int k, int n, T h2, T g_low,
                                   T g_high) {
    T nb, diag;
    a2_parts(x, i, j, k, n, g_low, g_high, nb, diag);
    return b[p] + (nb - diag * x[p]) / h2;
}
// full weighting onto coarse point (i, j, k) from the fine point (2i + 1, 2j + 1, 2k + 1) and its 26 neighbours
template <typename T> __device__ inline T mg_restrict_pt(const T *f, int i, int j, int k, int nf) {
    const long nff = (long)nf * nf;
    T s = T(0);
    for (int dz = -1; dz <= 1; dz++) {
        const int c = 2 * k + 1 + dz;
        if (c < 0 || c >= nf) continue;
        const T wz = dz == 0 ? T(0.5) : T(0.25);
        for (int dy = -1; dy <= 1; dy++) {
            const int bb = 2 * j + 1 + dy;
            if (bb < 0 || bb >= nf) continue;
            const T wy = dy == 0 ? T(0.5) : T(0.25);
            for (int dx = -1; dx <= 1; dx++) {
                const int a = 2 * i + 1 + dx;
                if (a < 0 || a >= nf) continue;
                const T wx = dx == 0 ? T(0.5) : T(0.25);
                s += wx * wy * wz * f[a + (long)nf * bb + nff * c];
            }
        }
    }
    return s;
}
// trilinear interpolation of the coarse correction at fine point (i, j, k): odd fine indices sit on a coarse point,
// even ones halfway between two
template <typename T> __device__ inline T mg_prolong_pt(const T *c, int i, int j, int k, int nc) {
    const long ncc = (long)nc * nc;
    const int t[3] = {i, j, k};
    int ci[3][2];
    T wi[3][2];
    int cnt[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const int cc = t[a] >> 1;
        if (t[a] & 1) {
            ci[a][0] = cc;
            wi[a][0] = T(1);
            ci[a][1] = -1;
            wi[a][1] = T(0);
            cnt[a] = 1;
        } else {
            ci[a][0] = cc - 1;
            wi[a][0] = T(0.5);
            ci[a][1] = cc;
            wi[a][1] = T(0.5);
            cnt[a] = 2;
        }
    }
    T s = T(0);
    for (int a = 0; a < cnt[0]; a++)
        for (int b = 0; b < cnt[1]; b++)
            for (int d = 0; d < cnt[2]; d++) {
                const int x = ci[0][a], y = ci[1][b], z = ci[2][d];
                if (x < 0 || y < 0 || z < 0 || x >= nc || y >= nc || z >= nc) continue;
                s += wi[0][a] * wi[1][b] * wi[2][d] * c[x + (long)nc * y + ncc * z];
            }
    return s;
}

template <typename T>
__global__ void k_mg_jacobi(T *dst, const T *x, const T *b, int n, T h2, T omega, T g_low, T g_high) {
    const long N = (long)n * n * n, nn = (long)n * n;
    const long p = blockIdx.x * 256L + threadIdx.x;
    if (p >= N) return;
    const int k = (int)(p / nn), rem = (int)(p - (long)k * nn), j = rem / n, i = rem - j * n;
    dst[p] = mg_jacobi_pt(x, b, p, i, j, k, n, h2, omega, g_low, g_high);
}

// the first sweep of a level from x = 0 (the zero fill folded in): the Jacobi formula with x and its neighbours 0
template <typename T> __global__ void k_mg_jacobi0(T *dst, const T *b, int n, T h2, T omega, T g_low, T g_high) {
    const long N = (long)n * n * n, nn = (long)n * n;
    const long p = blockIdx.x * 256L + threadIdx.x;
    if (p >= N) return;
    const int k = (int)(p / nn), rem = (int)(p - (long)k * nn), j = rem / n, i = rem - j * n;
    T g = T(0);
    if (i == 0) g += g_low;
    if (i == n - 1) g += g_high;
    if (j == 0) g += g_low;
    if (j == n - 1) g += g_high;
    if (k == 0) g += g_low;
    if (k == n - 1) g += g_high;
    const T diag = T(6) - g;
    dst[p] = T(0) + (omega / diag) * (h2 * b[p] + T(0) - diag * T(0));
}

template <typename T> __global__ void k_mg_residual(T *dst, const T *x, const T *b, int n, T h2, T g_low, T g_high) {
    const long N = (long)n * n * n, nn = (long)n * n;
    const long p = blockIdx.x * 256L + threadIdx.x;
    if (p >= N) return;
    const int k = (int)(p / nn), rem = (int)(p - (long)k * nn), j = rem / n, i = rem - j * n;
    dst[p] = mg_residual_pt(x, b, p, i, j, k, n, h2, g_low, g_high);
}

template <typename T> __global__ void k_mg_restrict(T *dst, const T *f, int nc, int nf) {
    const long Nc = (long)nc * nc * nc, ncc = (long)nc * nc;
    const long p = blockIdx.x * 256L + threadIdx.x;
    if (p >= Nc) return;
    const int k = (int)(p / ncc), rem = (int)(p - (long)k * ncc), j = rem / nc, i = rem - j * nc;
    dst[p] = mg_restrict_pt(f, i, j, k, nf);
}

template <typename T> __global__ void k_mg_prolong_add(T *fine, const T *c, int nf, int nc) {
    const long Nf = (long)nf * nf * nf, nff = (long)nf * nf;
    const long p = blockIdx.x * 256L + threadIdx.x;
    if (p >= Nf) return;
    const int k = (int)(p / nff), rem = (int)(p - (long)k * nff), j = rem / nf, i = rem - j * nf;
    fine[p] += mg_prolong_pt(c, i, j, k, nc);
}

// The V-cycle from a small level down, in one workgroup with every level's x, b and scratch in LDS: the same
// sequence as the per-level kernels (zero fill, pre-smoothing, residual, restriction, ..., the coarsest sweeps,
// prolongation, post-smoothing), each pass a loop over the level's points followed by a barrier. Replaces some 25
// launches per V-cycle whose fixed dispatch cost dominates these tiny grids.
template <typename T> __global__ __launch_bounds__(1024) void k_mg_coarse(T *x_out, const T *b_in, MgCoarse a) {
    extern __shared__ __align__(16) unsigned char mg_smem[];
    T *lds = reinterpret_cast<T *>(mg_smem);
    const int tid = threadIdx.x;
    const T omega = (T)a.omega;
    auto level = [&](int l, T *&x, T *&b, T *&tmp) {
        const int N = a.n[l] * a.n[l] * a.n[l];
        x = lds + a.off[l];
        b = x + N;
        tmp = b + N;
    };
    auto decode = [](int p, int n, int &i, int &j, int &k) {
        k = p / (n * n);
        const int rem = p - k * n * n;
        j = rem / n;
        i = rem - j * n;
    };
    auto smooth = [&](int l, int sweeps) {
        T *x, *b, *tmp;
        level(l, x, b, tmp);
        const int n = a.n[l], N = n * n * n;
        const T h2 = (T)a.h2[l], gl = (T)a.gl[l], gh = (T)a.gh[l];
        for (int sw = 0; sw < sweeps; sw += 2) {
            for (int p = tid; p < N; p += 1024) {
                int i, j, k;
                decode(p, n, i, j, k);
                tmp[p] = mg_jacobi_pt<T>(x, b, p, i, j, k, n, h2, omega, gl, gh);
            }
            __syncthreads();
            for (int p = tid; p < N; p += 1024) {
                int i, j, k;
                decode(p, n, i, j, k);
                x[p] = mg_jacobi_pt<T>(tmp, b, p, i, j, k, n, h2, omega, gl, gh);
            }
            __syncthreads();
        }
    };
    {
        T *x, *b, *tmp;
        level(0, x, b, tmp);
        const int N = a.n[0] * a.n[0] * a.n[0];
        for (int p = tid; p < N; p += 1024) b[p] = b_in[p];
    }
    __syncthreads();
    const int L = a.levels;
    for (int l = 0; l < L; l++) {
        T *x, *b, *tmp;
        level(l, x, b, tmp);
        const int n = a.n[l], N = n * n * n;
        for (int p = tid; p < N; p += 1024) x[p] = T(0);
        __syncthreads();
        if (l == L - 1) {
            smooth(l, a.coarse);
            break;
        }
        smooth(l, a.pre);
        const T h2 = (T)a.h2[l], gl = (T)a.gl[l], gh = (T)a.gh[l];
        for (int p = tid; p < N; p += 1024) {
            int i, j, k;
            decode(p, n, i, j, k);
            tmp[p] = mg_residual_pt<T>(x, b, p, i, j, k, n, h2, gl, gh);
        }
        __syncthreads();
        T *xc, *bc, *tc;
        level(l + 1, xc, bc, tc);
        const int nc = a.n[l + 1], Nc = nc * nc * nc;
        for (int p = tid; p < Nc; p += 1024) {
            int i, j, k;
            decode(p, nc, i, j, k);
            bc[p] = mg_restrict_pt<T>(tmp, i, j, k, n);
        }
        __syncthreads();
    }
    for (int l = L - 2; l >= 0; l--) {
        T *x, *b, *tmp, *xc, *bc, *tc;
        level(l, x, b, tmp);
        level(l + 1, xc, bc, tc);
        const int n = a.n[l], N = n * n * n;
        for (int p = tid; p < N; p += 1024) {
            int i, j, k;
            decode(p, n, i, j, k);
            x[p] += mg_prolong_pt<T>(xc, i, j, k, a.n[l + 1]);
        }
        __syncthreads();
        smooth(l, a.post);
    }
    {
        T *x, *b, *tmp;
        level(0, x, b, tmp);
        const int N = a.n[0] * a.n[0] * a.n[0];
        for (int p = tid; p < N; p += 1024) x_out[p] = x[p];
    }
}

#define LIRA_G1(N) dim3((unsigned)(((N) + 255) / 256)), dim3(256)

template <typename T>
void mg_jacobi(T *dst, const T *x, const T *b, int n, double h2, double omega, double g_low, double g_high,
               hipStream_t s) {
    const long N = (long)n * n * n;
    hipLaunchKernelGGL(k_mg_jacobi<T>, LIRA_G1(N), 0, s, dst, x, b, n, (T)h2, (T)omega, (T)g_low, (T)g_high);
    LIRA_HIP(hipGetLastError());
}
template <typename T>
void mg_residual(T *dst, const T *x, const T *b, int n, double h2, double g_low, double g_high, hipStream_t s) {
    const long N = (long)n * n * n;
    hipLaunchKernelGGL(k_mg_residual<T>, LIRA_G1(N), 0, s, dst, x, b, n, (T)h2, (T)g_low, (T)g_high);
    LIRA_HIP(hipGetLastError());
}
template <typename T> void mg_restrict(T *dst, const T *fine, int nc, int nf, hipStream_t s) {
    const long N = (long)nc * nc * nc;
    hipLaunchKernelGGL(k_mg_restrict<T>, LIRA_G1(N), 0, s, dst, fine, nc, nf);
    LIRA_HIP(hipGetLastError());
}
template <typename T> void mg_prolong_add(T *fine, const T *coarse, int nf, int nc, hipStream_t s) {
    const long N = (long)nf * nf * nf;
    hipLaunchKernelGGL(k_mg_prolong_add<T>, LIRA_G1(N), 0, s, fine, coarse, nf, nc);
    LIRA_HIP(hipGetLastError());
}
template <typename T>
void mg_jacobi0(T *dst, const T *b, int n, double h2, double omega, double g_low, double g_high, hipStream_t s) {
    const long N = (long)n * n * n;
    hipLaunchKernelGGL(k_mg_jacobi0<T>, LIRA_G1(N), 0, s, dst, b, n, (T)h2, (T)omega, (T)g_low, (T)g_high);
    LIRA_HIP(hipGetLastError());
}
// LDS bytes of the fused V-cycle over the levels of a (x, b and scratch per level)
size_t mg_coarse_bytes(const MgCoarse &a, int esize) {
    size_t e = 0;
    for (int l = 0; l < a.levels; l++) e += 3 * (size_t)a.n[l] * a.n[l] * a.n[l];
    return e * esize;
}
template <typename T> void mg_coarse(T *x, const T *b, const MgCoarse &a, hipStream_t s) {
    hipLaunchKernelGGL(k_mg_coarse<T>, dim3(1), dim3(1024), mg_coarse_bytes(a, (int)sizeof(T)), s, x, b, a);
    LIRA_HIP(hipGetLastError());
}

// ------------------------------------------------------------------------------------ CG on device ----
template <typename Tv, typename Tr>
__global__ void k_cg_update(Tv *x, const Tr *p, Tr *r, const Tr *q, const double *sc, int ia, int ib, long N) {
    const long i = blockIdx.x * 256L + threadIdx.x;
    if (i >= N) return;
    const double alpha = sc[ia] / sc[ib];
    x[i] = (Tv)((double)x[i] + alpha * (double)p[i]);
    r[i] = (Tr)((double)r[i] - alpha * (double)q[i]);
}
template <typename Tv, typename Tr>
void cg_update(Tv *x, const Tr *p, Tr *r, const Tr *q, const double *sc, int ia, int ib, long N, hipStream_t s) {
    hipLaunchKernelGGL((k_cg_update<Tv, Tr>), LIRA_G1(N), 0, s, x, p, r, q, sc, ia, ib, N);
    LIRA_HIP(hipGetLastError());
}

template <typename T> __global__ void k_cg_direction(T *p, const T *z, const double *sc, int ia, int ib, long N) {
    const long i = blockIdx.x * 256L + threadIdx.x;
    if (i >= N) return;
    const double beta = sc[ia] / sc[ib];
    p[i] = (T)((double)z[i] + beta * (double)p[i]);
}
template <typename T> void cg_direction(T *p, const T *z, const double *sc, int ia, int ib, long N, hipStream_t s) {
    hipLaunchKernelGGL(k_cg_direction<T>, LIRA_G1(N), 0, s, p, z, sc, ia, ib, N);
    LIRA_HIP(hipGetLastError());
}

template <typename T> __global__ void k_add_into(double *x, const T *e, long N) {
    const long i = blockIdx.x * 256L + threadIdx.x;
    if (i < N) x[i] += (double)e[i];
}
template <typename T> void add_into_f64(double *x, const T *e, long N, hipStream_t s) {
    hipLaunchKernelGGL(k_add_into<T>, LIRA_G1(N), 0, s, x, e, N);
    LIRA_HIP(hipGetLastError());
}

#define LIRA_PINST(T)                                                                                         \
    template void multipole_moments<T>(const T *, int, double, double *, double *, hipStream_t);              \
    template void poisson_rhs<T>(double *, const T *, const double *, int, double, hipStream_t);              \
    template void mg_jacobi<T>(T *, const T *, const T *, int, double, double, double, double, hipStream_t);  \
    template void mg_residual<T>(T *, const T *, const T *, int, double, double, double, hipStream_t);        \
    template void mg_restrict<T>(T *, const T *, int, int, hipStream_t);                                      \
    template void mg_prolong_add<T>(T *, const T *, int, int, hipStream_t);                                   \
    template void mg_jacobi0<T>(T *, const T *, int, double, double, double, double, hipStream_t);            \
    template void mg_coarse<T>(T *, const T *, const MgCoarse &, hipStream_t);                                \
    template void cg_direction<T>(T *, const T *, const double *, int, int, long, hipStream_t);               \
    template void add_into_f64<T>(double *, const T *, long, hipStream_t);
LIRA_PINST(float)
LIRA_PINST(double)
template void cg_update<double, double>(double *, const double *, double *, const double *, const double *, int, int,
                                        long, hipStream_t);
template void cg_update<float, float>(float *, const float *, float *, const float *, const double *, int, int, long,
                                      hipStream_t);
template void cg_update<double, float>(double *, const float *, float *, const float *, const double *, int, int,
                                       long, hipStream_t);

}  // namespace lira
