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

// ------------------------------------------------------------------------------------ multipole moments ----
// Q_lm = h^3 sum_p rho_p R_lm(r_p) (qm/src/poisson.rs:155-162). R_lm is a homogeneous polynomial of degree l in
// (x, y, z), so Q_lm = h^3 sum_{a+b+c=l} C_lm,abc M_abc with the raw moments M_abc = sum_p rho_p x^a y^b z^c
// (a + b + c <= 6, 84 of them), and those separate over the grid's axes: rows S_a(j,k) = sum_i x_i^a rho(i,j,k),
// planes T_ab(k) = sum_j y_j^b S_a(j,k), then M_abc = sum_k z_k^c T_ab(k). One streaming pass over rho with seven
// f64 accumulations per point, two small contractions, and the C table built on the host by running the recursion
// of solid_harmonics on polynomials (moment_coefficients below). Monomials are enumerated a outer, b, c inner.
constexpr int N_PAIRS = 28;      // (a, b) with a + b <= 6
constexpr int N_MONOMIALS = 84;  // (a, b, c) with a + b + c <= 6

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// one wave per row (j, k): rows[a * n^2 + j + n k] = S_a(j, k)
template <typename T> __global__ __launch_bounds__(256) void k_moment_rows(const T *rho, int n, double h, double *rows) {
    const int lane = threadIdx.x & 63;
    const long nn = (long)n * n;
    const long row = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= nn) return;  // a whole wave
    const double o = -((double)n - 1.0) * h / 2.0;
    const T *src = rho + row * n;
    double s[LMAX + 1];
#pragma unroll
    for (int a = 0; a <= LMAX; a++) s[a] = 0.0;
#pragma unroll 4
    for (int i = lane; i < n; i += 64) {
        const double x = o + (double)i * h;
        double t = (double)src[i];
#pragma unroll
        for (int a = 0; a <= LMAX; a++) {
            s[a] += t;
            t *= x;
        }
    }
#pragma unroll
    for (int a = 0; a <= LMAX; a++) {
        const double v = wave_sum(s[a]);
        if (lane == 0) rows[a * nn + row] = v;
    }
}

// one block per plane k: planes[pair * n + k] = T_ab(k)
__global__ __launch_bounds__(256) void k_moment_planes(const double *rows, int n, double h, double *planes) {
    __shared__ double sh[4][N_PAIRS];
    const long nn = (long)n * n;
    const int k = blockIdx.x;
    const double o = -((double)n - 1.0) * h / 2.0;
    double t[N_PAIRS];
#pragma unroll
    for (int q = 0; q < N_PAIRS; q++) t[q] = 0.0;
    for (int j = threadIdx.x; j < n; j += 256) {
        const double y = o + (double)j * h;
        double yp[LMAX + 1];
        yp[0] = 1.0;
#pragma unroll
        for (int b = 1; b <= LMAX; b++) yp[b] = yp[b - 1] * y;
        int q = 0;
#pragma unroll
        for (int a = 0; a <= LMAX; a++) {
            const double sa = rows[a * nn + j + (long)n * k];
#pragma unroll
            for (int b = 0; b <= LMAX - a; b++) t[q++] += sa * yp[b];
        }
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < N_PAIRS; q++) {
        const double v = wave_sum(t[q]);
        if (lane == 0) sh[w][q] = v;
    }
    __syncthreads();
    if (threadIdx.x < N_PAIRS) {
        const int q = threadIdx.x;
        planes[(long)q * n + k] = (sh[0][q] + sh[1][q]) + (sh[2][q] + sh[3][q]);
    }
}

// one block: M_abc = sum_k z_k^c T_ab(k) (monomial t on wave t % 4, its k spread over the lanes), then
// q[lm] = h^3 sum_t coef[lm * 84 + t] M_t
__global__ __launch_bounds__(256) void k_moment_finish(const double *planes, int n, double h, const double *coef,
                                                       double *q) {
    __shared__ double mono[N_MONOMIALS];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const double o = -((double)n - 1.0) * h / 2.0;
    int t = 0, pair = 0;
    for (int a = 0; a <= LMAX; a++)
        for (int b = 0; b <= LMAX - a; b++, pair++)
            for (int c = 0; c <= LMAX - a - b; c++, t++) {
                if ((t & 3) != w) continue;
                double m = 0.0;
                for (int k = lane; k < n; k += 64) {
                    const double z = o + (double)k * h;
                    double zc = 1.0;
                    for (int e = 0; e < c; e++) zc *= z;
                    m += planes[(long)pair * n + k] * zc;
                }
                m = wave_sum(m);
                if (lane == 0) mono[t] = m;
            }
    __syncthreads();
    if (threadIdx.x < NMOM) {
        const int lm = threadIdx.x;
        double acc = 0.0;
        for (int u = 0; u < N_MONOMIALS; u++) acc += coef[lm * N_MONOMIALS + u] * mono[u];
        q[lm] = acc * (h * h * h);
    }
}

// The monomial coefficients of R_lm: solid_harmonics' recursion run on polynomials in (x, y, z) (dense over
// exponents 0..6 per axis), then packed as coef[lm * 84 + t] in the monomial order above.
void moment_coefficients(double *coef) {
    constexpr int E = LMAX + 1;
    struct Poly {
        double c[E][E][E];
    };
    auto zero = [](Poly &p) {
        for (auto &a : p.c)
            for (auto &b : a)
                for (double &v : b) v = 0.0;
    };
    // dst = s1 * (axis-monomial u times p1) + s2 * (axis-monomial w times p2), u/w one of x, y, z, r^2 (0..3) or none
    auto times = [](const Poly &p, int axis, Poly &out) {  // out = m * p, m = x, y, z (axis 0..2) or r^2 (axis 3)
        for (auto &a : out.c)
            for (auto &b : a)
                for (double &v : b) v = 0.0;
        for (int a = 0; a < E; a++)
            for (int b = 0; b < E; b++)
                for (int c = 0; c < E; c++) {
                    const double v = p.c[a][b][c];
                    if (v == 0.0) continue;
                    if (axis == 0 && a + 1 < E) out.c[a + 1][b][c] += v;
                    if (axis == 1 && b + 1 < E) out.c[a][b + 1][c] += v;
                    if (axis == 2 && c + 1 < E) out.c[a][b][c + 1] += v;
                    if (axis == 3) {
                        if (a + 2 < E) out.c[a + 2][b][c] += v;
                        if (b + 2 < E) out.c[a][b + 2][c] += v;
                        if (c + 2 < E) out.c[a][b][c + 2] += v;
                    }
                }
    };
    static Poly cr[E][E], ci[E][E];
    for (int l = 0; l < E; l++)
        for (int m = 0; m < E; m++) {
            zero(cr[l][m]);
            zero(ci[l][m]);
        }
    cr[0][0].c[0][0][0] = 1.0;
    Poly t1, t2;
    for (int l = 1; l <= LMAX; l++) {
        const double lf = (double)l;
        const double f = std::sqrt((2.0 * lf - 1.0) / (2.0 * lf));
        // C_l^l = f (x + i y) C_{l-1}^{l-1}
        times(cr[l - 1][l - 1], 0, t1);
        times(ci[l - 1][l - 1], 1, t2);
        for (int a = 0; a < E; a++)
            for (int b = 0; b < E; b++)
                for (int c = 0; c < E; c++) cr[l][l].c[a][b][c] = f * (t1.c[a][b][c] - t2.c[a][b][c]);
        times(ci[l - 1][l - 1], 0, t1);
        times(cr[l - 1][l - 1], 1, t2);
        for (int a = 0; a < E; a++)
            for (int b = 0; b < E; b++)
                for (int c = 0; c < E; c++) ci[l][l].c[a][b][c] = f * (t1.c[a][b][c] + t2.c[a][b][c]);
        // C_l^m = ((2l - 1) z C_{l-1}^m - sqrt((l-1)^2 - m^2) r^2 C_{l-2}^m) / sqrt(l^2 - m^2)
        for (int m = 0; m < l; m++) {
            const double mf = (double)m;
            const double tt = (lf - 1.0) * (lf - 1.0) - mf * mf;
            const double bq = std::sqrt(tt > 0.0 ? tt : 0.0);
            const double d = std::sqrt(lf * lf - mf * mf);
            for (int part = 0; part < 2; part++) {
                Poly(&C)[E][E] = part == 0 ? cr : ci;
                times(C[l - 1][m], 2, t1);
                zero(t2);
                if (l >= 2 && m <= l - 2) times(C[l - 2][m], 3, t2);
                for (int a = 0; a < E; a++)
                    for (int b = 0; b < E; b++)
                        for (int c = 0; c < E; c++)
                            C[l][m].c[a][b][c] = ((2.0 * lf - 1.0) * t1.c[a][b][c] - bq * t2.c[a][b][c]) / d;
            }
        }
    }
    const double s2 = std::sqrt(2.0);
    for (int i = 0; i < NMOM * N_MONOMIALS; i++) coef[i] = 0.0;
    auto pack = [&](int lm, const Poly &p, double scale) {
        int idx = 0;
        for (int a = 0; a <= LMAX; a++)
            for (int b = 0; b <= LMAX - a; b++)
                for (int c = 0; c <= LMAX - a - b; c++, idx++) coef[lm * N_MONOMIALS + idx] = scale * p.c[a][b][c];
    };
    for (int l = 0; l <= LMAX; l++) {
        pack(l * l + l, cr[l][0], 1.0);
        for (int m = 1; m <= l; m++) {
            pack(l * l + l + m, cr[l][m], s2);
            pack(l * l + l - m, ci[l][m], s2);
        }
    }
}

size_t moment_work_doubles(int n) { return (size_t)(LMAX + 1) * n * n + (size_t)N_PAIRS * n; }

template <typename T>
void multipole_moments(const T *rho, int n, double h, double *work, const double *coef, double *q, hipStream_t s) {
    const long nn = (long)n * n;
    double *rows = work, *planes = work + (LMAX + 1) * nn;
    hipLaunchKernelGGL(k_moment_rows<T>, dim3((unsigned)((nn + 3) / 4)), dim3(256), 0, s, rho, n, h, rows);
    LIRA_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_moment_planes, dim3((unsigned)n), dim3(256), 0, s, rows, n, h, planes);
    LIRA_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_moment_finish, dim3(1), dim3(256), 0, s, planes, n, h, coef, q);
    LIRA_HIP(hipGetLastError());
}

// ----------------------------------------------------------------------------------- boundary terms ----
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

// The 8th-order Laplacian of an edge point reaches up to 4 points past a face; those ghost points carry V_bc. Each is
// evaluated once, one thread per ghost point: ghost[(slab * n^2 + u + n v)], slab = (axis * 2 + side) * 4 + (m - 1)
// for the point m steps past the low (side 0) or high (side 1) face of that axis, (u, v) the indices along the other
// two axes in increasing axis order.
__global__ __launch_bounds__(256) void k_boundary_ghosts(double *ghost, const double *qg, int n, double h) {
    __shared__ double q[NMOM];
    if (threadIdx.x < NMOM) q[threadIdx.x] = qg[threadIdx.x];
    __syncthreads();
    const long nn = (long)n * n;
    const long g = blockIdx.x * 256L + threadIdx.x;
    if (g >= 24 * nn) return;
    const int slab = (int)(g / nn);
    const int uv = (int)(g - (long)slab * nn);
    const int u = uv % n, v = uv / n;
    const int axis = slab >> 3, side = (slab >> 2) & 1, m = (slab & 3) + 1;
    const double o = -((double)n - 1.0) * h / 2.0;
    int idx[3];
    idx[axis] = side ? n - 1 + m : -m;
    idx[axis == 0 ? 1 : 0] = u;
    idx[axis == 2 ? 1 : 2] = v;
    ghost[g] = multipole_potential(q, o + (double)idx[0] * h, o + (double)idx[1] * h, o + (double)idx[2] * h);
}

// b = 4 pi rho + (sum over an edge point's ghost neighbours of c_m V_bc) / h^2 (qm/src/poisson.rs:174-197,
// 204-214): the terms in the order axis, step m = 1..4, below then above, as the oracle sums them.
template <typename T>
__global__ __launch_bounds__(256) void k_poisson_rhs(double *b, const T *rho, const double *ghost, int n, double h) {
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
    if (i >= n || j >= n) return;
    const long nn = (long)n * n;
    const long p = i + (long)n * j + nn * k;
    double val = 4.0 * M_PI * (double)rho[p];
    const bool near = i < 4 || j < 4 || k < 4 || i >= n - 4 || j >= n - 4 || k >= n - 4;
    if (near) {
        const double c[5] = {0.0, 8.0 / 5.0, -1.0 / 5.0, 8.0 / 315.0, -1.0 / 560.0};
        const int pos[3] = {i, j, k};
        double acc = 0.0;
        for (int a = 0; a < 3; a++) {
            const int u = a == 0 ? j : i, v = a == 2 ? j : k;
            for (int m = 1; m <= 4; m++)
                for (int sg = -1; sg <= 1; sg += 2) {
                    const int qa = pos[a] + sg * m;
                    if (qa >= 0 && qa < n) continue;
                    const int side = sg > 0, layer = side ? qa - (n - 1) : -qa;
                    acc += c[m] * ghost[(long)((a * 2 + side) * 4 + layer - 1) * nn + u + (long)n * v];
                }
        }
        val += acc / (h * h);
    }
    b[p] = val;
}

template <typename T>
void poisson_rhs(double *b, const T *rho, const double *q, double *ghost, int n, double h, hipStream_t s) {
    const long nn = (long)n * n;
    hipLaunchKernelGGL(k_boundary_ghosts, dim3((unsigned)((24 * nn + 255) / 256)), dim3(256), 0, s, ghost, q, n, h);
    LIRA_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_poisson_rhs<T>, dim3((unsigned)((n + 63) / 64), (unsigned)((n + 3) / 4), (unsigned)n),
                       dim3(64, 4), 0, s, b, rho, ghost, n, h);
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

// The per-level kernels of the large levels. One thread per point on a 3-D launch (64 x 4 threads: x fastest, a
// wave on 64 consecutive points of one row; no index divisions). Every neighbour load is issued unconditionally from a
// clamped, always valid address and a missing neighbour's term is replaced by an exact zero afterwards, so the loads of
// a thread are in flight together (a load under a bounds branch made the compiler wait for each in turn). Each sum runs
// over the same terms in the same order as the per-point functions above (an added exact zero leaves a float sum
// unchanged), so these kernels produce the values of the fused coarse kernel bit for bit.
constexpr int MG_BX = 64, MG_BY = 4;
inline dim3 mg_grid(int n) { return dim3((unsigned)((n + MG_BX - 1) / MG_BX), (unsigned)((n + MG_BY - 1) / MG_BY), (unsigned)n); }

template <typename T>
__device__ inline void a2_parts_bf(const T *x, int i, int j, int k, int n, T g_low, T g_high, T &nb, T &diag, T &xq) {
    const long nn = (long)n * n;
    const long id = i + (long)n * j + nn * k;
    const bool il = i > 0, ih = i < n - 1, jl = j > 0, jh = j < n - 1, kl = k > 0, kh = k < n - 1;
    const T a0 = x[il ? id - 1 : id], a1 = x[ih ? id + 1 : id], a2 = x[jl ? id - n : id], a3 = x[jh ? id + n : id];
    const T a4 = x[kl ? id - nn : id], a5 = x[kh ? id + nn : id];
    xq = x[id];
    T s = T(0), g = T(0);
    s += il ? a0 : T(0);
    g += il ? T(0) : g_low;
    s += ih ? a1 : T(0);
    g += ih ? T(0) : g_high;
    s += jl ? a2 : T(0);
    g += jl ? T(0) : g_low;
    s += jh ? a3 : T(0);
    g += jh ? T(0) : g_high;
    s += kl ? a4 : T(0);
    g += kl ? T(0) : g_low;
    s += kh ? a5 : T(0);
    g += kh ? T(0) : g_high;
    nb = s;
    diag = T(6) - g;
}

template <typename T>
__global__ __launch_bounds__(256) void k_mg_jacobi(T *dst, const T *x, const T *b, int n, T h2, T omega, T g_low,
                                                   T g_high) {
    const int i = blockIdx.x * MG_BX + threadIdx.x, j = blockIdx.y * MG_BY + threadIdx.y, k = blockIdx.z;
    if (i >= n || j >= n) return;
    const long p = i + (long)n * j + (long)n * n * k;
    T nb, diag, xq;
    a2_parts_bf(x, i, j, k, n, g_low, g_high, nb, diag, xq);
    dst[p] = xq + (omega / diag) * (h2 * b[p] + nb - diag * xq);
}

// a block's sum of one double per thread (256 threads, fixed order): thread 0 returns it
__device__ inline double block_sum256(double v, double *sh) {
    v = wave_sum(v);
    if ((threadIdx.x + threadIdx.y * blockDim.x) % 64 == 0) sh[(threadIdx.x + threadIdx.y * blockDim.x) / 64] = v;
    __syncthreads();
    return (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// k_mg_jacobi that also leaves each wave's share of sum_p b[p] dst[p] (the CG's r.z when b is the residual and dst the
// preconditioned one) in partials[block * 4 + wave]; every lane takes part in its wave's sum
template <typename T>
__global__ __launch_bounds__(256) void k_mg_jacobi_dot(T *dst, const T *x, const T *b, int n, T h2, T omega, T g_low,
                                                       T g_high, double *partials) {
    const int i = blockIdx.x * MG_BX + threadIdx.x, j = blockIdx.y * MG_BY + threadIdx.y, k = blockIdx.z;
    double d = 0.0;
    if (i < n && j < n) {
        const long p = i + (long)n * j + (long)n * n * k;
        T nb, diag, xq;
        a2_parts_bf(x, i, j, k, n, g_low, g_high, nb, diag, xq);
        const T bp = b[p];
        const T out = xq + (omega / diag) * (h2 * bp + nb - diag * xq);
        dst[p] = out;
        d = (double)bp * (double)out;
    }
    // one partial per wave (no block barrier in a streaming kernel)
    d = wave_sum(d);
    if (threadIdx.x == 0)
        partials[(blockIdx.x + (long)gridDim.x * (blockIdx.y + (long)gridDim.y * blockIdx.z)) * MG_BY + threadIdx.y] = d;
}

// the first sweep of a level from x = 0 (the zero fill folded in): the Jacobi formula with x and its neighbours 0
template <typename T>
__global__ __launch_bounds__(256) void k_mg_jacobi0(T *dst, const T *b, int n, T h2, T omega, T g_low, T g_high) {
    const int i = blockIdx.x * MG_BX + threadIdx.x, j = blockIdx.y * MG_BY + threadIdx.y, k = blockIdx.z;
    if (i >= n || j >= n) return;
    const long p = i + (long)n * j + (long)n * n * k;
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

template <typename T>
__global__ __launch_bounds__(256) void k_mg_residual(T *dst, const T *x, const T *b, int n, T h2, T g_low, T g_high) {
    const int i = blockIdx.x * MG_BX + threadIdx.x, j = blockIdx.y * MG_BY + threadIdx.y, k = blockIdx.z;
    if (i >= n || j >= n) return;
    const long p = i + (long)n * j + (long)n * n * k;
    T nb, diag, xq;
    a2_parts_bf(x, i, j, k, n, g_low, g_high, nb, diag, xq);
    dst[p] = b[p] + (nb - diag * xq) / h2;
}

// full weighting onto coarse point (i, j, k) from the 27 fine points about (2i + 1, 2j + 1, 2k + 1); a fine index past
// the grid (the top of an even level) reads a clamped address and contributes an exact zero
template <typename T> __global__ __launch_bounds__(256) void k_mg_restrict(T *dst, const T *f, int nc, int nf) {
    const int i = blockIdx.x * MG_BX + threadIdx.x, j = blockIdx.y * MG_BY + threadIdx.y, k = blockIdx.z;
    if (i >= nc || j >= nc) return;
    const long nff = (long)nf * nf;
    T s = T(0);
#pragma unroll
    for (int dz = -1; dz <= 1; dz++) {
        const int c = 2 * k + 1 + dz;
        const bool cz = c < nf;  // c >= 0 always
        const T wz = dz == 0 ? T(0.5) : T(0.25);
#pragma unroll
        for (int dy = -1; dy <= 1; dy++) {
            const int bb = 2 * j + 1 + dy;
            const bool cy = bb < nf;
            const T wy = dy == 0 ? T(0.5) : T(0.25);
#pragma unroll
            for (int dx = -1; dx <= 1; dx++) {
                const int a = 2 * i + 1 + dx;
                const bool cx = a < nf;
                const T wx = dx == 0 ? T(0.5) : T(0.25);
                const T v = f[(cx ? a : nf - 1) + (long)nf * (cy ? bb : nf - 1) + nff * (cz ? c : nf - 1)];
                s += (cx && cy && cz) ? wx * wy * wz * v : T(0);
            }
        }
    }
    dst[i + (long)nc * j + (long)nc * nc * k] = s;
}

// the trilinear interpolation of the coarse correction at fine point (i, j, k) added in place: per axis an odd fine
// index sits on coarse point t >> 1 (weight 1), an even one between (t >> 1) - 1 and t >> 1 (1/2 each); the eight
// terms in the order of mg_prolong_pt, a missing one an exact zero
template <typename T> __global__ __launch_bounds__(256) void k_mg_prolong_add(T *fine, const T *c, int nf, int nc) {
    const int i = blockIdx.x * MG_BX + threadIdx.x, j = blockIdx.y * MG_BY + threadIdx.y, k = blockIdx.z;
    if (i >= nf || j >= nf) return;
    const long ncc = (long)nc * nc;
    const int t[3] = {i, j, k};
    int ci[3][2];
    T wi[3][2];
    bool ok[3][2];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const bool odd = t[a] & 1;
        const int cc = t[a] >> 1;
        ci[a][0] = odd ? cc : cc - 1;
        ci[a][1] = cc;
        wi[a][0] = odd ? T(1) : T(0.5);
        wi[a][1] = T(0.5);
        ok[a][0] = ci[a][0] >= 0 && ci[a][0] < nc;
        ok[a][1] = !odd && cc < nc;  // an odd index has one term
    }
    T s = T(0);
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int d = 0; d < 2; d++) {
                const bool in = ok[0][a] && ok[1][b] && ok[2][d];
                const int x = in ? ci[0][a] : 0, y = in ? ci[1][b] : 0, z = in ? ci[2][d] : 0;
                const T v = c[x + (long)nc * y + ncc * z];
                s += in ? wi[0][a] * wi[1][b] * wi[2][d] * v : T(0);
            }
    fine[i + (long)nf * j + (long)nf * nf * k] += s;
}

// Restriction with the fine rows read as pairs: coarse point i takes fine 2i, 2i + 1 as one vector load and 2i + 2
// from its x neighbour's pair (a lane shuffle; the wave's last lane, and the last coarse point, load it). One vector
// load per fine row instead of three strided ones; the 27 terms in mg_restrict_pt's order. Needs an even fine n (the
// pairs aligned); every lane stays live for the shuffles and only stores inside the level.
template <typename T> struct Pair2;
template <> struct Pair2<float> { typedef float type __attribute__((ext_vector_type(2))); };
template <> struct Pair2<double> { typedef double type __attribute__((ext_vector_type(2))); };

template <typename T> __global__ __launch_bounds__(256) void k_mg_restrict_pairs(T *dst, const T *f, int nc, int nf) {
    typedef typename Pair2<T>::type P2;
    const int i = blockIdx.x * MG_BX + threadIdx.x, j = blockIdx.y * MG_BY + threadIdx.y, k = blockIdx.z;
    const int ic = i < nc ? i : nc - 1, jc = j < nc ? j : nc - 1;  // loads of lanes past the level stay in it
    const long nff = (long)nf * nf;
    const bool last = threadIdx.x == MG_BX - 1 || i + 1 >= nc;
    T s = T(0);
#pragma unroll
    for (int dz = -1; dz <= 1; dz++) {
        const int c = 2 * k + 1 + dz;
        const bool cz = c < nf;
        const T wz = dz == 0 ? T(0.5) : T(0.25);
#pragma unroll
        for (int dy = -1; dy <= 1; dy++) {
            const int bb = 2 * jc + 1 + dy;
            const bool cy = bb < nf;
            const T wy = dy == 0 ? T(0.5) : T(0.25);
            const T *row = f + (long)nf * (cy ? bb : nf - 1) + nff * (cz ? c : nf - 1);
            const P2 pr = *reinterpret_cast<const P2 *>(row + 2 * ic);
            const int a2 = 2 * ic + 2;
            T x2 = __shfl_down(pr.x, 1, 64);
            if (last) x2 = row[a2 < nf ? a2 : nf - 1];
            const bool c2 = a2 < nf;
            s += (cy && cz) ? T(0.25) * wy * wz * pr.x : T(0);
            s += (cy && cz) ? T(0.5) * wy * wz * pr.y : T(0);
            s += (c2 && cy && cz) ? T(0.25) * wy * wz * x2 : T(0);
        }
    }
    if (i < nc && j < nc) dst[i + (long)nc * j + (long)nc * nc * k] = s;
}

// Prolongation-add with one thread per coarse cell: the cell's (up to) 8 fine children 2c + (0, 1) per axis share the
// 8 coarse values at c - 1 and c, loaded once; each child's terms in mg_prolong_pt's order and weights (an odd fine
// index takes coarse c with weight 1; an even one c - 1 and c with 1/2 each; a missing term an exact zero).
template <typename T> __global__ __launch_bounds__(256) void k_mg_prolong_cells(T *fine, const T *c, int nf, int nc) {
    const int hx = (nf + 1) / 2;
    const int cx = blockIdx.x * MG_BX + threadIdx.x, cy = blockIdx.y * MG_BY + threadIdx.y, cz = blockIdx.z;
    if (cx >= hx || cy >= hx) return;
    const long ncc = (long)nc * nc;
    const int cc[3] = {cx, cy, cz};
    int idx[3][2];
    bool ok[3][2];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        idx[a][0] = cc[a] - 1 < 0 ? 0 : (cc[a] - 1 < nc ? cc[a] - 1 : nc - 1);
        idx[a][1] = cc[a] < nc ? cc[a] : nc - 1;
        ok[a][0] = cc[a] >= 1;  // c - 1 <= nc - 1 always (c <= nc)
        ok[a][1] = cc[a] < nc;
    }
    T v[2][2][2];  // [x slot][y slot][z slot]
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int d = 0; d < 2; d++) v[a][b][d] = c[idx[0][a] + (long)nc * idx[1][b] + ncc * idx[2][d]];
#pragma unroll
    for (int sz = 0; sz < 2; sz++) {
        const int k = 2 * cz + sz;
        if (k >= nf) continue;
#pragma unroll
        for (int sy = 0; sy < 2; sy++) {
            const int j = 2 * cy + sy;
            if (j >= nf) continue;
#pragma unroll
            for (int sx = 0; sx < 2; sx++) {
                const int i = 2 * cx + sx;
                if (i >= nf) continue;
                const int sv[3] = {sx, sy, sz};
                // per axis the two terms of mg_prolong_pt: (slot, weight, present)
                int sl[3][2];
                T wt[3][2];
                bool in[3][2];
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    const bool odd = sv[a] == 1;
                    sl[a][0] = odd ? 1 : 0;
                    wt[a][0] = odd ? T(1) : T(0.5);
                    in[a][0] = odd ? ok[a][1] : ok[a][0];
                    sl[a][1] = 1;
                    wt[a][1] = T(0.5);
                    in[a][1] = !odd && ok[a][1];
                }
                T s = T(0);
#pragma unroll
                for (int a = 0; a < 2; a++)
#pragma unroll
                    for (int b = 0; b < 2; b++)
#pragma unroll
                        for (int d = 0; d < 2; d++) {
                            const bool present = in[0][a] && in[1][b] && in[2][d];
                            const T val = v[sl[0][a]][sl[1][b]][sl[2][d]];
                            s += present ? wt[0][a] * wt[1][b] * wt[2][d] * val : T(0);
                        }
                fine[i + (long)nf * j + (long)nf * nf * k] += s;
            }
        }
    }
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
    hipLaunchKernelGGL(k_mg_jacobi<T>, mg_grid(n), dim3(MG_BX, MG_BY), 0, s, dst, x, b, n, (T)h2, (T)omega, (T)g_low,
                       (T)g_high);
    LIRA_HIP(hipGetLastError());
}
template <typename T>
void mg_residual(T *dst, const T *x, const T *b, int n, double h2, double g_low, double g_high, hipStream_t s) {
    hipLaunchKernelGGL(k_mg_residual<T>, mg_grid(n), dim3(MG_BX, MG_BY), 0, s, dst, x, b, n, (T)h2, (T)g_low, (T)g_high);
    LIRA_HIP(hipGetLastError());
}
template <typename T>
long mg_jacobi_dot(T *dst, const T *x, const T *b, int n, double h2, double omega, double g_low, double g_high,
                   double *partials, hipStream_t s) {
    const dim3 g = mg_grid(n);
    hipLaunchKernelGGL(k_mg_jacobi_dot<T>, g, dim3(MG_BX, MG_BY), 0, s, dst, x, b, n, (T)h2, (T)omega, (T)g_low,
                       (T)g_high, partials);
    LIRA_HIP(hipGetLastError());
    return (long)g.x * g.y * g.z * MG_BY;
}
long mg_dot_blocks(int n) {
    const dim3 g = mg_grid(n);
    return (long)g.x * g.y * g.z * MG_BY;
}
template <typename T> void mg_restrict(T *dst, const T *fine, int nc, int nf, hipStream_t s) {
    if (nf % 2 == 0)  // fine rows start at even offsets: the pair loads are aligned
        hipLaunchKernelGGL(k_mg_restrict_pairs<T>, mg_grid(nc), dim3(MG_BX, MG_BY), 0, s, dst, fine, nc, nf);
    else
        hipLaunchKernelGGL(k_mg_restrict<T>, mg_grid(nc), dim3(MG_BX, MG_BY), 0, s, dst, fine, nc, nf);
    LIRA_HIP(hipGetLastError());
}
template <typename T> void mg_prolong_add(T *fine, const T *coarse, int nf, int nc, hipStream_t s) {
    hipLaunchKernelGGL(k_mg_prolong_cells<T>, mg_grid((nf + 1) / 2), dim3(MG_BX, MG_BY), 0, s, fine, coarse, nf, nc);
    LIRA_HIP(hipGetLastError());
}
template <typename T>
void mg_jacobi0(T *dst, const T *b, int n, double h2, double omega, double g_low, double g_high, hipStream_t s) {
    hipLaunchKernelGGL(k_mg_jacobi0<T>, mg_grid(n), dim3(MG_BX, MG_BY), 0, s, dst, b, n, (T)h2, (T)omega, (T)g_low,
                       (T)g_high);
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
// k_cg_update that also leaves each wave's share of the new r.r in partials[block * 4 + wave]
// (each thread updates DOT_PER_THREAD points 256 apart, so the wave sums that end the kernel are few)
constexpr int DOT_PER_THREAD = 4;
template <typename Tv, typename Tr>
__global__ __launch_bounds__(256) void k_cg_update_dot(Tv *x, const Tr *p, Tr *r, const Tr *q, const double *sc, int ia,
                                                       int ib, long N, double *partials) {
    const double alpha = sc[ia] / sc[ib];
    double d = 0.0;
#pragma unroll
    for (int u = 0; u < DOT_PER_THREAD; u++) {
        const long i = (blockIdx.x * (long)DOT_PER_THREAD + u) * 256L + threadIdx.x;
        if (i < N) {
            x[i] = (Tv)((double)x[i] + alpha * (double)p[i]);
            const Tr rn = (Tr)((double)r[i] - alpha * (double)q[i]);
            r[i] = rn;
            d += (double)rn * (double)rn;
        }
    }
    d = wave_sum(d);
    if ((threadIdx.x & 63) == 0) partials[blockIdx.x * 4L + (threadIdx.x >> 6)] = d;
}

// out[0] = the sum of a's na partials, out[1] = that of b's nb, in a fixed order: DOT_FINISH_BLOCKS blocks per sum each
// take a contiguous share of the partials into stage[], and the last block to finish (a device counter, reset by it)
// sums the stages in block order.
// each wave's share of sum a[i] b[i] into partials[block * 4 + wave] (a dot at full-grid parallelism)
template <typename T>
__global__ __launch_bounds__(256) void k_dot_partials(const T *a, const T *b, long N, double *partials) {
    double d = 0.0;
#pragma unroll
    for (int u = 0; u < DOT_PER_THREAD; u++) {
        const long i = (blockIdx.x * (long)DOT_PER_THREAD + u) * 256L + threadIdx.x;
        if (i < N) d += (double)a[i] * (double)b[i];
    }
    d = wave_sum(d);
    if ((threadIdx.x & 63) == 0) partials[blockIdx.x * 4L + (threadIdx.x >> 6)] = d;
}

constexpr int DOT_FINISH_BLOCKS = 32;
__global__ __launch_bounds__(256) void k_dot_finish(const double *a, long na, const double *b, long nb, double *out,
                                                    double *stage, unsigned *counter) {
    __shared__ double sh[4];
    __shared__ bool last;
    const int which = blockIdx.x / DOT_FINISH_BLOCKS, part = blockIdx.x % DOT_FINISH_BLOCKS;
    const double *src = which == 0 ? a : b;
    const long cnt = which == 0 ? na : nb;
    const long per = (cnt + DOT_FINISH_BLOCKS - 1) / DOT_FINISH_BLOCKS;
    const long lo = part * per, hi = lo + per < cnt ? lo + per : cnt;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    long p = lo + threadIdx.x;
    for (; p + 3 * 256 < hi; p += 4 * 256) {
#pragma unroll
        for (int u = 0; u < 4; u++) acc[u] += src[p + u * 256];
    }
    for (int u = 0; p < hi; p += 256, u++) acc[u & 3] += src[p];
    const double t = block_sum256((acc[0] + acc[1]) + (acc[2] + acc[3]), sh);
    if (threadIdx.x == 0) {
        stage[blockIdx.x] = t;
        __threadfence();
        last = atomicAdd(counter, 1u) == gridDim.x - 1u;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    if (threadIdx.x < gridDim.x / DOT_FINISH_BLOCKS) {
        double s = 0.0;
        for (int k = 0; k < DOT_FINISH_BLOCKS; k++) s += stage[threadIdx.x * DOT_FINISH_BLOCKS + k];
        out[threadIdx.x] = s;
    }
    if (threadIdx.x == 0) *counter = 0u;
}

template <typename Tv, typename Tr>
void cg_update(Tv *x, const Tr *p, Tr *r, const Tr *q, const double *sc, int ia, int ib, long N, hipStream_t s) {
    hipLaunchKernelGGL((k_cg_update<Tv, Tr>), LIRA_G1(N), 0, s, x, p, r, q, sc, ia, ib, N);
    LIRA_HIP(hipGetLastError());
}

template <typename Tv, typename Tr>
long cg_update_dot(Tv *x, const Tr *p, Tr *r, const Tr *q, const double *sc, int ia, int ib, long N, double *partials,
                   hipStream_t s) {
    const long blocks = (N + 256L * DOT_PER_THREAD - 1) / (256L * DOT_PER_THREAD);
    hipLaunchKernelGGL((k_cg_update_dot<Tv, Tr>), dim3((unsigned)blocks), dim3(256), 0, s, x, p, r, q, sc, ia, ib, N,
                       partials);
    LIRA_HIP(hipGetLastError());
    return blocks * 4;
}
size_t dot_finish_scratch_doubles() { return 2 * DOT_FINISH_BLOCKS + 1; }
void dot_finish(const double *a, long na, const double *b, long nb, double *out, double *scratch, hipStream_t s) {
    // scratch: the stage sums of the (one or two: b == nullptr for one) sums, then the counter (zero between calls)
    const int sums = b ? 2 : 1;
    hipLaunchKernelGGL(k_dot_finish, dim3(sums * DOT_FINISH_BLOCKS), dim3(256), 0, s, a, na, b, nb, out, scratch,
                       reinterpret_cast<unsigned *>(scratch + 2 * DOT_FINISH_BLOCKS));
    LIRA_HIP(hipGetLastError());
}
template <typename T> long dot_partials(const T *a, const T *b, long N, double *partials, hipStream_t s) {
    const long blocks = (N + 256L * DOT_PER_THREAD - 1) / (256L * DOT_PER_THREAD);
    hipLaunchKernelGGL(k_dot_partials<T>, dim3((unsigned)blocks), dim3(256), 0, s, a, b, N, partials);
    LIRA_HIP(hipGetLastError());
    return blocks * 4;
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
    template void multipole_moments<T>(const T *, int, double, double *, const double *, double *, hipStream_t); \
    template void poisson_rhs<T>(double *, const T *, const double *, double *, int, double, hipStream_t);     \
    template void mg_jacobi<T>(T *, const T *, const T *, int, double, double, double, double, hipStream_t);  \
    template void mg_residual<T>(T *, const T *, const T *, int, double, double, double, hipStream_t);        \
    template void mg_restrict<T>(T *, const T *, int, int, hipStream_t);                                      \
    template void mg_prolong_add<T>(T *, const T *, int, int, hipStream_t);                                   \
    template void mg_jacobi0<T>(T *, const T *, int, double, double, double, double, hipStream_t);            \
    template long mg_jacobi_dot<T>(T *, const T *, const T *, int, double, double, double, double, double *,  \
                                   hipStream_t);                                                              \
    template long cg_update_dot<T, T>(T *, const T *, T *, const T *, const double *, int, int, long, double *, \
                                      hipStream_t);                                                           \
    template long dot_partials<T>(const T *, const T *, long, double *, hipStream_t);                          \
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
