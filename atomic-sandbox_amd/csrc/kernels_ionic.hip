// Pseudised-ion kernels for gfx950: the ionic local potential and partial core, the Kleinman–Bylander nonlocal
// operator on per-ion boxes, and the pseudised force terms.
//
// Reference: qm/src/nonlocal.rs:81-174 (V_NL), qm/src/nuclear.rs:290-368 (V_loc, rho_pc), qm/src/forces.rs:259-350
// (force terms), qm/src/radial/filter.rs:302-363 (uniform tables).
//
// Layout. An ion with projectors owns the bounding box of its largest projector sphere, clipped to the grid
// (lo[3], dim[3], pts = dim0*dim1*dim2 points, x fastest). Its beta_lm are stored row-major over that box
// (rows = sum_j (2 l_j + 1), row r at beta + beta_off + r * pts), exactly 0 outside each projector's radius, so a
// box sum equals the reference's sum over the sphere's point list. Box rows are contiguous in x, so reads of a
// band's values over a box coalesce. Projections C[row][band] are f64.
#include "lira_internal.hpp"

#include <cmath>

namespace lira {

// 4-point Lagrange read of a uniform even table (index -1 mirrors 1), value and derivative; 0 beyond radius
// (qm/src/radial/filter.rs:302-318, 355)
__device__ inline void table_read(const double *f, int len, double du, double radius, double d, double &v,
                                  double &dv) {
    v = 0.0;
    dv = 0.0;
    if (len <= 0 || d > radius) return;
    const double sv = d / du;
    const long i0 = (long)floor(sv) - 1;
    if (i0 + 3 >= (long)len) return;
    const double t = sv - (double)i0;
    const long j0 = i0 < 0 ? -i0 : i0;
    const double a = f[j0], b = f[i0 + 1], c = f[i0 + 2], e = f[i0 + 3];
    const double l0 = -(t - 1.0) * (t - 2.0) * (t - 3.0) / 6.0, l1 = t * (t - 2.0) * (t - 3.0) / 2.0;
    const double l2 = -t * (t - 1.0) * (t - 3.0) / 2.0, l3 = t * (t - 1.0) * (t - 2.0) / 6.0;
    const double d0 = -((t - 2.0) * (t - 3.0) + (t - 1.0) * (t - 3.0) + (t - 1.0) * (t - 2.0)) / 6.0;
    const double d1 = ((t - 2.0) * (t - 3.0) + t * (t - 3.0) + t * (t - 2.0)) / 2.0;
    const double d2 = -((t - 1.0) * (t - 3.0) + t * (t - 3.0) + t * (t - 1.0)) / 2.0;
    const double d3 = ((t - 1.0) * (t - 2.0) + t * (t - 2.0) + t * (t - 1.0)) / 6.0;
    v = a * l0 + b * l1 + c * l2 + e * l3;
    dv = (a * d0 + b * d1 + c * d2 + e * d3) / du;
}

// Racah-normalised real solid harmonics R_lm for l <= 2 in m = -l..l order, and their gradients
// (qm/src/poisson.rs lm_index order; qm/src/nonlocal.rs:63-73)
__device__ inline void solid_l(int l, double x, double y, double z, double *R) {
    const double s3 = 1.7320508075688772;
    if (l == 0) {
        R[0] = 1.0;
    } else if (l == 1) {
        R[0] = y;
        R[1] = z;
        R[2] = x;
    } else {
        R[0] = s3 * x * y;
        R[1] = s3 * y * z;
        R[2] = (2.0 * z * z - x * x - y * y) / 2.0;
        R[3] = s3 * x * z;
        R[4] = s3 / 2.0 * (x * x - y * y);
    }
}
__device__ inline void solid_grad_l(int l, double x, double y, double z, double (*G)[3]) {
    const double s3 = 1.7320508075688772;
    if (l == 0) {
        G[0][0] = G[0][1] = G[0][2] = 0.0;
    } else if (l == 1) {
        G[0][0] = 0.0, G[0][1] = 1.0, G[0][2] = 0.0;
        G[1][0] = 0.0, G[1][1] = 0.0, G[1][2] = 1.0;
        G[2][0] = 1.0, G[2][1] = 0.0, G[2][2] = 0.0;
    } else {
        G[0][0] = s3 * y, G[0][1] = s3 * x, G[0][2] = 0.0;
        G[1][0] = 0.0, G[1][1] = s3 * z, G[1][2] = s3 * y;
        G[2][0] = -x, G[2][1] = -y, G[2][2] = 2.0 * z;
        G[3][0] = s3 * z, G[3][1] = 0.0, G[3][2] = s3 * x;
        G[4][0] = s3 * x, G[4][1] = -s3 * y, G[4][2] = 0.0;
    }
}

__device__ inline double origin(int n, double h) { return -((double)n - 1.0) * h / 2.0; }

// box point p of an ion -> grid index and coordinates relative to the ion
__device__ inline long box_point(const NlIon &ion, long p, int n, double h, double &dx, double &dy, double &dz) {
    const long px = p % ion.dim[0], t = p / ion.dim[0];
    const long py = t % ion.dim[1], pz = t / ion.dim[1];
    const long i = ion.lo[0] + px, j = ion.lo[1] + py, k = ion.lo[2] + pz;
    const double o = origin(n, h);
    dx = o + (double)i * h - ion.pos[0];
    dy = o + (double)j * h - ion.pos[1];
    dz = o + (double)k * h - ion.pos[2];
    return i + (long)n * (j + (long)n * k);
}

// ---------------------------------------------------------------------------------- local potential ----
// V_loc = sum_I -q_I erf(d/(sigma_I sqrt2))/d (on site -q sqrt(2/pi)/sigma) + sum_pseudised V_SR(d); rho_pc likewise
// (qm/src/nuclear.rs:290-302, 342-368). One thread per grid point, the centres in input order.
template <typename T>
__global__ void k_ionic_fields(T *vloc, T *pc, int n, double h, const IonicCentre *c, int nat, const double *tab) {
    const long N = (long)n * n * n;
    const long i = blockIdx.x * 256L + threadIdx.x;
    if (i >= N) return;
    const long nn = (long)n * n;
    const long k = i / nn, rem = i - k * nn, j = rem / n, ii = rem - j * n;
    const double o = origin(n, h);
    const double x = o + (double)ii * h, y = o + (double)j * h, z = o + (double)k * h;
    double acc = 0.0;
    for (int a = 0; a < nat; a++) {
        const double dx = x - c[a].pos[0], dy = y - c[a].pos[1], dz = z - c[a].pos[2];
        const double d = sqrt(dx * dx + dy * dy + dz * dz);
        acc += d < 1e-10 ? -c[a].q * sqrt(2.0 / M_PI) / c[a].sigma : -c[a].q * erf(d / (c[a].sigma * sqrt(2.0))) / d;
    }
    double core = 0.0;
    for (int a = 0; a < nat; a++) {
        if (c[a].sr_len == 0 && c[a].pc_len == 0) continue;
        const double dx = x - c[a].pos[0], dy = y - c[a].pos[1], dz = z - c[a].pos[2];
        const double d = sqrt(dx * dx + dy * dy + dz * dz);
        double v, dv;
        if (c[a].sr_len > 0 && d <= c[a].r_sr) {
            table_read(tab + c[a].sr_off, c[a].sr_len, c[a].du, c[a].r_sr, d, v, dv);
            acc += v;
        }
        if (c[a].pc_len > 0 && d <= c[a].pc_r) {
            table_read(tab + c[a].pc_off, c[a].pc_len, c[a].du, c[a].pc_r, d, v, dv);
            core += v;
        }
    }
    vloc[i] = (T)acc;
    if (pc) pc[i] = (T)core;
}

template <typename T>
void ionic_fields(T *vloc, T *pc, int n, double h, const IonicCentre *c, int nat, const double *tab, hipStream_t s) {
    const long N = (long)n * n * n;
    hipLaunchKernelGGL(k_ionic_fields<T>, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, vloc, pc, n, h, c, nat,
                       tab);
    LIRA_HIP(hipGetLastError());
}

// ------------------------------------------------------------------------------------- free atoms ----
// rho(p) = sum_A w_A rho_A(|r_p - R_A|) (log-linear interpolation of each atom's radial profile, its innermost value
// inside the grid, 0 beyond it) + sum_B q_B (2 pi)^{-3/2} e^{-d^2/2} (sigma = 1 blobs) (qm/src/nuclear.rs:153-226)
template <typename T>
__global__ void k_atom_density(T *out, int n, double h, const AtomProfile *atoms, int natoms, const double *tab,
                               const double *blobs, int nblobs, double scale) {
    const long N = (long)n * n * n;
    const long i = blockIdx.x * 256L + threadIdx.x;
    if (i >= N) return;
    const long nn = (long)n * n;
    const long k = i / nn, rem = i - k * nn, j = rem / n, ii = rem - j * n;
    const double o = origin(n, h);
    const double x = o + (double)ii * h, y = o + (double)j * h, z = o + (double)k * h;
    double acc = 0.0;
    for (int a = 0; a < natoms; a++) {
        const AtomProfile &A = atoms[a];
        const double dx = x - A.pos[0], dy = y - A.pos[1], dz = z - A.pos[2];
        const double d = sqrt(dx * dx + dy * dy + dz * dz);
        const double *rho = tab + A.off;
        double v;
        const double r0 = exp(A.x0), rl = exp(A.x0 + (A.len - 1) * A.dx);
        if (d <= r0) {
            v = rho[0];
        } else if (d >= rl) {
            v = 0.0;
        } else {
            const double s = (log(d) - A.x0) / A.dx;
            int q = (int)floor(s);
            q = q < 0 ? 0 : (q > A.len - 2 ? A.len - 2 : q);
            const double t = s - (double)q;
            v = rho[q] + t * (rho[q + 1] - rho[q]);
        }
        acc += A.w * v;
    }
    const double norm = 1.0 / pow(2.0 * M_PI, 1.5);
    for (int b = 0; b < nblobs; b++) {
        const double dx = x - blobs[4 * b], dy = y - blobs[4 * b + 1], dz = z - blobs[4 * b + 2];
        acc += blobs[4 * b + 3] * norm * exp(-(dx * dx + dy * dy + dz * dz) / 2.0);
    }
    out[i] = (T)(acc * scale);
}

template <typename T>
void atom_density(T *out, int n, double h, const AtomProfile *atoms, int natoms, const double *tab,
                  const double *blobs, int nblobs, double scale, hipStream_t s) {
    const long N = (long)n * n * n;
    hipLaunchKernelGGL(k_atom_density<T>, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, out, n, h, atoms, natoms,
                       tab, blobs, nblobs, scale);
    LIRA_HIP(hipGetLastError());
}

// one free-atom orbital sampled on the grid: R(d) (log-linear in ln d, the innermost value inside the grid, 0 beyond
// it) times the real solid harmonic of (l, m) divided by d^l (an initial LOBPCG column; its norm does not matter)
template <typename T>
__global__ void k_atom_orbital(T *out, int n, double h, AtomProfile A, int l, int m, const double *tab) {
    const long N = (long)n * n * n;
    const long i = blockIdx.x * 256L + threadIdx.x;
    if (i >= N) return;
    const long nn = (long)n * n;
    const long k = i / nn, rem = i - k * nn, j = rem / n, ii = rem - j * n;
    const double o = origin(n, h);
    const double dx = o + (double)ii * h - A.pos[0], dy = o + (double)j * h - A.pos[1], dz = o + (double)k * h - A.pos[2];
    const double d = sqrt(dx * dx + dy * dy + dz * dz);
    const double *R = tab + A.off;
    const double r0 = exp(A.x0), rl = exp(A.x0 + (A.len - 1) * A.dx);
    double v;
    if (d <= r0) {
        v = R[0];
    } else if (d >= rl) {
        v = 0.0;
    } else {
        const double s = (log(d) - A.x0) / A.dx;
        int q = (int)floor(s);
        q = q < 0 ? 0 : (q > A.len - 2 ? A.len - 2 : q);
        const double t = s - (double)q;
        v = R[q] + t * (R[q + 1] - R[q]);
    }
    double ang = 1.0;
    if (l > 0) {
        double Y[5];
        solid_l(l, dx, dy, dz, Y);
        const double dl = l == 1 ? d : d * d;
        ang = d > 1e-12 ? Y[m + l] / dl : 0.0;
    }
    out[i] = (T)(v * ang);
}

template <typename T>
void atom_orbital(T *out, int n, double h, const AtomProfile &a, int l, int m, const double *tab, hipStream_t s) {
    const long N = (long)n * n * n;
    hipLaunchKernelGGL(k_atom_orbital<T>, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, out, n, h, a, l, m, tab);
    LIRA_HIP(hipGetLastError());
}

// --------------------------------------------------------------------------------------- projectors ----
// beta_lm = t_j(d) sqrt((2l+1)/4pi) R_lm(d) on the ion's box (qm/src/nonlocal.rs:94-128)
template <typename T> __global__ void k_nl_beta(T *beta, NlIon ion, int n, double h, const double *tab) {
    const long p = blockIdx.x * 256L + threadIdx.x;
    if (p >= ion.pts) return;
    double dx, dy, dz;
    (void)box_point(ion, p, n, h, dx, dy, dz);
    const double d = sqrt(dx * dx + dy * dy + dz * dz);
    int row = 0;
    for (int j = 0; j < ion.nproj; j++) {
        const int l = ion.l[j];
        double t, dt, R[5];
        table_read(tab + ion.t_off[j], ion.t_len[j], ion.du, ion.radius[j], d, t, dt);
        const double tn = t * sqrt((double)(2 * l + 1) / (4.0 * M_PI));
        solid_l(l, dx, dy, dz, R);
        for (int m = 0; m < 2 * l + 1; m++) beta[ion.beta_off + (long)(row + m) * ion.pts + p] = (T)(tn * R[m]);
        row += 2 * l + 1;
    }
}

template <typename T> void nl_beta(T *beta, const NlIon &ion, int n, double h, const double *tab, hipStream_t s) {
    if (ion.pts <= 0) return;
    hipLaunchKernelGGL(k_nl_beta<T>, dim3((unsigned)((ion.pts + 255) / 256)), dim3(256), 0, s, beta, ion, n, h, tab);
    LIRA_HIP(hipGetLastError());
}

// deterministic sum of R values over the 256 lanes of a block: wave shuffles, then the 4 waves in order
template <int RMAX>
__device__ inline void block_sum_rows(double *acc, int rows, double (*lds)[RMAX]) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int r = 0; r < RMAX; r++) {
        if (r < rows) {
            double v = acc[r];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
            if (lane == 0) lds[wave][r] = v;
        }
    }
    __syncthreads();
}

// Projections over chunks of each ion's box. Block (chunk, ion, band group) writes the partial sums of its chunk for
// NL_BANDS bands, P[((row_off + r) * nb + b) * chunks + chunk] = sum over the chunk's points of beta_r(p) x_b(p) (a
// chunk is NL_CHUNK consecutive box points, 8 per thread; chunks past an ion's box write zeros); k_nl_finish then sums
// each (row, band)'s chunks in a fixed order into C[(row_off + r) * nb + b] = dv * sum. Bands b < n1 read x1, the
// others x2 (two blocks in one launch). (Folding the finish into the last chunk block to arrive costs more: each
// block would pay a device-scope release fence, a write-back of its XCD's dirty L2 lines.)
constexpr int NL_CHUNK = 2048, NL_BANDS = 4;
template <typename T>
__global__ __launch_bounds__(256) void k_nl_project(const T *x1, int n1, const T *x2, int nb, long N, int n,
                                                    const T *beta, const NlIon *ions, int chunks, double *P) {
    __shared__ double lds[4][NL_MAX_ROWS];
    const NlIon &ion = ions[blockIdx.y];
    const int b0 = blockIdx.z * NL_BANDS, chunk = blockIdx.x;
    const int bands = nb - b0 < NL_BANDS ? nb - b0 : NL_BANDS;
    const int rows = ion.rows;
    const int pts = (int)ion.pts, dx = ion.dim[0], dy = ion.dim[1];
    const T *xb[NL_BANDS];
#pragma unroll
    for (int j = 0; j < NL_BANDS; j++) {
        const int b = b0 + (j < bands ? j : 0);
        xb[j] = b < n1 ? x1 + (long)b * N : x2 + (long)(b - n1) * N;
    }
    const T *bb = beta + ion.beta_off;
    double acc[NL_BANDS][NL_MAX_ROWS];
#pragma unroll
    for (int j = 0; j < NL_BANDS; j++)
#pragma unroll
        for (int r = 0; r < NL_MAX_ROWS; r++) acc[j][r] = 0.0;
    const int p0 = chunk * NL_CHUNK + (int)threadIdx.x;
    // points past the box read the box's last point with weight 0: the loads carry no branch, so the unrolled
    // iterations issue them together
#pragma unroll 4
    for (int k = 0; k < (pts > 0 ? NL_CHUNK / 256 : 0); k++) {
        const int p = p0 + k * 256;
        const bool in = p < pts;
        const int q = in ? p : pts - 1;
        const double w = in ? 1.0 : 0.0;
        const int px = q % dx, t = q / dx;
        const int py = t % dy, pz = t / dy;
        const long g = (ion.lo[0] + px) + (long)n * ((ion.lo[1] + py) + (long)n * (ion.lo[2] + pz));
        double xv[NL_BANDS];
#pragma unroll
        for (int j = 0; j < NL_BANDS; j++) xv[j] = w * (double)xb[j][g];
#pragma unroll
        for (int r = 0; r < NL_MAX_ROWS; r++)
            if (r < rows) {
                const double bv = (double)bb[(long)r * pts + q];
#pragma unroll
                for (int j = 0; j < NL_BANDS; j++) acc[j][r] += bv * xv[j];
            }
    }
#pragma unroll
    for (int j = 0; j < NL_BANDS; j++) {
        if (j < bands) {  // uniform over the block
            block_sum_rows<NL_MAX_ROWS>(acc[j], rows, lds);
            if ((int)threadIdx.x < rows) {
                const int r = threadIdx.x;
                P[((long)(ion.row_off + r) * nb + b0 + j) * chunks + chunk] =
                    ((lds[0][r] + lds[1][r]) + lds[2][r]) + lds[3][r];
            }
            __syncthreads();
        }
    }
}

// C[pair] = dv * the sum of P[pair * chunks + c] over the chunks, pair = row * nb + band, one wave per pair: lane l
// sums chunks l, l + 64, ... in order, then the wave's fixed shuffle tree
__global__ __launch_bounds__(256) void k_nl_finish(const double *P, int chunks, long pairs, double dv, double *C) {
    const int lane = threadIdx.x & 63;
    const long pair = blockIdx.x * 4L + (threadIdx.x >> 6);
    if (pair >= pairs) return;  // uniform over the wave
    const double *pp = P + pair * chunks;
    double v = 0.0;
    for (int c = lane; c < chunks; c += 64) v += pp[c];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if (lane == 0) C[pair] = dv * v;
}

// out_b(p) += sum_r E_r C[r][b] beta_r(p) over the boxes of the ions of one colour (disjoint boxes: no two threads
// touch one point); blockIdx.y picks the ion of the colour, blockIdx.z the band (b < n1 in out1, the rest in out2)
template <typename T>
__global__ __launch_bounds__(256) void k_nl_apply(T *out1, int n1, T *out2, int nb, long N, int n, const T *beta,
                                                  const NlIon *ions, const int *members, const double *C,
                                                  const double *erow) {
    const NlIon &ion = ions[members[blockIdx.y]];
    const long p = blockIdx.x * 256L + threadIdx.x;
    if (p >= ion.pts) return;
    const int b = blockIdx.z;
    const long px = p % ion.dim[0], t = p / ion.dim[0];
    const long py = t % ion.dim[1], pz = t / ion.dim[1];
    const long g = (ion.lo[0] + px) + (long)n * ((ion.lo[1] + py) + (long)n * (ion.lo[2] + pz));
    double acc = 0.0;
    for (int r = 0; r < ion.rows; r++) {
        const long row = ion.row_off + r;
        acc += erow[row] * C[row * nb + b] * (double)beta[ion.beta_off + (long)r * ion.pts + p];
    }
    T *ob = b < n1 ? out1 + (long)b * N : out2 + (long)(b - n1) * N;
    ob[g] = (T)((double)ob[g] + acc);
}

int nl_project_chunks(long max_pts) { return (int)std::max(1L, (max_pts + NL_CHUNK - 1) / NL_CHUNK); }

template <typename T>
void nl_project(const T *x1, int n1, const T *x2, int n2, long N, int n, const T *beta, const NlIon *ions, int nions,
                long max_pts, int rows, double dv, double *C, double *work, hipStream_t s) {
    const int nb = n1 + n2;
    if (nions <= 0 || nb <= 0) return;
    const int chunks = nl_project_chunks(max_pts);
    hipLaunchKernelGGL(k_nl_project<T>, dim3(chunks, nions, (nb + NL_BANDS - 1) / NL_BANDS), dim3(256), 0, s, x1, n1,
                       x2, nb, N, n, beta, ions, chunks, work);
    LIRA_HIP(hipGetLastError());
    const long pairs = (long)rows * nb;
    hipLaunchKernelGGL(k_nl_finish, dim3((unsigned)((pairs + 3) / 4)), dim3(256), 0, s, work, chunks, pairs, dv, C);
    LIRA_HIP(hipGetLastError());
}

template <typename T>
void nl_apply(T *out1, int n1, T *out2, int n2, long N, int n, const T *beta, const NlIon *ions, const int *members,
              int nmembers, long max_pts, const double *C, const double *erow, hipStream_t s) {
    const int nb = n1 + n2;
    if (nmembers <= 0 || nb <= 0 || max_pts <= 0) return;
    hipLaunchKernelGGL(k_nl_apply<T>, dim3((unsigned)((max_pts + 255) / 256), nmembers, nb), dim3(256), 0, s, out1, n1,
                       out2, nb, N, n, beta, ions, members, C, erow);
    LIRA_HIP(hipGetLastError());
}

// ------------------------------------------------------------------------------------------- forces ----
__device__ inline double block_sum256_ionic(double v, double *sh) {
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

// Radial force partials (qm/src/forces.rs:232-246, 281-333): term 0 = local LR over every point (weight rho,
// dV_LR/dd of the smeared erf with the centre's q, sigma), term 1 = local SR over the V_SR ball (weight rho), term 2 =
// partial core over its ball (weight (v_xc_up + v_xc_dn)/2). partials[((term * nat + I) * 3 + a) * RED_BLOCKS + blk].
template <typename T>
__global__ __launch_bounds__(256) void k_ionic_forces(const T *rho, const T *vu, const T *vd, int n, double h,
                                                      const IonicCentre *cs, const double *tab, int nat,
                                                      double *partials) {
    __shared__ double sh[256];
    const int I = blockIdx.y, term = blockIdx.z;
    const IonicCentre &c = cs[I];
    const double o = origin(n, h);
    long lo[3] = {0, 0, 0}, dim[3] = {n, n, n};
    double radius = 0.0;
    bool active = true;
    if (term > 0) {
        const int len = term == 1 ? c.sr_len : c.pc_len;
        radius = term == 1 ? c.r_sr : c.pc_r;
        active = len > 0;
        for (int a = 0; a < 3 && active; a++) {
            double l = floor((c.pos[a] - radius - o) / h), u = ceil((c.pos[a] + radius - o) / h);
            l = l < 0.0 ? 0.0 : l;
            u = u > (double)(n - 1) ? (double)(n - 1) : u;
            if (!(u >= l)) active = false;
            lo[a] = (long)l;
            dim[a] = (long)u - (long)l + 1;
        }
    }
    double f0 = 0.0, f1 = 0.0, f2 = 0.0;
    if (active) {
        const long pts = dim[0] * dim[1] * dim[2];
        const double a2 = c.sigma * sqrt(2.0);
        for (long p = blockIdx.x * 256L + threadIdx.x; p < pts; p += (long)gridDim.x * 256) {
            const long px = p % dim[0], t = p / dim[0];
            const long py = t % dim[1], pz = t / dim[1];
            const long i = lo[0] + px, j = lo[1] + py, k = lo[2] + pz;
            const long g = i + (long)n * (j + (long)n * k);
            const double dx = o + (double)i * h - c.pos[0], dy = o + (double)j * h - c.pos[1],
                         dz = o + (double)k * h - c.pos[2];
            const double d = sqrt(dx * dx + dy * dy + dz * dz);
            if (d < 1e-8) continue;
            double dvdd, w;
            if (term == 0) {
                const double u = d / a2;
                dvdd = -c.q * ((2.0 / sqrt(M_PI)) * u * exp(-u * u) - erf(u)) / (d * d);
                w = (double)rho[g];
            } else {
                if (d > radius) continue;
                double v;
                table_read(tab + (term == 1 ? c.sr_off : c.pc_off), term == 1 ? c.sr_len : c.pc_len, c.du, radius, d,
                           v, dvdd);
                w = term == 1 ? (double)rho[g] : 0.5 * ((double)vu[g] + (double)vd[g]);
            }
            const double sc = w * dvdd / d;
            f0 += sc * dx;
            f1 += sc * dy;
            f2 += sc * dz;
        }
    }
    const double r0 = block_sum256_ionic(f0, sh), r1 = block_sum256_ionic(f1, sh), r2 = block_sum256_ionic(f2, sh);
    if (threadIdx.x == 0) {
        const long base = ((long)term * nat + I) * 3;
        partials[(base + 0) * RED_BLOCKS + blockIdx.x] = r0;
        partials[(base + 1) * RED_BLOCKS + blockIdx.x] = r1;
        partials[(base + 2) * RED_BLOCKS + blockIdx.x] = r2;
    }
}

__global__ __launch_bounds__(256) void k_ionic_sum_partials(const double *partials, double *out, double scale) {
    __shared__ double sh[256];
    double acc = 0.0;
    for (int p = threadIdx.x; p < RED_BLOCKS; p += 256) acc += partials[(long)blockIdx.x * RED_BLOCKS + p];
    const double r = block_sum256_ionic(acc, sh);
    if (threadIdx.x == 0) out[blockIdx.x] = r * scale;
}

template <typename T>
void ionic_forces(const T *rho, const T *vu, const T *vd, int n, double h, const IonicCentre *cs, const double *tab,
                  int nat, double *partials, double *out, hipStream_t s) {
    if (nat <= 0) return;
    hipLaunchKernelGGL(k_ionic_forces<T>, dim3(RED_BLOCKS, nat, 3), dim3(256), 0, s, rho, vu, vd, n, h, cs, tab, nat,
                       partials);
    LIRA_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_ionic_sum_partials, dim3(9 * nat), dim3(256), 0, s, partials, out, h * h * h);
    LIRA_HIP(hipGetLastError());
}

// Nonlocal force projections (qm/src/forces.rs:299-318): per (ion, band, row r) c_r = dv sum beta_r psi and
// g_{r,a} = dv sum (d beta_r / d r_a) psi, with grad beta = t'(d) N R_lm d/|d| + t(d) N grad R_lm computed on the fly.
// CG[((row_off + r) * 4 + q) * nb + b], q = 0 for c, 1..3 for g.
template <typename T>
__global__ __launch_bounds__(256) void k_nl_force(const T *psi, int nb, long N, int n, double h, const NlIon *ions,
                                                  const double *tab, double *CG) {
    __shared__ double lds[4][4];
    const NlIon &ion = ions[blockIdx.x];
    const int b = blockIdx.y, row = blockIdx.z;
    if (row >= ion.rows) return;  // uniform over the block
    // the projector j and magnetic index m of this row
    int j = 0, base = 0;
    while (base + 2 * ion.l[j] + 1 <= row) {
        base += 2 * ion.l[j] + 1;
        j++;
    }
    const int l = ion.l[j], m = row - base;
    const double norm = sqrt((double)(2 * l + 1) / (4.0 * M_PI));
    const T *pb = psi + (long)b * N;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (long p = threadIdx.x; p < ion.pts; p += 256) {
        double dx, dy, dz;
        const long g = box_point(ion, p, n, h, dx, dy, dz);
        const double d = sqrt(dx * dx + dy * dy + dz * dz);
        const double pv = (double)pb[g];
        double t, dt, R[5], G[5][3];
        table_read(tab + ion.t_off[j], ion.t_len[j], ion.du, ion.radius[j], d, t, dt);
        const double tn = t * norm;
        solid_l(l, dx, dy, dz, R);
        solid_grad_l(l, dx, dy, dz, G);
        const double radial = d > 1e-12 ? dt * norm * R[m] / d : 0.0;
        acc[0] += tn * R[m] * pv;
        acc[1] += (radial * dx + tn * G[m][0]) * pv;
        acc[2] += (radial * dy + tn * G[m][1]) * pv;
        acc[3] += (radial * dz + tn * G[m][2]) * pv;
    }
    block_sum_rows<4>(acc, 4, lds);
    if (threadIdx.x < 4) {
        const int q = threadIdx.x;
        CG[((long)(ion.row_off + row) * 4 + q) * nb + b] =
            h * h * h * (((lds[0][q] + lds[1][q]) + lds[2][q]) + lds[3][q]);
    }
}

template <typename T>
void nl_force(const T *psi, int nb, long N, int n, double h, const NlIon *ions, int nions, const double *tab,
              double *CG, hipStream_t s) {
    if (nions <= 0 || nb <= 0) return;
    hipLaunchKernelGGL(k_nl_force<T>, dim3(nions, nb, NL_MAX_ROWS), dim3(256), 0, s, psi, nb, N, n, h, ions, tab, CG);
    LIRA_HIP(hipGetLastError());
}

#define LIRA_IONIC_INST(T)                                                                                            \
    template void atom_orbital<T>(T *, int, double, const AtomProfile &, int, int, const double *, hipStream_t);     \
    template void atom_density<T>(T *, int, double, const AtomProfile *, int, const double *, const double *, int,   \
                                  double, hipStream_t);                                                               \
    template void ionic_fields<T>(T *, T *, int, double, const IonicCentre *, int, const double *, hipStream_t);      \
    template void nl_beta<T>(T *, const NlIon &, int, double, const double *, hipStream_t);                          \
    template void nl_project<T>(const T *, int, const T *, int, long, int, const T *, const NlIon *, int, long, int,  \
                                double, double *, double *, hipStream_t);                                             \
    template void nl_apply<T>(T *, int, T *, int, long, int, const T *, const NlIon *, const int *, int, long,          \
                              const double *, const double *, hipStream_t);                                           \
    template void ionic_forces<T>(const T *, const T *, const T *, int, double, const IonicCentre *, const double *, \
                                  int, double *, double *, hipStream_t);                                              \
    template void nl_force<T>(const T *, int, long, int, double, const NlIon *, int, const double *, double *,       \
                              hipStream_t);
LIRA_IONIC_INST(float)
LIRA_IONIC_INST(double)

}  // namespace lira
