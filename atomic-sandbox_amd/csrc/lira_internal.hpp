// Internal declarations of liblira: device buffers, launch wrappers of the gfx950 kernels, and the host-side
// numerics shared by the Poisson, LOBPCG and SCF drivers. Nothing here is part of the C ABI (include/lira.h).
//
// Layout in HBM: every grid field is a dense cube n^3 in the reference's qm layout idx = i + n(j + n k)
// (qm/src/grid.rs:16-17), zero outside the box (Dirichlet). A block of nb bands is nb such fields back to back
// (band-major), so band b of a block starts at b * N.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace lira {

struct DeviceError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct NumericalError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct InvalidError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void check_hip(hipError_t e, const char *what);
#define LIRA_HIP(call) ::lira::check_hip((call), #call)

// A device allocation owned by the solver (hipMalloc'd, freed on destruction).
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    explicit DevBuf(size_t b) { alloc(b); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : ptr(o.ptr), bytes(o.bytes) { o.ptr = nullptr; o.bytes = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) { release(); ptr = o.ptr; bytes = o.bytes; o.ptr = nullptr; o.bytes = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t b);
    void release();
    template <typename T> T *as() const { return static_cast<T *>(ptr); }
};

// ------------------------------------------------------------------ reductions ----
// Deterministic two-level sums: a fixed number of workgroup partials in f64, then one fixed-order finish.
constexpr int RED_BLOCKS = 512;
constexpr int MAX_JOBS = 24;

enum RedKind : int { RED_DOT = 0, RED_SUM = 1, RED_ABSDIFF = 2, RED_FACEMAX = 3, RED_MAX_ABS = 4 };

template <typename T> struct RedJobs {
    const T *a[MAX_JOBS];
    const T *b[MAX_JOBS];
    int kind[MAX_JOBS];
    int count;
};

// out[j] = scale * reduction j over N points (FACEMAX: max |a| on the 6 faces of the n^3 cube; MAX_ABS: max |a|).
// partials: RED_BLOCKS * MAX_JOBS doubles of scratch. Results stay on the device (out).
template <typename T>
void reduce_jobs(const RedJobs<T> &jobs, int n, double scale, double *partials, double *out, hipStream_t s);

// G[i * q + j] = scale * sum_p S_i[p] * U_j[p] for i < m, j < q; S columns are segments of band-major blocks.
struct Seg {
    const void *base;
    int count;
};
constexpr int MAX_SEGS = 8;
struct Cols {
    Seg seg[MAX_SEGS];
    int nseg;
    int total() const { int t = 0; for (int i = 0; i < nseg; i++) t += seg[i].count; return t; }
};
template <typename T>
void gram(const Cols &S, const Cols &U, long N, double scale, double *partials, size_t partial_cap, double *G,
          hipStream_t s);
size_t gram_partials_needed(int m, int q, long N);

// Y_j = sum_i S_i * C[i * q + j] (C f64 on the device, m x q), accumulated in f64. Outputs j < qsplit go to
// Y + j N, the others to Y2 + (j - qsplit) N.
template <typename T>
void lincomb(const Cols &S, const double *C, int ldc, int q, T *Y, long N, hipStream_t s, int qsplit = -1,
             T *Y2 = nullptr);
// R = HX - theta X and W = R / max(diag_t + V - lam, 0.1) for nb bands in one pass
template <typename T>
void residual_precond(T *r, T *w, const T *hx, const T *x, const double *theta, const T *v, const double *lam,
                      double diag_t, int nb, long N, hipStream_t s);
// the padded shape lincomb's coefficients must have: lincomb_rows(m) rows of lincomb_ldc(q) doubles, zero beyond m x q
int lincomb_rows(int m);
int lincomb_ldc(int q);

// ------------------------------------------------------------------- stencil ----
enum StencilMode : int { ST_LAP = 0, ST_HAM = 1, ST_JAC = 2 };
struct StencilParams {
    int n, nb;
    double h;
    double scale;   // ST_LAP: out = scale * lap x
    double omega;   // ST_JAC damping
    double diag_t;  // ST_JAC: diagonal of -1/2 lap, -1.5 c0 / h^2
    // a second band segment (ST_HAM, ST_LAP): bands nb1 .. nb - 1 read x2 and write out2 (band b at (b - nb1) N);
    // nb1 < 0: every band in the first segment
    int nb1 = -1;
    const void *x2 = nullptr;
    void *out2 = nullptr;
};
// ST_HAM: out_b = -1/2 lap x_b + v x_b. ST_JAC: out_b = x_b + omega/max(diag_t + v - lam_b, 0.1) (r_b - (H - lam_b) x_b).
template <typename T>
void stencil(int mode, const StencilParams &p, const T *x, T *out, const T *v, const T *r, const double *lam,
             hipStream_t s, int variant = -1);
int stencil_variants();
bool stencil_variant_usable(int variant, int n, int dtype_size);
// true when an n^3 field of esize-byte values (plus 4 planes either side) fits the stencils' 32-bit buffer offsets:
// (n^3 + 4 n^2) esize < 2^31 (n <= 644 in f64, 812 in f32)
inline bool stencil_buffer_ok(int n, int esize) {
    const long nn = (long)n * n;
    return ((long)n * nn + 4 * nn) * esize < 0x7fffffffL;
}
const char *stencil_variant_name(int v);
void stencil_set_default_variant(int dtype_size, int variant);

// -------------------------------------------------------------- pointwise ----
template <typename T> void fill(T *y, long count, double a, hipStream_t s);
template <typename T> void axpy(T *y, const T *x, long count, double a, hipStream_t s);  // y += a x
// r_b = hx_b - theta_b x_b, theta on the device (f64)
template <typename T> void residual(T *r, const T *hx, const T *x, const double *theta, int nb, long N, hipStream_t s);
// w_b = r_b / max(diag_t + v - lam_b, 0.1)
template <typename T>
void precond_init(T *w, const T *r, const T *v, const double *lam, double diag_t, int nb, long N, hipStream_t s);
// rho = sum_b f_b psi_b^2 (f on the device, f64)
template <typename T> void density(T *rho, const T *psi, const double *f, int nb, long N, hipStream_t s);
// v_sigma = vnuc + vh + vxc_sigma
template <typename T> void sum3(T *out, const T *a, const T *b, const T *c, long N, hipStream_t s);
template <typename T> void add2(T *out, const T *a, const T *b, long N, hipStream_t s);
template <typename T> void sub2(T *out, const T *a, const T *b, long N, hipStream_t s);
template <typename T> void copy(T *out, const T *a, long N, hipStream_t s);
template <typename Td, typename Ts> void convert(Td *out, const Ts *a, long N, hipStream_t s);
// LSDA (Dirac + PW92), math in f64 (qm/src/xc.rs:120-210); with a partial core pc (nullable) on n_sigma = rho_sigma + pc/2
template <typename T>
void xc_lsda(const T *nu, const T *nd, T *e, T *vu, T *vd, long N, hipStream_t s, const T *pc = nullptr);
// V = -sum_I Z_I erf(d/(sigma sqrt2))/d (qm/src/nuclear.rs:258-273); pos/z on the device
template <typename T>
void nuclear_potential(T *out, int n, double h, const double *pos, const double *zf, int nat, double sigma,
                       hipStream_t s);
template <typename T>
void deposit_charge(T *out, int n, double h, const double *pos, const double *zf, int nat, double sigma, double scale,
                    hipStream_t s);
// the reference's LCG orbital guess (qm/src/scf.rs:626-634), generated on the device by jump-ahead
template <typename T>
void initial_orbitals(T *out, int n, double h, const double *pos, int nat, int k, uint64_t seed0, hipStream_t s);
// DIIS mix: out = max(0, sum_i c_i (rho_i + alpha r_i)) with c on the device
template <typename T>
void diis_mix(T *out, const T *const *rho, const T *const *res, const double *c, int m, double alpha, long N,
              hipStream_t s);
// HF force partials: per nucleus I and axis a, sum_p rho_p (-Z g'(d)/d) (r_p - R_I)_a (qm/src/forces.rs:189-236)
template <typename T>
void hf_forces(const T *rho, int n, double h, const double *pos, const double *zf, int nat, double sigma,
               double *partials, double *out, hipStream_t s);

// ------------------------------------------------------------ pseudised ions ----
// One nucleus as the local potential sees it (qm/src/nuclear.rs:304-311 Centre + the tables of a pseudised ion):
// the smeared charge q at width sigma (Z, sigma_AE all-electron; Z_v, sigma_LR pseudised), the V_SR and partial-core
// tables (offsets into a f64 table arena, length 0 = none) and their radii.
struct IonicCentre {
    double pos[3];
    double q, sigma, du;
    int sr_off, sr_len;
    double r_sr;
    int pc_off, pc_len;
    double pc_r;
};
constexpr int NL_MAX_PROJ = 3;  // one projector per non-local channel, l <= 2 (qm/src/nonlocal.rs:100)
constexpr int NL_MAX_ROWS = 9;  // sum of 2l+1 over the projectors of one ion
// One nucleus with projectors: its box (largest projector radius, clipped), beta rows and tables.
struct NlIon {
    double pos[3];
    double du;
    int lo[3], dim[3];
    long pts, beta_off;
    int rows, row_off, nproj;
    int l[NL_MAX_PROJ], t_off[NL_MAX_PROJ], t_len[NL_MAX_PROJ];
    double radius[NL_MAX_PROJ], e_kb[NL_MAX_PROJ];
};
template <typename T>
void ionic_fields(T *vloc, T *pc, int n, double h, const IonicCentre *c, int nat, const double *tab, hipStream_t s);
template <typename T> void nl_beta(T *beta, const NlIon &ion, int n, double h, const double *tab, hipStream_t s);
// C[(row_off + r) * nb + b] = dv sum_box beta_r x_b for every ion with projectors, nb = n1 + n2 bands (the first n1
// from x1, the rest from x2); the box sums run in chunks (nl_project_chunks(max_pts) per box), their partials in work
// (rows * nb * chunks doubles, rows = all ions' projector rows) summed per (row, band) in a fixed order by a second
// launch
int nl_project_chunks(long max_pts);
template <typename T>
void nl_project(const T *x1, int n1, const T *x2, int n2, long N, int n, const T *beta, const NlIon *ions, int nions,
                long max_pts, int rows, double dv, double *C, double *work, hipStream_t s);
// out_b += sum_r erow_r C_rb beta_r on the boxes of the listed ions (boxes must be disjoint), bands b < n1 in out1,
// the rest in out2
template <typename T>
void nl_apply(T *out1, int n1, T *out2, int n2, long N, int n, const T *beta, const NlIon *ions, const int *members,
              int nmembers, long max_pts, const double *C, const double *erow, hipStream_t s);
// out[(term * nat + I) * 3 + a], term 0 local LR, 1 local SR, 2 partial core (qm/src/forces.rs:281-333)
template <typename T>
void ionic_forces(const T *rho, const T *vu, const T *vd, int n, double h, const IonicCentre *cs, const double *tab,
                  int nat, double *partials, double *out, hipStream_t s);
// CG[((row_off + r) * 4 + q) * nb + b]: q = 0 the projection c, q = 1..3 <grad_a beta_r | psi_b>
template <typename T>
void nl_force(const T *psi, int nb, long N, int n, double h, const NlIon *ions, int nions, const double *tab,
              double *CG, hipStream_t s);
// out = max(2 a - b, 0) (warm-start density extrapolation)
template <typename T> void extrapolate(T *out, const T *a, const T *b, long N, hipStream_t s);
// y *= (*num / *den) when both are > 0 (device scalars)
template <typename T> void scale_ratio(T *y, long N, const double *num, const double *den, hipStream_t s);
// out[q] = (u + d) at pts[3q..3q+2] by trilinear interpolation (0 outside the grid)
template <typename T>
void sample_density(const T *u, const T *d, int n, double h, const double *pts, int count, double *out, hipStream_t s);
// --------------------------------------------------------------- Poisson ----
// Q_lm = dv * sum_p rho_p R_lm(r_p) for the 49 real solid harmonics l <= 6 (qm/src/poisson.rs:97-162), from the raw
// moments x^a y^b z^c (a + b + c <= 6) by separable sums; work holds moment_work_doubles(n), coef the 49 x 84 table
// moment_coefficients() writes (on the device)
constexpr int MOMENT_COEFS = 49 * 84;
void moment_coefficients(double *coef);
size_t moment_work_doubles(int n);
template <typename T>
void multipole_moments(const T *rho, int n, double h, double *work, const double *coef, double *q, hipStream_t s);
// b = 4 pi rho + L_bc (qm/src/poisson.rs:175-214) with the moments q on the device; ghost holds 24 n^2 doubles (the
// boundary potential on the 4 layers past each face)
template <typename T>
void poisson_rhs(double *b, const T *rho, const double *q, double *ghost, int n, double h, hipStream_t s);
// multigrid level operators (7-point A2 with boundary-distance ghost factors)
// the last post-smoothing sweep of the finest level with the block partials of sum b dst (returns the partial count);
// the CG update with the block partials of the new r.r; their two sums into out[0], out[1] in a fixed order
template <typename T>
long mg_jacobi_dot(T *dst, const T *x, const T *b, int n, double h2, double omega, double g_low, double g_high,
                   double *partials, hipStream_t s);
long mg_dot_blocks(int n);  // the partial count of mg_jacobi_dot on an n^3 level (one per wave)
template <typename Tv, typename Tr>
long cg_update_dot(Tv *x, const Tr *p, Tr *r, const Tr *q, const double *sc, int ia, int ib, long N, double *partials,
                   hipStream_t s);
void dot_finish(const double *a, long na, const double *b, long nb, double *out, double *scratch, hipStream_t s);
// the per-wave partials of a . b (returns their count; 4 per 256 points), finished by dot_finish
template <typename T> long dot_partials(const T *a, const T *b, long N, double *partials, hipStream_t s);
size_t dot_finish_scratch_doubles();  // dot_finish's scratch (zeroed once before the first call)
template <typename T>
void mg_jacobi(T *dst, const T *x, const T *b, int n, double h2, double omega, double g_low, double g_high,
               hipStream_t s);
template <typename T>
void mg_residual(T *dst, const T *x, const T *b, int n, double h2, double g_low, double g_high, hipStream_t s);
template <typename T> void mg_restrict(T *dst, const T *fine, int nc, int nf, hipStream_t s);
template <typename T> void mg_prolong_add(T *fine, const T *coarse, int nf, int nc, hipStream_t s);
template <typename T>
void mg_jacobi0(T *dst, const T *b, int n, double h2, double omega, double g_low, double g_high, hipStream_t s);
// the V-cycle from a small level down in one workgroup with every level in LDS (level 0 = the entry level)
constexpr int MG_MAX_LEVELS = 8;
struct MgCoarse {
    int levels;
    int n[MG_MAX_LEVELS];
    double h2[MG_MAX_LEVELS], gl[MG_MAX_LEVELS], gh[MG_MAX_LEVELS];
    int off[MG_MAX_LEVELS];  // element offset of each level's x, b, scratch block in LDS
    int pre, post, coarse;
    double omega;
};
size_t mg_coarse_bytes(const MgCoarse &a, int esize);
constexpr size_t MG_COARSE_LDS = 65536;  // the fused levels' LDS budget (the default dynamic LDS limit)
template <typename T> void mg_coarse(T *x, const T *b, const MgCoarse &a, hipStream_t s);
// device-side CG scalar updates
// x += (num/den) p ; r -= (num/den) q    (alpha = sc[ia]/sc[ib])
template <typename Tv, typename Tr>
void cg_update(Tv *x, const Tr *p, Tr *r, const Tr *q, const double *sc, int ia, int ib, long N, hipStream_t s);
// p = z + (sc[ia]/sc[ib]) p
template <typename T> void cg_direction(T *p, const T *z, const double *sc, int ia, int ib, long N, hipStream_t s);
// r = b + lap(v) computed outside; r64 -> r32 etc. via convert
// x += e (f64 += T)
template <typename T> void add_into_f64(double *x, const T *e, long N, hipStream_t s);

// ------------------------------------------------------------- free atoms ----
// The sigma-smeared spherical LSDA free-atom density of Z on the log grid r_i = e^{x0 + i dx} (cached per (Z, sigma));
// ok = false when the atom does not solve (the caller keeps its blob).
struct RadialProfile {
    double x0 = 0.0, dx = 0.0;
    std::vector<double> rho;
    bool ok = false;
    struct Orbital {
        int l;
        double eps;
        std::vector<double> R;  // the radial function u / r on the log grid
    };
    std::vector<Orbital> orbitals;  // the bound levels (n <= 4, l <= 2), ascending in energy
};
RadialProfile free_atom_profile(int z, double sigma);
// one free atom on the device: its profile at tab + off (len values), weight w, position
struct AtomProfile {
    double pos[3];
    double x0, dx, w;
    int off, len;
};
// out = R(|r - R_A|) * S_lm(r - R_A) / |r - R_A|^l (the real solid harmonic of index m in -l..l, Racah order)
template <typename T>
void atom_orbital(T *out, int n, double h, const AtomProfile &a, int l, int m, const double *tab, hipStream_t s);
// out = scale * (sum_A w_A rho_A(|r - R_A|) + sum_B q_B (2 pi)^{-3/2} e^{-|r - R_B|^2 / 2}); blobs = (x, y, z, q) rows
template <typename T>
void atom_density(T *out, int n, double h, const AtomProfile *atoms, int natoms, const double *tab,
                  const double *blobs, int nblobs, double scale, hipStream_t s);
bool free_atom_levels(int z, std::vector<double> &eps_out, double &electrons);
// ------------------------------------------------------------- host utils ----
void sym_eigen(int m, std::vector<double> &a, std::vector<double> &evals, std::vector<double> &evecs);
bool lu_solve(int m, std::vector<double> a, std::vector<double> &b);
bool cholesky(int m, std::vector<double> &a);  // in place, lower; false if not SPD
void chol_solve(int m, const std::vector<double> &l, std::vector<double> &b, int nrhs);

// generalised Rayleigh-Ritz on (A, B) m x m, lowest nk pairs; C is m x nk (row-major), returns the amplification
// alpha = max_i ||T y_i||. Throws NumericalError on rank loss.
double rayleigh_ritz(int m, const std::vector<double> &A, const std::vector<double> &B, int nk, double drop,
                     std::vector<double> &theta, std::vector<double> &C);

// ------------------------------------------------------- the Rayleigh-Ritz step on the device ----
// One LOBPCG iteration's stop test and Rayleigh-Ritz step in one workgroup, for m <= RR_DEV_MAX subspace columns.
// G: the Gram (m x q row-major, q = 2m) followed by the k squared residual norms; theta: the k Ritz values (read, and
// the moving ones rewritten); C: the lincomb's padded coefficient block (mp x ldc, [C | C_p] in the first 2 nm
// columns); record: [stop, error, head residual, tail residual, dropped P, amplification, Jacobi sweeps, 0,
// lambda[k], residual[k]]
// (error: 1 non-finite, 2 rank lost, 3 Cholesky failed, 4 a non-positive column norm).
constexpr int RR_DEV_MAX = 36;
constexpr int RR_REC_HEAD = 8;
constexpr int RR_REC_STRIDE = 128;  // >= RR_REC_HEAD + 2 RR_DEV_MAX
struct RrStepArgs {
    const double *G;
    double *theta, *C, *work, *record;
    int m, q, k, na, np, nm, locked, head, checked, iter, max_iter, is_f32, mp, ldc;
    double tol, guess_tol, tail_scale, floor, alpha_max, alpha_fraction, drop;
};
size_t rr_step_work_doubles();
void rr_step(const RrStepArgs &args, hipStream_t s);

}  // namespace lira
