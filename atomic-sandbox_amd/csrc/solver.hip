// liblira host driver: the Poisson MG-PCG solve, LOBPCG, the Kohn–Sham SCF loop with DIIS and Fermi–Dirac
// occupations, energies and Hellmann–Feynman forces, and the C ABI of include/lira.h.
//
// Precision: Solver<double> keeps every grid field in f64 (LIRA_FP64); Solver<float> keeps the grid fields in f32
// (stencil applies, V-cycles, block products) while V_H, the Poisson right-hand side and residual corrections, every
// reduction, the Rayleigh–Ritz and Cholesky steps, DIIS coefficients and the energy sums are f64 (LIRA_MIXED).
#include <algorithm>
#include <array>
#include <mutex>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <numeric>
#include <thread>
#include <type_traits>

#include "lira.h"
#include "lira_internal.hpp"
#include "provider_policy.hpp"

namespace lira {

static const double COEFF0 = -205.0 / 72.0;
static const double COEFF[5] = {-205.0 / 72.0, 8.0 / 5.0, -1.0 / 5.0, 8.0 / 315.0, -1.0 / 560.0};

// qm/src/eigen.rs:218-222: omega = 4 / (3 rho), rho = symbol(pi) / c0
// stencil variants tuned per (element size, n, bands, mode) in this process
static std::mutex g_tune_mu;
static std::map<std::array<int, 4>, int> g_tuned;

static double jacobi_omega() {
    const double symbol_pi = COEFF[0] + 2.0 * (-COEFF[1] + COEFF[2] - COEFF[3] + COEFF[4]);
    return 4.0 / (3.0 * (symbol_pi / COEFF[0]));
}

// f32 residual floor of ||(H - lambda) x|| for ||x|| = 1 (crates/qm-gpu/src/eigen.rs:313-318): the RMS rounding of a
// stored f32 vector (2.4e-8) filtered by the kinetic symbol.
static double residual_floor_f32(double h) {
    const double s = 1.0 / (2.0 * h * h);
    const double mean = 3.0 * std::fabs(COEFF[0]) * s;
    double c2 = 0.0;
    for (int i = 1; i < 5; i++) c2 += COEFF[i] * COEFF[i];
    const double var = 3.0 * 2.0 * c2 * s * s;
    return 2.4e-8 * std::sqrt(mean * mean + var);
}

using Clock = std::chrono::steady_clock;
// LIRA_TRACE=1 prints one line per SCF step and per eigen solve to stderr (the reference's SCF_TRACE / LOBPCG_TRACE)
static bool trace_on() {
    static const bool on = [] {
        const char *e = std::getenv("LIRA_TRACE");
        return e && *e && *e != '0';
    }();
    return on;
}
static double secs(Clock::time_point a) { return std::chrono::duration<double>(Clock::now() - a).count(); }

// Multigrid constants (crates/qm-gpu/src/multigrid.rs:70-78): V(2,2), 12 sweeps on the coarsest level (n <= 2),
// omega = 6/7 for the 7-point operator.
constexpr int MG_PRE = 2, MG_POST = 2, MG_COARSE = 12, MG_COARSEST_MAX = 2;
constexpr double MG_OMEGA = 6.0 / 7.0;
// LOBPCG constants (crates/qm-gpu/src/eigen.rs:231-262)
constexpr double RR_DROP = 1e-6;
constexpr double ALPHA_MAX = 2.0;
constexpr double ALPHA_RESIDUAL_FRACTION = 0.1;
// SCF constants (crates/qm-gpu/src/scf.rs:357-386; crates/qm/src/scf.rs:368-704)
constexpr double EIG_TOL_FLOOR_FACTOR = 3.0;
constexpr double EIG_TOL_FRACTION = 0.0015;
constexpr double EIG_TOL_MAX = 1e-3;
constexpr double DIIS_REG = 1e-12;
constexpr int MAX_WITNESS_GROWTH = 32;
constexpr double WITNESS_GAP = 0.1, WITNESS_TAIL_KT = 36.0, WITNESS_GUARD_TOL = 1e-3, WITNESS_RESIDUAL_FACTOR = 1.5;
constexpr double STALL_LEVEL_PER_ELECTRON = 1e-3;
// a warm guess within this factor of the step's eigen tolerance is kept without an iteration
// (EIG_GUESS_FACTOR, crates/qm-gpu/src/scf.rs; crates/qm-gpu/src/eigen.rs:323-326)
constexpr double EIG_GUESS_FACTOR = 1.5;
// a guess column within sqrt(GUESS_DEP) of the span of the earlier ones is re-seeded (crates/qm-gpu/src/eigen.rs:258)
constexpr double GUESS_DEP = 1e-4;
constexpr int MAX_RESEED_ROUNDS = 8;
// the step-accuracy witness tail on a loosened step (crates/qm/src/occupations.rs:174-177, 319-326)
constexpr double STEP_TAIL_C = 1e-3, STEP_TAIL_MIN_KT = 6.0;
// a level with min(f, 1 - f) at or above this is partly occupied (crates/qm/src/occupations.rs:173)
constexpr double PARTIAL_OCCUPATION_MIN = 1e-6;
// the f32 solve's stop on a step with a partly occupied level (GPU_PARTIAL_TOL, crates/qm/src/occupations.rs:180)
constexpr double GPU_PARTIAL_TOL = 1e-5;
// density on a box face above this (1/Bohr^3) warns that the box is too small (crates/qm/src/scf.rs:577)
constexpr double BOUNDARY_DENSITY_WARN = 1e-5;
// display solves: guard levels above the shown ones and their tolerance (WITNESS_GAP / 100), and the level-matching
// tolerance tau = max(100 tol, 1e-4) Ha (crates/qm/src/orbitals.rs:368-415; crates/qm-gpu/src/harris.rs)
constexpr int UNOCC_GUARD = 2;
constexpr double GUARD_TOL = 1e-3;
constexpr double TAU_PER_TOL = 100.0, TAU_MIN = 1e-4;
// a display solve whose occupied levels disagree with the SCF's: reported, the converged state kept
struct DisplayRejected : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// the witness vectors per spin a solve allocates at its start (WITNESS_GROWTH_COUNTED, crates/providers/src/ks.rs:375)
constexpr int WITNESS_GROWTH_COUNTED = 3;

// Device bytes one solve holds on an n^3 grid (the provider's footprint model, crates/providers/src/ks.rs:360-400,
// restated for this library's buffers): per spin 10 block fields (X, HX, R, W, HW, P, HP, X', P', scratch) per column
// for n_sigma + witness + 1 columns (the buffer), 12 SCF fields, 4 per DIIS slot (history + 1), 4 for density
// extrapolation, 7 of the Poisson solve and about 4/7 of a field for its coarse levels, all in precision esize; 5 f64
// fields (V_H, the f64 density copy, the Poisson right-hand side, residual and scratch); 1.1 for the small buffers.
static unsigned long long solve_footprint(int n, int esize, int n_up, int n_dn, int witness, int diis_history) {
    const double N = (double)n * n * n;
    const int h1 = std::max(diis_history, 1) + 1;
    const double cols = (n_up > 0 ? n_up + witness + 1 : 1) + (n_dn > 0 ? n_dn + witness + 1 : 1);
    const double t_fields = 10.0 * cols + 12.0 + 4.0 * h1 + 4.0 + 7.0 + 4.0 / 7.0;
    const double boundary = 8.0 * (31.0 * n * n + 28.0 * n);  // the moments' row sums and the ghost layers
    return (unsigned long long)(1.1 * (N * (esize * t_fields + 8.0 * 5.0) + boundary));
}

// ------------------------------------------------------------------------------------- occupations ----
static double occupation(double x) {
    if (x >= 0.0) {
        const double e = std::exp(-x);
        return e / (1.0 + e);
    }
    return 1.0 / (1.0 + std::exp(x));
}
static double neumaier(const std::vector<double> &v) {
    double sum = 0.0, comp = 0.0;
    for (double x : v) {
        const double t = sum + x;
        if (std::fabs(sum) >= std::fabs(x)) comp += (sum - t) + x;
        else comp += (x - t) + sum;
        sum = t;
    }
    return sum + comp;
}
static double fd_entropy(double x) {
    const double a = std::fabs(x);
    if (std::isinf(a)) return 0.0;
    return std::log1p(std::exp(-a)) + a / (1.0 + std::exp(a));
}
struct Fd {
    std::vector<double> f;
    double mu = NAN, entropy = 0.0;
};
// qm/src/occupations.rs:244-270: mu by two tie-directed bisections, reported at their midpoint
static Fd fermi_dirac(const std::vector<double> &ev, int n, double kt) {
    Fd r;
    const int m = (int)ev.size();
    r.f.assign(m, 0.0);
    if (n == 0) return r;
    if (m < n + 1) throw NumericalError("Fermi-Dirac needs at least n + 1 levels");
    for (double e : ev)
        if (!std::isfinite(e)) throw NumericalError("non-finite eigenvalue in the Fermi-Dirac occupations");
    if (!(std::isfinite(kt) && kt > 0.0)) throw NumericalError("Fermi-Dirac kT is not finite and > 0");
    auto total = [&](double mu) {
        std::vector<double> f(m);
        for (int i = 0; i < m; i++) f[i] = occupation((ev[i] - mu) / kt);
        return neumaier(f);
    };
    const double nf = n;
    double lo = *std::min_element(ev.begin(), ev.end()) - 50.0 * kt;
    double hi = *std::max_element(ev.begin(), ev.end()) + 50.0 * kt;
    if (!(total(lo) < nf && total(hi) > nf)) throw NumericalError("the Fermi-Dirac mu bracket does not bracket n");
    auto bisect = [&](double a, double b, bool ge) {
        for (int hv = 0; hv < 2048; hv++) {
            const double mid = 0.5 * a + 0.5 * b;
            if (mid <= a || mid >= b) break;
            const double d = total(mid) - nf;
            if (ge ? d >= 0.0 : d > 0.0) b = mid;
            else a = mid;
        }
        return ge ? b : a;
    };
    const double mu = 0.5 * bisect(lo, hi, true) + 0.5 * bisect(lo, hi, false);
    std::vector<double> s(m);
    for (int i = 0; i < m; i++) {
        r.f[i] = occupation((ev[i] - mu) / kt);
        s[i] = fd_entropy((ev[i] - mu) / kt);
    }
    r.mu = mu;
    r.entropy = neumaier(s);
    return r;
}
static std::vector<int> ascending(const std::vector<double> &ev) {
    std::vector<int> o(ev.size());
    std::iota(o.begin(), o.end(), 0);
    std::stable_sort(o.begin(), o.end(), [&](int a, int b) { return ev[a] < ev[b]; });
    return o;
}
// qm/src/occupations.rs:299-317: the top level's lower bound theta - r clears both witness arms (WITNESS_GAP above
// eps_N and tail_kt kT above mu)
static bool witness_bound_reached(const std::vector<double> &ev, const std::vector<double> &res, int n, double mu,
                                  double kt, double tail_kt = WITNESS_TAIL_KT) {
    const int m = (int)ev.size();
    if (n == 0) return true;
    if (m <= n) return false;
    auto o = ascending(ev);
    const double eps_n = ev[o[n - 1]];
    const int top = o[m - 1];
    const double lower = ev[top] - res[top];
    return lower >= eps_n + WITNESS_GAP && lower - mu >= tail_kt * kt;
}
// qm/src/occupations.rs:372-399: t + 1 for the lowest rank t >= n whose lower bound clears both arms, t + 2 with a
// spare above it that clears them too
static int witness_keep(const std::vector<double> &ev, const std::vector<double> &res, int n, double mu, double kt,
                        double tail_kt = WITNESS_TAIL_KT) {
    const int m = (int)ev.size();
    if (n == 0 || m <= n) return m;
    auto o = ascending(ev);
    const double eps_n = ev[o[n - 1]];
    auto passes = [&](int rank) {
        const double lower = ev[o[rank]] - res[o[rank]];
        return lower >= eps_n + WITNESS_GAP && lower - mu >= tail_kt * kt;
    };
    for (int t = n; t < m; t++)
        if (passes(t)) return (t + 1 < m && passes(t + 1)) ? t + 2 : t + 1;
    return m;
}

// qm/src/occupations.rs:401-421: the lowest rank t >= n whose lower bound clears both witness arms with r_t <= r_max;
// returns t + 1 (t + 2 with a spare above), or -1 when none qualifies
static int witness_keep_within(const std::vector<double> &ev, const std::vector<double> &res, int n, double mu,
                               double kt, double r_max, double tail_kt = WITNESS_TAIL_KT) {
    const int m = (int)ev.size();
    if (n == 0) return m;
    if (m <= n) return -1;
    auto o = ascending(ev);
    const double eps_n = ev[o[n - 1]];
    auto passes = [&](int rank) {
        const double lower = ev[o[rank]] - res[o[rank]];
        return lower >= eps_n + WITNESS_GAP && lower - mu >= tail_kt * kt;
    };
    for (int t = n; t < m; t++)
        if (passes(t) && res[o[t]] <= r_max) return (t + 1 < m && passes(t + 1)) ? t + 2 : t + 1;
    return -1;
}

// qm/src/occupations.rs:319-326: ln(n_w / (STEP_TAIL_C * delta_rho_prev)) clamped to [STEP_TAIL_MIN_KT, WITNESS_TAIL_KT]
static double step_tail_kt(int n_w, double delta_rho_prev) {
    if (!(std::isfinite(delta_rho_prev) && delta_rho_prev > 0.0)) return WITNESS_TAIL_KT;
    const double x = std::log((double)std::max(n_w, 1) / (STEP_TAIL_C * delta_rho_prev));
    return std::min(std::max(x, STEP_TAIL_MIN_KT), WITNESS_TAIL_KT);
}

// qm/src/occupations.rs:352-366: the tight count a warm solve starts from, at least n
static int tight_start(const std::vector<double> &f, int n, double tol) {
    const double limit = 1e-2 * tol;
    double tail = 0.0;
    int loose_from = (int)f.size();
    for (int i = (int)f.size() - 1; i >= 0; i--) {
        tail += f[i];
        if (!(tail < limit)) break;
        loose_from = i;
    }
    return std::max(loose_from, n);
}

// Sequential elimination on a Gram matrix B (m x m, row-major): the columns whose Schur complement against the
// earlier kept columns is at or below rel x their norm^2 (crates/qm-gpu/src/eigen.rs:491-509)
static std::vector<int> dependent_columns(int m, const std::vector<double> &B, double rel) {
    std::vector<double> L((size_t)m * m, 0.0);
    std::vector<int> kept, dep;
    for (int j = 0; j < m; j++) {
        const double bjj = B[(size_t)j * m + j];
        std::vector<double> lj(kept.size());
        double s = bjj;
        for (size_t a = 0; a < kept.size(); a++) {
            const int kk = kept[a];
            double v = B[(size_t)j * m + kk];
            for (size_t b = 0; b < a; b++) v -= lj[b] * L[(size_t)kk * m + b];
            lj[a] = v / L[(size_t)kk * m + a];
            s -= lj[a] * lj[a];
        }
        if (!(bjj > 0.0) || !(s > rel * bjj)) {
            dep.push_back(j);
            continue;
        }
        for (size_t a = 0; a < kept.size(); a++) L[(size_t)j * m + a] = lj[a];
        L[(size_t)j * m + kept.size()] = std::sqrt(s);
        kept.push_back(j);
    }
    return dep;
}

// The Rayleigh-Ritz step with the first `locked` columns of S held fixed (crates/qm-gpu/src/eigen.rs:669-681): the
// pencil restricted to the B-orthogonal complement of the locked columns, Q = [-Byy^-1 Byz; I], and C = Q C'
// (m x na, row-major). Returns the amplification of the reduced step.
static double rayleigh_ritz_locked(int m, const std::vector<double> &A, const std::vector<double> &B, int locked,
                                   int na, double drop, std::vector<double> &th, std::vector<double> &C) {
    if (locked == 0) return rayleigh_ritz(m, A, B, na, drop, th, C);
    const int r = m - locked;
    std::vector<double> byy((size_t)locked * locked), K((size_t)locked * r);
    for (int i = 0; i < locked; i++) {
        for (int j = 0; j < locked; j++) byy[(size_t)i * locked + j] = B[(size_t)i * m + j];
        for (int j = 0; j < r; j++) K[(size_t)i * r + j] = B[(size_t)i * m + locked + j];
    }
    if (!cholesky(locked, byy)) throw NumericalError("lobpcg: Cholesky of the locked columns' Gram failed");
    chol_solve(locked, byy, K, r);
    auto reduce = [&](const std::vector<double> &M) {
        // M Q (m x r), then Q^T (M Q), symmetrised
        std::vector<double> MQ((size_t)m * r), out((size_t)r * r);
        for (int i = 0; i < m; i++)
            for (int j = 0; j < r; j++) {
                double v = M[(size_t)i * m + locked + j];
                for (int l = 0; l < locked; l++) v -= M[(size_t)i * m + l] * K[(size_t)l * r + j];
                MQ[(size_t)i * r + j] = v;
            }
        for (int a = 0; a < r; a++)
            for (int b = 0; b < r; b++) {
                double v = MQ[(size_t)(locked + a) * r + b];
                for (int l = 0; l < locked; l++) v -= K[(size_t)l * r + a] * MQ[(size_t)l * r + b];
                out[(size_t)a * r + b] = v;
            }
        for (int a = 0; a < r; a++)
            for (int b = 0; b < a; b++) {
                const double v = 0.5 * (out[(size_t)a * r + b] + out[(size_t)b * r + a]);
                out[(size_t)a * r + b] = out[(size_t)b * r + a] = v;
            }
        return out;
    };
    const std::vector<double> Ap = reduce(A), Bp = reduce(B);
    std::vector<double> Cp;
    const double alpha = rayleigh_ritz(r, Ap, Bp, na, drop, th, Cp);
    C.assign((size_t)m * na, 0.0);
    for (int j = 0; j < na; j++) {
        for (int i = 0; i < locked; i++) {
            double v = 0.0;
            for (int b = 0; b < r; b++) v -= K[(size_t)i * r + b] * Cp[(size_t)b * na + j];
            C[(size_t)i * na + j] = v;
        }
        for (int a = 0; a < r; a++) C[(size_t)(locked + a) * na + j] = Cp[(size_t)a * na + j];
    }
    return alpha;
}

// qm/src/occupations.rs:334-350: the columns whose upper-bound occupation f(theta - r; mu) the tail sum needs to stay
// below LOOSE_F_FACTOR * tol are tight; returns 1 + the largest column position among them (0 if none)
static int tight_columns(const std::vector<double> &ev, const std::vector<double> &res, double mu, double kt,
                         double tol) {
    if (std::isnan(mu)) return 0;
    auto o = ascending(ev);
    const double limit = 1e-2 * tol;
    double tail = 0.0;
    int loose_from = (int)o.size();
    for (int rank = (int)o.size() - 1; rank >= 0; rank--) {
        tail += occupation((ev[o[rank]] - res[o[rank]] - mu) / kt);
        if (!(tail < limit)) break;
        loose_from = rank;
    }
    int need = 0;
    for (int r = 0; r < loose_from; r++) need = std::max(need, o[r] + 1);
    return need;
}

// ------------------------------------------------------------------------------------ solver state ----
struct PoissonOut {
    int iterations = 0, outer = 0;
    double residual = 0.0, seconds = 0.0;
};
struct EigenOut {
    int iterations = 0, fallbacks = 0;
    double max_residual = 0.0;   // the tight (head) columns' worst exit residual
    double guard_residual = 0.0; // the loose tested columns' worst exit residual
    bool head_kept = true;       // the head came back bit for bit (a 0-iteration accept or a locked head)
    bool aborted = false;        // the inner abort stopped the solve
    int locked_calls = 0;
};

struct SolverIface {
    virtual ~SolverIface() = default;
    virtual lira_status solve(const lira_system *sys, const lira_settings *st, int warm, lira_abort_fn abort,
                              lira_abort_fn inner_abort, lira_observe_fn observe, void *user, lira_scf_stats *out) = 0;
    virtual void forces(const lira_nucleus *nuc, int nat, double sigma, double *out) = 0;
    virtual void get_field(int which, double *out) = 0;
    virtual void get_levels(int spin, double *evals, double *occ, int cap, int *count) = 0;
    virtual void get_orbitals(int spin, double *out, int cap, int *count) = 0;
    virtual PoissonOut poisson_host(const double *rho, double *v, double tol, int max_iter) = 0;
    virtual void mg_vcycle_host(const double *r, double *z) = 0;
    virtual void unoccupied(int m, lira_unoccupied_result *out, double *orbitals) = 0;
    virtual void harris(const lira_system *sys, const lira_settings *st, const double *rho_in_up,
                        const double *rho_in_dn, int m, lira_harris_result *out, double *rho_out_up,
                        double *rho_out_dn) = 0;
    virtual EigenOut lobpcg_host(const double *v, int k, double *x, double *evals, double lam, double tol,
                                 int max_iter) = 0;
    virtual double effective_eig_tol(double requested) const = 0;
    virtual void reset() = 0;
    virtual void inject_readback_fault(int after) = 0;
    virtual int precision() const = 0;
    virtual void set_ionic(const lira_species *sp, int nsp) = 0;
    virtual void force_terms(const lira_nucleus *nuc, int nat, double sigma, double *out) = 0;
    virtual void ionic_fields_host(const lira_nucleus *nuc, int nat, double sigma, double *vloc, double *pc) = 0;
    virtual void nonlocal_apply(const double *x, int nb, double *out, double *c, int c_cap, int *rows) = 0;
    virtual void set_witness_cap(int cap) = 0;
    virtual int witness_cap() const = 0;
    virtual void set_density_extrapolation(bool on) = 0;
    virtual void get_profile(lira_profile *out, double *drho, double *etol, int cap, int *n) const = 0;
    virtual void sample(const double *pts, int count, double *out) = 0;
    virtual void set_graphs(bool on) = 0;
    virtual void set_device_rr(bool on) = 0;
    virtual void state_summary(double *spin_abs, double *rho_max) = 0;
    virtual bool ionic() const = 0;
    virtual bool density_extrapolation() const = 0;
};

// One pseudised species as the solver keeps it: scalars plus offsets of its tables in the f64 table arena.
struct SpeciesHost {
    int z;
    double z_v, sigma_lr, du, r_sr, pc_r, r_core;
    int sr_off, sr_len, pc_off, pc_len;
    int nproj;
    int l[NL_MAX_PROJ], t_off[NL_MAX_PROJ], t_len[NL_MAX_PROJ];
    double radius[NL_MAX_PROJ], e_kb[NL_MAX_PROJ];
};

template <typename T> struct Channel {
    int kcap = 0, k = 0;
    int limit = 0;  // the most columns the channel may need (electrons + witness cap + buffer); 0 = unbounded
    int tight = 0;  // Fermi-Dirac: the leading columns held to the eigen tolerance (the rest to the guard tolerance)
    DevBuf X, HX, R, W, HW, P, HP, XN, PN, TMP;
    std::vector<double> evals, occ, res;
};

// Waits for e by polling it. The solve's host syncs are short (a Gram product, one CG iteration: tens of
// microseconds) and a blocking wait adds an interrupt's wake-up of about that order to each; past 2 ms of polling the
// wait blocks instead.
static void wait_event(hipEvent_t e) {
    const auto t0 = std::chrono::steady_clock::now();
    for (unsigned i = 1;; i++) {
        const hipError_t r = hipEventQuery(e);
        if (r == hipSuccess) return;
        if (r != hipErrorNotReady) LIRA_HIP(r);
        if ((i & 63) == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(2)) {
            LIRA_HIP(hipEventSynchronize(e));
            return;
        }
    }
}

template <typename T> class Solver final : public SolverIface {
  public:
    Solver(int n, double h);
    ~Solver() override {
        if (stream_) (void)hipStreamSynchronize(stream_);
        drop_graphs();
        for (hipEvent_t &e : rr_ev_)
            if (e) (void)hipEventDestroy(e);
        if (rb_ev_) (void)hipEventDestroy(rb_ev_);
        if (rr_host_) (void)hipHostFree(rr_host_);
        if (hbuf_) (void)hipHostFree(hbuf_);
        if (up_) (void)hipHostFree(up_);
        if (stream_) (void)hipStreamDestroy(stream_);
    }
    lira_status solve(const lira_system *sys, const lira_settings *st, int warm, lira_abort_fn abort,
                      lira_abort_fn inner_abort, lira_observe_fn observe, void *user, lira_scf_stats *out) override;
    void forces(const lira_nucleus *nuc, int nat, double sigma, double *out) override;
    void get_field(int which, double *out) override;
    void get_levels(int spin, double *evals, double *occ, int cap, int *count) override;
    void get_orbitals(int spin, double *out, int cap, int *count) override;
    PoissonOut poisson_host(const double *rho, double *v, double tol, int max_iter) override;
    void mg_vcycle_host(const double *r, double *z) override;
    void unoccupied(int m, lira_unoccupied_result *out, double *orbitals) override;
    void harris(const lira_system *sys, const lira_settings *st, const double *rho_in_up, const double *rho_in_dn, int m,
                lira_harris_result *out, double *rho_out_up, double *rho_out_dn) override;
    EigenOut lobpcg_host(const double *v, int k, double *x, double *evals, double lam, double tol,
                         int max_iter) override;
    double effective_eig_tol(double requested) const override {
        return sizeof(T) == 4 ? std::max(requested, EIG_TOL_FLOOR_FACTOR * residual_floor_f32(h_)) : requested;
    }
    void reset() override {
        valid_ = false;
        parked_ = false;
    }
    void inject_readback_fault(int after) override { fault_after_ = after < 0 ? -1 : after; }
    int precision() const override { return sizeof(T) == 8 ? LIRA_FP64 : LIRA_MIXED; }
    void set_ionic(const lira_species *sp, int nsp) override;
    void force_terms(const lira_nucleus *nuc, int nat, double sigma, double *out) override;
    void ionic_fields_host(const lira_nucleus *nuc, int nat, double sigma, double *vloc, double *pc) override;
    void nonlocal_apply(const double *x, int nb, double *out, double *c, int c_cap, int *rows) override;
    void set_witness_cap(int cap) override { witness_cap_ = std::min(std::max(cap, 1), MAX_WITNESS_GROWTH); }
    int witness_cap() const override { return witness_cap_; }
    void set_density_extrapolation(bool on) override {
        extrapolate_ = on;
        if (!on) have_prev_ = false;
    }
    void get_profile(lira_profile *out, double *drho, double *etol, int cap, int *n) const override {
        *out = profile_;
        const int cnt = (int)drho_steps_.size();
        if (n) *n = cnt;
        for (int i = 0; i < std::min(cap, cnt); i++) {
            if (drho) drho[i] = drho_steps_[i];
            if (etol) etol[i] = etol_steps_[i];
        }
    }
    void sample(const double *pts, int count, double *out) override;
    // GpuSolver::spin_abs_integral and density_max (crates/qm-gpu/src/scf.rs:1834-1845): one reduction pass over the
    // converged state, one readback
    void state_summary(double *spin_abs, double *rho_max) override {
        if (!valid_) throw InvalidError("lira_state_summary: no converged state");
        add2<T>(rho_t_.template as<T>(), rho_out_[0].template as<T>(), rho_out_[1].template as<T>(), N_, stream_);
        RedJobs<T> j{};
        j.a[0] = rho_out_[0].template as<T>();
        j.b[0] = rho_out_[1].template as<T>();
        j.kind[0] = RED_ABSDIFF;
        j.a[1] = j.b[1] = rho_t_.template as<T>();
        j.kind[1] = RED_MAX_ABS;
        j.count = 2;
        const std::vector<double> r = reduce(j, dv_);
        *spin_abs = r[0];
        *rho_max = r[1];
    }
    bool ionic() const override { return !species_.empty(); }
    bool density_extrapolation() const override { return extrapolate_; }
    void set_graphs(bool on) override {
        if (!on) drop_graphs();
        graphs_ = on;
    }
    void set_device_rr(bool on) override { device_rr_ = on; }

  private:
    // ---- HIP graphs: a fixed kernel sequence between two host syncs is captured once and replayed ----
    // A slot holds one captured sequence and the key it was captured under: the device pointers and sizes its
    // kernels were given. A different key (another buffer, a grown block) captures again. Captured bodies hold
    // kernel launches and device-to-device copies only: no host wait, no host copy, no memset.
    struct GraphSlot {
        std::vector<uintptr_t> key;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
    };
    bool graphs_ = false;  // opt-in (lira_set_graphs): one solver per thread may capture
    // LOBPCG's Rayleigh-Ritz step on the device from the second iteration on (subspaces up to RR_DEV_MAX columns),
    // with one iteration queued ahead of the host's read of the last iteration's record
    bool device_rr_ = false;  // opt-in (lira_set_device_rr): at the water's 27 columns the host step is faster
    DevBuf rr_work_, rr_rec_;
    double *rr_host_ = nullptr;  // pinned: two record slots
    hipEvent_t rr_ev_[2] = {nullptr, nullptr};
    void ensure_device_rr() {
        if (rr_host_) return;
        rr_work_.alloc(rr_step_work_doubles() * sizeof(double));
        rr_rec_.alloc(2 * RR_REC_STRIDE * sizeof(double));
        LIRA_HIP(hipHostMalloc((void **)&rr_host_, 2 * RR_REC_STRIDE * sizeof(double), hipHostMallocDefault));
        for (hipEvent_t &e : rr_ev_) LIRA_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    GraphSlot pcg_graph_[2], pcg_start_graph_;
    GraphSlot lobpcg_graph_[2];  // one LOBPCG iteration's device segment, per arrangement of the W / TMP buffers
    static void drop_graph(GraphSlot &g) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        if (g.graph) (void)hipGraphDestroy(g.graph);
        g.exec = nullptr;
        g.graph = nullptr;
        g.key.clear();
    }
    void drop_graphs() {
        for (GraphSlot *g : {&pcg_graph_[0], &pcg_graph_[1], &pcg_start_graph_, &lobpcg_graph_[0], &lobpcg_graph_[1]})
            drop_graph(*g);
    }
    // runs body's launches on stream_: replayed from g when g was captured under key, else captured into g first
    template <typename F> void run_graph(GraphSlot &g, const std::vector<uintptr_t> &key, F &&body) {
        if (!graphs_) {
            body();
            return;
        }
        if (g.exec == nullptr || g.key != key) {
            drop_graph(g);
            LIRA_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
            try {
                body();
            } catch (...) {
                hipGraph_t partial = nullptr;
                (void)hipStreamEndCapture(stream_, &partial);
                if (partial) (void)hipGraphDestroy(partial);
                throw;
            }
            LIRA_HIP(hipStreamEndCapture(stream_, &g.graph));
            LIRA_HIP(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
            g.key = key;
        }
        LIRA_HIP(hipGraphLaunch(g.exec, stream_));
    }

    int n_;
    double h_, dv_;
    long N_;
    hipStream_t stream_ = nullptr;
    double *hbuf_ = nullptr;
    size_t hcap_ = 0;
    hipEvent_t rb_ev_ = nullptr;  // marks a readback's copy
    double *up_ = nullptr;  // pinned upload arena: small host->device copies stay asynchronous
    size_t up_cap_ = 0, up_pos_ = 0;
    DevBuf partials_, dres_, gpart_, dG_, dC_, dsmall_;
    size_t gpart_cap_ = 0, dG_cap_ = 0, dC_cap_ = 0;
    int readbacks_ = 0;
    // Poisson
    struct Level {
        int n;
        double h, d_low, d_high;
        DevBuf x, b, res, tmp;
    };
    std::vector<Level> levels_;
    int mg_entry_ = 0;  // the level from which the V-cycle runs fused in one workgroup (mgc_)
    MgCoarse mgc_{};
    DevBuf pb_, pr64_, ptmp64_, pr_, pz_, pp_, pq_, pe_, psc_;
    DevBuf pmom_, pghost_, pcoef_;  // the moments' row/plane sums, the boundary ghost layers, R_lm's monomial table
    DevBuf pdot_;                   // the CG's r.z and r.r block partials (from the kernels that produce r and z)
    bool poisson_ready_ = false;
    int cg_total_ = 0;
    // SCF
    DevBuf vnuc_, vh_, vh_t_, rho_in_[2], rho_out_[2], vxc_[2], exc_, veff_[2], rho_t_, rho64_, pos_, zf_;
    struct Slot {
        DevBuf rho[2], res[2];
    };
    std::vector<Slot> slots_;
    Channel<T> ch_[2];
    bool valid_ = false;
    bool parked_ = false;  // an aborted solve's iterate, usable as the next warm start (not a converged state)
    int fault_after_ = -1; // the readback that is reported failed (0 = the next one); -1 = none
    int nat_ = 0;
    std::vector<double> pos_h_, z_h_;
    double sigma_ = 0.0;
    int occ_scheme_ = -1;
    double kt_ = 0.0;
    int nel_[2] = {0, 0};
    int lobpcg_total_ = 0;
    // pseudised ions: the setup (tables copied at lira_set_ionic) and the terms of the current geometry
    std::vector<SpeciesHost> species_;
    std::vector<double> tab_h_;
    DevBuf tab_, centres_, nl_ions_, beta_, nl_C_, nl_erow_, nl_members_, pc_, qf_, fpart_, fout_, nl_P_;
    size_t nl_C_cap_ = 0, nl_P_cap_ = 0;
    int ionic_version_ = 0, state_ionic_version_ = -1;
    std::vector<NlIon> nl_h_;
    std::vector<int> nl_atom_;  // the nucleus of each NlIon
    std::vector<double> nl_e_h_;
    struct Colour {
        int off, count;
        long max_pts;
    };
    std::vector<Colour> colours_;
    int nl_rows_ = 0;
    bool has_pc_ = false;
    std::vector<double> q_h_;  // the charge each nucleus carries (Z_v pseudised, Z all-electron)
    DevBuf atom_tab_, atoms_dev_, blobs_dev_, ao_tab_;
    // provider controls (crates/qm-gpu/src/scf.rs:1122-1170) and the last solve's profile
    int witness_cap_ = MAX_WITNESS_GROWTH;
    bool extrapolate_ = false, have_prev_ = false;
    DevBuf prev_rho_[2], save_rho_[2], pts_dev_, samp_dev_;
    lira_profile profile_{};
    std::vector<double> drho_steps_, etol_steps_;
    void cold_density(bool free_atoms, double sigma, const double *charges, double scale0, double scale1);
    void seed_atomic_orbitals(double sigma);
    // the free-atom orbitals of the current cold start, lowest first: column c of a channel's block starts from
    // entry c (cold columns and Fermi-Dirac growth alike) while entries last
    std::vector<AtomProfile> ao_cols_;
    std::vector<std::pair<int, int>> ao_lm_;
    void seed_column(T *dst, int col, int spin);
    const SpeciesHost *species_of(int z) const {
        for (const auto &sp : species_)
            if (sp.z == z) return &sp;
        return nullptr;
    }
    void ionic_geometry(double sigma);
    void nl_apply_block(const T *x, T *out, int nb, const T *x2 = nullptr, T *out2 = nullptr, int nb2 = 0);
    void nl_projections(const T *x, int nb, const T *x2 = nullptr, int nb2 = 0);
    void ensure_nl(int nb);  // the projections' buffers for nb bands
    std::vector<double> nl_energy_terms(Channel<T> &c, bool fd);

    void ensure_host(size_t count) {
        if (count <= hcap_) return;
        if (hbuf_) LIRA_HIP(hipHostFree(hbuf_));
        hcap_ = std::max(count, (size_t)4096);
        LIRA_HIP(hipHostMalloc((void **)&hbuf_, hcap_ * sizeof(double), hipHostMallocDefault));
    }
    // copies count doubles from the device and waits (the solve's only sync points)
    // A readback marked failed by lira_inject_readback_fault completes its copy and wait like any other, then reports a
    // device failure (the reference's readback-fault latch, crates/qm-gpu/src/grid.rs:113-157): no kernel is disturbed.
    const double *readback(const double *d, size_t count) {
        ensure_host(count);
        if (!rb_ev_) LIRA_HIP(hipEventCreateWithFlags(&rb_ev_, hipEventDisableTiming));
        LIRA_HIP(hipMemcpyAsync(hbuf_, d, count * sizeof(double), hipMemcpyDeviceToHost, stream_));
        LIRA_HIP(hipEventRecord(rb_ev_, stream_));
        wait_event(rb_ev_);  // the copy is the stream's last command: the stream is idle
        up_pos_ = 0;  // the stream is idle: every queued upload has been consumed
        readbacks_++;
        if (fault_after_ >= 0 && fault_after_-- == 0) throw DeviceError("a readback failed (injected)");
        return hbuf_;
    }
    // a blocking copy ordered on this solver's stream (which does not synchronise with the null stream)
    void copy_now(void *dst, const void *src, size_t bytes, hipMemcpyKind kind) {
        LIRA_HIP(hipMemcpyAsync(dst, src, bytes, kind, stream_));
        LIRA_HIP(hipStreamSynchronize(stream_));
    }
    // queues a host->device copy of count doubles through the pinned arena (no wait)
    void upload(double *d, const double *h, size_t count) {
        if (count == 0) return;
        if (up_pos_ + count > up_cap_) {
            LIRA_HIP(hipStreamSynchronize(stream_));
            up_pos_ = 0;
            if (count > up_cap_) {
                if (up_) LIRA_HIP(hipHostFree(up_));
                up_cap_ = std::max(count, (size_t)1 << 18);
                LIRA_HIP(hipHostMalloc((void **)&up_, up_cap_ * sizeof(double), hipHostMallocDefault));
            }
        }
        double *src = up_ + up_pos_;
        std::memcpy(src, h, count * sizeof(double));
        LIRA_HIP(hipMemcpyAsync(d, src, count * sizeof(double), hipMemcpyHostToDevice, stream_));
        up_pos_ += count;
    }
    T *fld(DevBuf &b) { return b.template as<T>(); }
    void alloc_field(DevBuf &b, int count = 1) {
        if (b.bytes < (size_t)N_ * count * sizeof(T)) b.alloc((size_t)N_ * count * sizeof(T));
    }
    void ensure_gram(int m, int q) {
        const size_t need = gram_partials_needed(m, q, N_);
        if (need > gpart_cap_) {
            gpart_.alloc(need * sizeof(double));
            gpart_cap_ = need;
        }
        const size_t g = (size_t)m * q + m + 64;
        if (g > dG_cap_) {
            dG_.alloc(g * sizeof(double));
            dG_cap_ = g;
        }
    }
    void ensure_coeff(size_t count) {
        if (count > dC_cap_) {
            dC_.alloc(count * sizeof(double));
            dC_cap_ = count;
        }
    }
    // uploads the m x q row-major coefficients of a lincomb into dC_, zero-padded to lincomb's shape; returns ldc
    // tail (optional) rides along in the same copy, right after the padded coefficients; *tail_dev receives its
    // device address (valid until the next upload_coeffs)
    int upload_coeffs(const std::vector<double> &C, int m, int q, const std::vector<double> *tail = nullptr,
                      const double **tail_dev = nullptr) {
        const int mp = lincomb_rows(m), ldc = lincomb_ldc(q);
        std::vector<double> pad((size_t)mp * ldc, 0.0);
        for (int i = 0; i < m; i++)
            for (int j = 0; j < q; j++) pad[(size_t)i * ldc + j] = C[(size_t)i * q + j];
        const size_t off = pad.size();
        if (tail) pad.insert(pad.end(), tail->begin(), tail->end());
        ensure_coeff(pad.size() + 64);
        upload(dC_.as<double>(), pad.data(), pad.size());
        if (tail_dev) *tail_dev = dC_.as<double>() + off;
        return ldc;
    }
    void ensure_channel(Channel<T> &c, int kc);
    // the stencil tile variant the tuner picked for H.X / Jacobi on this grid with nb bands (tuned on first use of a
    // block size, then kept; block sizes change with the witness columns)
    std::map<std::pair<int, int>, int> variants_;
    int variant(int mode, int nb) {
        const std::pair<int, int> key{mode, nb};
        auto it = variants_.find(key);
        if (it != variants_.end()) return it->second;
        const int v = nb > 0 ? tune_stencil_mode(mode, nb) : -1;
        variants_[key] = v;
        return v;
    }
    void hamiltonian(const T *v, const T *x, T *out, int nb) {
        StencilParams sp{n_, nb, h_, 1.0, 0.0, 0.0};
        stencil<T>(ST_HAM, sp, x, out, v, nullptr, nullptr, stream_, variant(ST_HAM, nb));
        if (nl_rows_ > 0) nl_apply_block(x, out, nb);  // + V_NL (qm/src/nonlocal.rs:165-174)
    }
    // H applied to two blocks in one stencil launch (x1: n1 bands, x2: n2 bands); V_NL per block
    void hamiltonian2(const T *v, const T *x1, T *out1, int n1, const T *x2, T *out2, int n2) {
        if (n2 <= 0) {
            hamiltonian(v, x1, out1, n1);
            return;
        }
        StencilParams sp{n_, n1 + n2, h_, 1.0, 0.0, 0.0};
        sp.nb1 = n1;
        sp.x2 = x2;
        sp.out2 = out2;
        stencil<T>(ST_HAM, sp, x1, out1, v, nullptr, nullptr, stream_, variant(ST_HAM, n1 + n2));
        if (nl_rows_ > 0) nl_apply_block(x1, out1, n1, x2, out2, n2);
    }
    void tune_stencil(int nb) {
        if (nb <= 0) return;
        variant(ST_HAM, nb);
        variant(ST_JAC, nb);
    }
    // times every usable stencil tile variant in one mode on this grid with nb bands (the block size of the solve)
    // and returns the fastest; the choice is kept per (precision, n, nb, mode) for later solvers in the process
    int tune_stencil_mode(int mode, int nb) {
        const std::array<int, 4> key{(int)sizeof(T), n_, nb, mode};
        // one tuning at a time in the process: solvers on other threads that want the same key wait for its result
        std::lock_guard<std::mutex> g(g_tune_mu);
        auto it = g_tuned.find(key);
        if (it != g_tuned.end()) return it->second;
        DevBuf x((size_t)N_ * nb * sizeof(T)), o((size_t)N_ * nb * sizeof(T)), v((size_t)N_ * sizeof(T));
        DevBuf lam((size_t)nb * sizeof(double));
        fill<T>(x.template as<T>(), N_ * nb, 0.0, stream_);
        fill<T>(v.template as<T>(), N_, 0.0, stream_);
        fill<double>(lam.as<double>(), nb, 0.0, stream_);
        hipEvent_t a, b;
        LIRA_HIP(hipEventCreate(&a));
        LIRA_HIP(hipEventCreate(&b));
        StencilParams sp{n_, nb, h_, 1.0, 1.0, 1.0};
        auto run = [&](int var) {
            stencil<T>(mode, sp, x.template as<T>(), o.template as<T>(), v.template as<T>(),
                       mode == ST_JAC ? x.template as<T>() : nullptr, mode == ST_JAC ? lam.as<double>() : nullptr,
                       stream_, var);
        };
        float best = 1e30f;
        int pick = 0;
        for (int var = 0; var < stencil_variants(); var++) {
            if (!stencil_variant_usable(var, n_, (int)sizeof(T))) continue;
            run(var);
            LIRA_HIP(hipEventRecord(a, stream_));
            for (int r = 0; r < 8; r++) run(var);
            LIRA_HIP(hipEventRecord(b, stream_));
            LIRA_HIP(hipEventSynchronize(b));
            float ms = 0.f;
            LIRA_HIP(hipEventElapsedTime(&ms, a, b));
            if (ms < best) {
                best = ms;
                pick = var;
            }
        }
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        g_tuned[key] = pick;
        return pick;
    }
    std::vector<double> reduce(RedJobs<T> &jobs, double scale) {
        reduce_jobs<T>(jobs, n_, scale, partials_.as<double>(), dres_.as<double>(), stream_);
        const double *r = readback(dres_.as<double>(), jobs.count);
        return std::vector<double>(r, r + jobs.count);
    }
    // any number of reductions (up to RES_SLOTS) in launches of MAX_JOBS, results in consecutive slots, one readback
    struct Job {
        const T *a, *b;
        int kind;
    };
    static constexpr int RES_SLOTS = 960;  // dres_ holds 1024 doubles; the slots above serve single reductions
    std::vector<double> reduce_list(const std::vector<Job> &jobs, double scale) {
        if (jobs.empty()) return {};
        if ((int)jobs.size() > RES_SLOTS) throw InvalidError("too many reductions in one readback");
        for (size_t c0 = 0; c0 < jobs.size(); c0 += MAX_JOBS) {
            RedJobs<T> rj{};
            for (size_t j = c0; j < std::min(jobs.size(), c0 + MAX_JOBS); j++) {
                rj.a[j - c0] = jobs[j].a;
                rj.b[j - c0] = jobs[j].b;
                rj.kind[j - c0] = jobs[j].kind;
                rj.count = (int)(j - c0 + 1);
            }
            reduce_jobs<T>(rj, n_, scale, partials_.as<double>(), dres_.as<double>() + c0, stream_);
        }
        const double *r = readback(dres_.as<double>(), jobs.size());
        return std::vector<double>(r, r + jobs.size());
    }
    // Poisson
    void poisson_setup();
    void vcycle(int l, T *x, const T *rhs, double *dot_partials = nullptr);
    PoissonOut poisson(const T *rho, double *v, double tol, int max_iter);
    int pcg(const T *r0, T *e, double rel_tol, int max_iter, double *rel_out);
    // LOBPCG on a channel's X (k columns) for H = -1/2 lap + v
    EigenOut lobpcg(Channel<T> &c, const T *v, int k, double lam, double tol, int max_iter, double tail_tol, int head,
                    int checked, double guess_tol, bool lock);
    bool solve_channel(int spin, const T *v, int nel, bool fd, double kt, double tol, double scf_tol, bool guess_band,
                       double loosened_delta, EigenOut &eo, Fd &occ);
    void permute_columns(Channel<T> &c, const std::vector<int> &order);
    int buffer_last(Channel<T> &c);
    void sort_block(Channel<T> &c);
    // display solves: the unoccupied view's block and kept vectors, the Harris solve's channels and fields
    Channel<T> disp_, hch_[2];
    DevBuf unocc_[2];
    int unocc_k_[2] = {0, 0};
    DevBuf hr_[2], hrho_, hvh_, hvht_, hexc_, hxc_[2], hv_[2], hvnuc_, hvnuc_save_;
    double last_tol_ = 1e-6, last_eig_tol_ = 1e-6;  // the converged solve's settings (display tolerances)
    // the SCF's inner abort, polled once per LOBPCG iteration under Fermi-Dirac (solve_polled's inner_abort)
    lira_abort_fn inner_abort_ = nullptr;
    void *abort_user_ = nullptr;
    double step_tail_ = 0.0;  // the widest witness tail of the current SCF step (kT)
    int doublings_ = 0;       // witness growths that doubled the extra block (the widened-window rule reads them)
    void upload_nuclei(const lira_nucleus *nuc, int nat);
};

template <typename T> Solver<T>::Solver(int n, double h) : n_(n), h_(h), dv_(h * h * h), N_((long)n * n * n) {
    // non-blocking: solvers driven from several host threads must not serialise on null-stream work
    LIRA_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    up_cap_ = (size_t)1 << 18;
    LIRA_HIP(hipHostMalloc((void **)&up_, up_cap_ * sizeof(double), hipHostMallocDefault));
    partials_.alloc((size_t)RED_BLOCKS * std::max(MAX_JOBS, 49) * sizeof(double) * 4);
    dres_.alloc(1024 * sizeof(double));
    dsmall_.alloc(4096 * sizeof(double));
    ensure_host(1 << 16);
}

template <typename T> void Solver<T>::ensure_channel(Channel<T> &c, int kc) {
    if (kc <= c.kcap) return;
    // headroom for witness growth, never beyond what the witness cap lets the channel hold
    if (c.kcap > 0) kc = std::max(kc, std::min({2 * c.kcap, kc + 8, std::max(kc, c.limit)}));
    const size_t bytes = (size_t)N_ * kc * sizeof(T);
    DevBuf nx(bytes);
    if (c.k > 0) LIRA_HIP(hipMemcpyAsync(nx.ptr, c.X.ptr, (size_t)N_ * c.k * sizeof(T), hipMemcpyDeviceToDevice, stream_));
    c.X = std::move(nx);
    for (DevBuf *b : {&c.HX, &c.R, &c.W, &c.HW, &c.P, &c.HP, &c.XN, &c.PN, &c.TMP}) b->alloc(bytes);
    c.kcap = kc;
}

// -------------------------------------------------------------------------------------------- Poisson ----
template <typename T> void Solver<T>::poisson_setup() {
    if (poisson_ready_) return;
    pb_.alloc(N_ * sizeof(double));
    pr64_.alloc(N_ * sizeof(double));
    ptmp64_.alloc(N_ * sizeof(double));
    pmom_.alloc(moment_work_doubles(n_) * sizeof(double));
    pghost_.alloc((size_t)24 * n_ * n_ * sizeof(double));
    {
        const size_t cnt = (size_t)(mg_dot_blocks(n_) + 4 * ((N_ + 255) / 256)) + dot_finish_scratch_doubles();
        pdot_.alloc(cnt * sizeof(double));
        LIRA_HIP(hipMemset(pdot_.ptr, 0, cnt * sizeof(double)));  // the finish's counter starts at zero
    }
    {
        std::vector<double> coef(MOMENT_COEFS);
        moment_coefficients(coef.data());
        pcoef_.alloc(coef.size() * sizeof(double));
        LIRA_HIP(hipMemcpy(pcoef_.ptr, coef.data(), coef.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    for (DevBuf *b : {&pr_, &pz_, &pp_, &pq_, &pe_}) alloc_field(*b);
    psc_.alloc(64 * sizeof(double));
    // level 0 works in pz_ (x) with rhs pr_; coarse levels own x and b
    levels_.clear();
    Level l0{n_, h_, 1.0, 1.0, {}, {}, {}, {}};
    l0.res.alloc(N_ * sizeof(T));
    l0.tmp.alloc(N_ * sizeof(T));
    levels_.push_back(std::move(l0));
    while (levels_.back().n > MG_COARSEST_MAX) {
        const Level &f = levels_.back();
        const int nc = f.n / 2;
        Level c{nc, 2.0 * f.h, (f.d_low + 1.0) / 2.0, (f.d_high + (double)(f.n - 2 * nc)) / 2.0, {}, {}, {}, {}};
        const size_t bytes = (size_t)nc * nc * nc * sizeof(T);
        c.x.alloc(bytes);
        c.b.alloc(bytes);
        c.res.alloc(bytes);
        c.tmp.alloc(bytes);
        levels_.push_back(std::move(c));
    }
    // the finest level from which every coarser level fits the fused kernel's LDS budget
    mg_entry_ = (int)levels_.size();
    for (int l = (int)levels_.size() - 1; l >= 0 && (int)levels_.size() - l <= MG_MAX_LEVELS; l--) {
        MgCoarse a{};
        a.levels = (int)levels_.size() - l;
        int off = 0;
        for (int q = 0; q < a.levels; q++) {
            const Level &L = levels_[l + q];
            a.n[q] = L.n;
            a.h2[q] = L.h * L.h;
            a.gl[q] = 1.0 - 1.0 / L.d_low;
            a.gh[q] = 1.0 - 1.0 / L.d_high;
            a.off[q] = off;
            off += 3 * L.n * L.n * L.n;
        }
        a.pre = MG_PRE;
        a.post = MG_POST;
        a.coarse = MG_COARSE;
        a.omega = MG_OMEGA;
        if (mg_coarse_bytes(a, (int)sizeof(T)) > MG_COARSE_LDS) break;
        mg_entry_ = l;
        mgc_ = a;
    }
    poisson_ready_ = true;
}

// one V-cycle from level l: x = M_l^{-1} rhs (crates/qm-gpu/src/multigrid.rs:165-181)
// dot_partials (finest level only, never the fused coarse kernel): the last post-smoothing sweep also leaves the
// block partials of rhs . x
template <typename T> void Solver<T>::vcycle(int l, T *x, const T *rhs, double *dot_partials) {
    if (l == mg_entry_) {
        mg_coarse<T>(x, rhs, mgc_, stream_);  // this level and every coarser one in one workgroup
        return;
    }
    Level &L = levels_[l];
    const long Nl = (long)L.n * L.n * L.n;
    const double h2 = L.h * L.h, gl = 1.0 - 1.0 / L.d_low, gh = 1.0 - 1.0 / L.d_high;
    T *tmp = L.tmp.template as<T>();
    // x = 0, then sweeps in pairs; the first sweep of the first pair reads no x (the zero fill folded into it)
    auto smooth = [&](int sweeps, bool from_zero, double *dot) {
        for (int s = 0; s < sweeps; s += 2) {
            if (from_zero && s == 0) mg_jacobi0<T>(tmp, rhs, L.n, h2, MG_OMEGA, gl, gh, stream_);
            else mg_jacobi<T>(tmp, x, rhs, L.n, h2, MG_OMEGA, gl, gh, stream_);
            if (dot && s + 2 >= sweeps) mg_jacobi_dot<T>(x, tmp, rhs, L.n, h2, MG_OMEGA, gl, gh, dot, stream_);
            else mg_jacobi<T>(x, tmp, rhs, L.n, h2, MG_OMEGA, gl, gh, stream_);
        }
    };
    if (l + 1 == (int)levels_.size()) {
        if (MG_COARSE == 0) fill<T>(x, Nl, 0.0, stream_);
        smooth(MG_COARSE, true, dot_partials);
        return;
    }
    if (MG_PRE == 0) fill<T>(x, Nl, 0.0, stream_);
    smooth(MG_PRE, true, nullptr);
    mg_residual<T>(L.res.template as<T>(), x, rhs, L.n, h2, gl, gh, stream_);
    Level &C = levels_[l + 1];
    mg_restrict<T>(C.b.template as<T>(), L.res.template as<T>(), C.n, L.n, stream_);
    vcycle(l + 1, C.x.template as<T>(), C.b.template as<T>());
    mg_prolong_add<T>(x, C.x.template as<T>(), L.n, C.n, stream_);
    smooth(MG_POST, false, dot_partials);
}

// Preconditioned CG in precision T on A e = r0 (A = -lap, 8th order, Dirichlet) from e = 0, with the V-cycle as M.
// The CG scalars live on the device (alpha, beta never cross to the host); the residual norm is read back every
// iteration with a one-iteration lag, so the host never waits on the iteration it just queued.
template <typename T> int Solver<T>::pcg(const T *r0, T *e, double rel_tol, int max_iter, double *rel_out) {
    T *r = pr_.template as<T>(), *z = pz_.template as<T>(), *p = pp_.template as<T>(), *q = pq_.template as<T>();
    double *sc = psc_.as<double>();
    // every launch below reads and writes buffers fixed for the solver's life (the MG levels, the CG vectors, the
    // scalar slots, the reduction partials); the key adds the two that vary between callers
    auto key = [&](uintptr_t a, uintptr_t b) {
        return std::vector<uintptr_t>{a, b, (uintptr_t)r, (uintptr_t)z, (uintptr_t)p, (uintptr_t)q, (uintptr_t)sc,
                                      (uintptr_t)partials_.ptr, (uintptr_t)dres_.ptr, (uintptr_t)levels_[0].tmp.ptr};
    };
    // sc: [4 c] r.z and [4 c + 1] r.r of the current slot pair c (c alternates between 0 and 1), [2] p.q, [8] r0.r0
    run_graph(pcg_start_graph_, key((uintptr_t)r0, (uintptr_t)e), [&] {
        if (r0 != r) copy<T>(r, r0, N_, stream_);
        fill<T>(e, N_, 0.0, stream_);
        RedJobs<T> j0{};
        j0.a[0] = r;
        j0.b[0] = r;
        j0.kind[0] = RED_DOT;
        j0.count = 1;
        reduce_jobs<T>(j0, n_, 1.0, partials_.as<double>(), sc + 8, stream_);
        vcycle(0, p, r);  // the first direction is the preconditioned residual itself
        RedJobs<T> jrz{};
        jrz.a[0] = r;
        jrz.b[0] = p;
        jrz.kind[0] = RED_DOT;
        jrz.count = 1;
        reduce_jobs<T>(jrz, n_, 1.0, partials_.as<double>(), sc + 0, stream_);
    });
    const double r00 = readback(sc + 8, 1)[0];
    if (!(r00 > 0.0)) {
        *rel_out = 0.0;
        return 0;
    }
    const double r0n = std::sqrt(r00);
    StencilParams sp{n_, 1, h_, -1.0, 0.0, 0.0};
    const int lap_variant = variant(ST_LAP, 1);  // tuned on first use, outside any graph capture
    int cur = 0, it = 0;
    double rel = 1.0;
    hipEvent_t ev[2];
    LIRA_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    LIRA_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    double *hrr = hbuf_;  // two lagged slots
    bool pending[2] = {false, false};
    bool done = false;
    while (!done && it < max_iter) {
        // one iteration; cur (which rz slot is current) alternates, so two captured forms
        run_graph(pcg_graph_[cur], key((uintptr_t)e, (uintptr_t)cur), [&] {
            stencil<T>(ST_LAP, sp, p, q, nullptr, nullptr, nullptr, stream_, lap_variant);  // q = A p = -lap p
            if (mg_entry_ > 0) {
                double *pr = pdot_.as<double>(), *pz = pr + 4 * ((N_ + 255) / 256);
                const long np = dot_partials<T>(p, q, N_, pr, stream_);
                dot_finish(pr, np, nullptr, 0, sc + 2, pz + mg_dot_blocks(n_), stream_);
            } else {
                RedJobs<T> jpq{};
                jpq.a[0] = p;
                jpq.b[0] = q;
                jpq.kind[0] = RED_DOT;
                jpq.count = 1;
                reduce_jobs<T>(jpq, n_, 1.0, partials_.as<double>(), sc + 2, stream_);
            }
            if (mg_entry_ > 0) {
                // the new r.r and r.z as block partials of the kernels that write r and z, summed in one launch
                double *pr = pdot_.as<double>(), *pz = pr + 4 * ((N_ + 255) / 256);
                double *fin = pz + mg_dot_blocks(n_);
                const long nr = cg_update_dot<T, T>(e, p, r, q, sc, 4 * cur, 2, N_, pr, stream_);
                vcycle(0, z, r, pz);
                dot_finish(pz, mg_dot_blocks(n_), pr, nr, sc + 4 * (1 - cur), fin, stream_);
            } else {
                cg_update<T, T>(e, p, r, q, sc, 4 * cur, 2, N_, stream_);
                vcycle(0, z, r);
                RedJobs<T> jn{};
                jn.a[0] = r;
                jn.b[0] = z;
                jn.kind[0] = RED_DOT;
                jn.a[1] = r;
                jn.b[1] = r;
                jn.kind[1] = RED_DOT;
                jn.count = 2;
                // the new r.z and r.r into the other slot pair (consecutive slots, as the reduction writes them)
                reduce_jobs<T>(jn, n_, 1.0, partials_.as<double>(), sc + 4 * (1 - cur), stream_);
            }
            cg_direction<T>(p, z, sc, 4 * (1 - cur), 4 * cur, N_, stream_);
        });
        const int slot = it & 1;
        LIRA_HIP(hipMemcpyAsync(hrr + slot, sc + 4 * (1 - cur) + 1, sizeof(double), hipMemcpyDeviceToHost, stream_));
        LIRA_HIP(hipEventRecord(ev[slot], stream_));
        pending[slot] = true;
        cur = 1 - cur;
        it++;
        // check the previous iteration's residual while this one runs
        const int prev = slot ^ 1;
        if (pending[prev]) {
            wait_event(ev[prev]);
            pending[prev] = false;
            rel = std::sqrt(std::max(hrr[prev], 0.0)) / r0n;
            if (!std::isfinite(rel)) throw NumericalError("poisson: non-finite CG residual");
            if (rel <= rel_tol) done = true;
        }
    }
    LIRA_HIP(hipStreamSynchronize(stream_));
    const int last = (it - 1) & 1;
    if (it > 0) rel = std::sqrt(std::max(hrr[last], 0.0)) / r0n;
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    readbacks_ += it;
    *rel_out = rel;
    cg_total_ += it;
    return it;
}

// del^2 V = -4 pi rho with the multipole boundary (qm/src/poisson.rs:200-255): b = 4 pi rho + L_bc in f64, then
// f64 residual corrections V += e with e from the precision-T MG-PCG, until ||b - A V|| / ||b|| <= tol.
template <typename T> PoissonOut Solver<T>::poisson(const T *rho, double *v, double tol, int max_iter) {
    auto t0 = Clock::now();
    poisson_setup();
    PoissonOut out;
    double *q = dsmall_.as<double>();
    multipole_moments<T>(rho, n_, h_, pmom_.as<double>(), pcoef_.as<double>(), q, stream_);
    double *b = pb_.as<double>(), *r64 = pr64_.as<double>(), *t64 = ptmp64_.as<double>();
    poisson_rhs<T>(b, rho, q, pghost_.as<double>(), n_, h_, stream_);
    RedJobs<double> jb{};
    jb.a[0] = b;
    jb.b[0] = b;
    jb.kind[0] = RED_DOT;
    jb.count = 1;
    reduce_jobs<double>(jb, n_, 1.0, partials_.as<double>(), dres_.as<double>(), stream_);
    double bb = readback(dres_.as<double>(), 1)[0];
    const double bnorm = std::max(std::sqrt(bb), 1e-300);
    StencilParams sp{n_, 1, h_, 1.0, 0.0, 0.0};
    const double inner_floor = sizeof(T) == 4 ? 2e-5 : 1e-14;
    for (int outer = 0; outer < 12; outer++) {
        stencil<double>(ST_LAP, sp, v, t64, nullptr, nullptr, nullptr, stream_);  // lap V
        add2<double>(r64, b, t64, N_, stream_);                                   // r = b - A V = b + lap V
        RedJobs<double> jr{};
        jr.a[0] = r64;
        jr.b[0] = r64;
        jr.kind[0] = RED_DOT;
        jr.count = 1;
        reduce_jobs<double>(jr, n_, 1.0, partials_.as<double>(), dres_.as<double>(), stream_);
        const double rr = readback(dres_.as<double>(), 1)[0];
        out.residual = std::sqrt(std::max(rr, 0.0)) / bnorm;
        if (!std::isfinite(out.residual)) throw NumericalError("poisson: non-finite residual (NaN in rho?)");
        if (out.residual <= tol || out.iterations >= max_iter) break;
        out.outer++;
        const double rel = std::max(inner_floor, 0.3 * tol / out.residual);
        T *e = pe_.template as<T>();
        const T *r0;
        if (sizeof(T) == 8) {
            r0 = reinterpret_cast<const T *>(r64);
        } else {
            convert<T, double>(pr_.template as<T>(), r64, N_, stream_);  // straight into the CG's residual
            r0 = pr_.template as<T>();
        }
        double rel_out = 0.0;
        out.iterations += pcg(r0, e, rel, max_iter - out.iterations, &rel_out);
        add_into_f64<T>(v, e, N_, stream_);
    }
    LIRA_HIP(hipStreamSynchronize(stream_));
    out.seconds = secs(t0);
    return out;
}

// ------------------------------------------------------------------------------------------- LOBPCG ----
// Block LOBPCG with one readback per iteration (crates/qm-gpu/src/eigen.rs:685-943, restated): R = HX - theta X;
// W = 4-sweep damped Jacobi on (H - lam_i) w = r (lam_i the column's Ritz value, or one fixed shift); one Gram of S = [X, W, P] against [S, R, HW, HP] gives
// B = S^T S and A = S^T H S (A's X columns restored from R); host f64 generalised Rayleigh–Ritz with the
// conditioning guard; X, P updated by one block lincomb on the device; HX recomputed by the stencil.
// LOBPCG on the channel's first k columns for H = -1/2 lap + v (+ V_NL) (crates/qm-gpu/src/eigen.rs:685-943). The
// guess is measured (X^T X and <x_i, H x_i>, one readback) and neither rotated nor rescaled, so a guess that already
// meets guess_tol (>= tol) at the first test comes back bit for bit; a guess column within sqrt(GUESS_DEP) of the span
// of the earlier ones is re-seeded. Split stop: the first `head` columns at tol, the rest below `checked` at tail_tol,
// the columns from `checked` on untested. lock: when the head meets guess_tol at the first test and the tested tail
// does not, the head is locked, keeps its columns bitwise, and only the others move (eigen.rs:815-818).
template <typename T>
EigenOut Solver<T>::lobpcg(Channel<T> &c, const T *v, int k, double lam, double tol, int max_iter, double tail_tol,
                           int head, int checked, double guess_tol, bool lock) {
    EigenOut out;
    if (k <= 0) return out;
    const double diag_t = -1.5 * COEFF0 / (h_ * h_);
    const double omega = jacobi_omega();
    const double floor = sizeof(T) == 4 ? residual_floor_f32(h_) : 0.0;
    guess_tol = std::max(guess_tol, tol);
    double *dth = dsmall_.as<double>() + 256, *dlam = dsmall_.as<double>() + 512;
    // the Jacobi preconditioner's shift: per column its current Ritz value (lam NaN: the reference's GPU solver,
    // crates/qm-gpu/src/eigen.rs:740-757), or one fixed shift for every column (crates/qm/src/eigen.rs:227)
    const bool per_column = std::isnan(lam);
    if (!per_column) {
        std::vector<double> l(k, lam);
        upload(dlam, l.data(), k);
    }
    auto X = [&]() { return c.X.template as<T>(); };
    auto col = [&](DevBuf &b, int i) { return b.template as<T>() + (long)i * N_; };
    ensure_gram(3 * k, 6 * k);
    // the guess: B = X^T X and the Rayleigh quotients theta_i = <x_i, H x_i> / <x_i, x_i>
    std::vector<double> theta(k);
    for (int round = 0;; round++) {
        hamiltonian(v, X(), c.HX.template as<T>(), k);
        Cols S{{{X(), k}}, 1}, U{{{X(), k}, {c.HX.ptr, k}}, 2};
        gram<T>(S, U, N_, dv_, gpart_.as<double>(), gpart_cap_, dG_.as<double>(), stream_);
        const double *g = readback(dG_.as<double>(), (size_t)k * 2 * k);
        for (size_t i = 0; i < (size_t)k * 2 * k; i++)
            if (!std::isfinite(g[i])) throw NumericalError("lobpcg: non-finite Gram of the guess (NaN in V or X?)");
        std::vector<double> B((size_t)k * k);
        for (int i = 0; i < k; i++) {
            for (int j = 0; j <= i; j++) B[(size_t)i * k + j] = B[(size_t)j * k + i] = g[(size_t)i * 2 * k + j];
            theta[i] = g[(size_t)i * 2 * k + k + i] / B[(size_t)i * k + i];
        }
        const std::vector<int> dep = dependent_columns(k, B, GUESS_DEP);
        if (dep.empty()) break;
        if (round + 1 >= MAX_RESEED_ROUNDS)
            throw NumericalError("lobpcg: could not generate " + std::to_string(k) + " independent guess vectors");
        // deterministic noise under a Gaussian envelope about the nuclei (about the box centre without nuclei)
        const double *cpos = pos_.as<double>();
        int cnat = nat_;
        if (cnat == 0 || !cpos) {
            const double zero[3] = {0.0, 0.0, 0.0};
            upload(dsmall_.as<double>() + 3072, zero, 3);
            cpos = dsmall_.as<double>() + 3072;
            cnat = 1;
        }
        for (int i : dep)
            initial_orbitals<T>(col(c.X, i), n_, h_, cpos, cnat, 1, 7919ULL * (uint64_t)(round + 1) + 2ULL * (uint64_t)i,
                                stream_);
    }
    std::vector<double> lambda(theta), res(k, 0.0);
    int np = 0;      // valid P columns: P[locked .. locked + np)
    int locked = 0;  // the locked head; 0 = none
    const double *dth_cur = nullptr;
    for (;;) {
        const int lk0 = locked;     // the lock this iteration's S was built under
        const int na = k - lk0;     // its W columns
        if (device_rr_ && out.iterations >= 1 && np == na && k + na + np <= RR_DEV_MAX) {
            // From here on the step runs on the device (the lock and the block's shape are now fixed): each queued
            // iteration measures, tests and updates on the GPU and leaves a record; the host reads iteration t's
            // record while iteration t + 1 runs. An iteration whose test stops writes the identity, so the one queued
            // behind it changes nothing, and the records of both equal the stopping iteration's.
            ensure_device_rr();
            const int m = k + na + np, q = 2 * m, nm = na;
            const int mp = lincomb_rows(m), ldc = lincomb_ldc(2 * nm);
            ensure_gram(m, q);
            ensure_coeff((size_t)mp * ldc + 64);
            upload(dth, theta.data(), k);
            const double *dshift = per_column ? dth : dlam;
            RrStepArgs ra{};
            ra.G = dG_.as<double>();
            ra.theta = dth;
            ra.C = dC_.as<double>();
            ra.work = rr_work_.as<double>();
            ra.m = m, ra.q = q, ra.k = k, ra.na = na, ra.np = np, ra.nm = nm, ra.locked = locked;
            ra.head = head, ra.checked = checked, ra.max_iter = max_iter, ra.is_f32 = sizeof(T) == 4;
            ra.mp = mp, ra.ldc = ldc;
            ra.tol = tol, ra.guess_tol = guess_tol, ra.tail_scale = tol / tail_tol, ra.floor = floor;
            ra.alpha_max = ALPHA_MAX, ra.alpha_fraction = ALPHA_RESIDUAL_FRACTION, ra.drop = RR_DROP;
            auto enqueue = [&](int t) {
                residual_precond<T>(c.R.template as<T>(), c.W.template as<T>(), c.HX.template as<T>(), X(), dth, v,
                                    dshift, diag_t, k, N_, stream_);
                StencilParams jp{n_, na, h_, 1.0, omega, diag_t};
                for (int sw = 0; sw < 3; sw++) {
                    stencil<T>(ST_JAC, jp, col(c.W, locked), col(c.TMP, locked), v, col(c.R, locked),
                               dshift + locked, stream_, variant(ST_JAC, na));
                    std::swap(c.W, c.TMP);
                }
                hamiltonian2(v, col(c.W, locked), col(c.HW, locked), na, col(c.P, locked), col(c.HP, locked), np);
                Cols S{{{X(), k}, {col(c.W, locked), na}, {col(c.P, locked), np}}, 3};
                Cols U{{{X(), k}, {col(c.W, locked), na}, {col(c.P, locked), np}, {c.R.ptr, k},
                        {col(c.HW, locked), na}, {col(c.HP, locked), np}}, 6};
                double *dG = dG_.as<double>();
                gram<T>(S, U, N_, dv_, gpart_.as<double>(), gpart_cap_, dG, stream_);
                RedJobs<T> jr{};
                for (int i = 0; i < k; i++) {
                    jr.a[i % MAX_JOBS] = col(c.R, i);
                    jr.b[i % MAX_JOBS] = jr.a[i % MAX_JOBS];
                    jr.kind[i % MAX_JOBS] = RED_DOT;
                    jr.count = i % MAX_JOBS + 1;
                    if (jr.count == MAX_JOBS || i == k - 1) {
                        reduce_jobs<T>(jr, n_, dv_, partials_.as<double>(),
                                       dG + (size_t)m * q + (i / MAX_JOBS) * MAX_JOBS, stream_);
                        jr.count = 0;
                    }
                }
                ra.iter = t;
                ra.record = rr_rec_.as<double>() + (t & 1) * RR_REC_STRIDE;
                rr_step(ra, stream_);
                LIRA_HIP(hipMemcpyAsync(rr_host_ + (t & 1) * RR_REC_STRIDE, ra.record,
                                        (RR_REC_HEAD + 2 * k) * sizeof(double), hipMemcpyDeviceToHost, stream_));
                LIRA_HIP(hipEventRecord(rr_ev_[t & 1], stream_));
                Cols Sm{{{X(), k}, {col(c.W, locked), na}, {col(c.P, locked), np}}, 3};
                lincomb<T>(Sm, dC_.as<double>(), ldc, 2 * nm, col(c.XN, locked), N_, stream_, nm, col(c.PN, locked));
                if (locked > 0)
                    LIRA_HIP(hipMemcpyAsync(c.XN.ptr, c.X.ptr, (size_t)N_ * locked * sizeof(T),
                                            hipMemcpyDeviceToDevice, stream_));
                std::swap(c.X, c.XN);
                std::swap(c.P, c.PN);
                hamiltonian(v, col(c.X, locked), col(c.HX, locked), nm);
            };
            int t = out.iterations, queued = t;
            enqueue(t);
            for (;;) {
                if (queued == t && t < max_iter) enqueue(++queued);
                wait_event(rr_ev_[t & 1]);
                readbacks_++;
                if (fault_after_ >= 0 && fault_after_-- == 0) throw DeviceError("a readback failed (injected)");
                const double *r = rr_host_ + (t & 1) * RR_REC_STRIDE;
                for (int i = 0; i < k; i++) {
                    lambda[i] = r[RR_REC_HEAD + i];
                    res[i] = r[RR_REC_HEAD + k + i];
                }
                out.max_residual = r[2];
                out.guard_residual = r[3];
                if (r[1] != 0.0) {
                    const int code = (int)r[1];
                    throw NumericalError(code == 1 ? "lobpcg: non-finite Gram entry or residual (NaN in a vector or in V)"
                                         : code == 2 ? "lobpcg: the subspace lost rank in the Rayleigh-Ritz step"
                                         : code == 4 ? "lobpcg: a column of the subspace has a non-positive norm"
                                                     : "lobpcg: Cholesky of X^T X failed (X lost rank)");
                }
                if (r[0] != 0.0) break;
                if (trace_on())
                    std::fprintf(stderr, "lobpcg device step %d: m %d, Jacobi sweeps %d, amplification %.3g%s\n", t, m,
                                 (int)r[6], r[5], r[4] != 0.0 ? ", without P" : "");
                if (inner_abort_ && inner_abort_(abort_user_)) {
                    out.aborted = true;
                    break;
                }
                out.iterations = t + 1;
                if (r[4] != 0.0) out.fallbacks++;
                t++;
            }
            break;
        }
        if (!dth_cur) {
            upload(dth, theta.data(), k);
            dth_cur = dth;
        }
        const double *dshift = per_column ? dth_cur : dlam;
        // 1. R = HX - theta X for every column (the head's residuals are measured too), W = T R for the moving ones;
        // three preconditioner sweeps W -> TMP -> W -> TMP (the result in TMP's buffer, which then becomes W); H W and
        // H P; 2. the Gram B = S^T S and S^T T with S = [X, W, P], T = [R, HW, HP] and ||r_i||^2 for the readback.
        // This device segment is one fixed launch sequence: with graphs on it is captured once per arrangement of
        // its buffers and sizes and replayed (every allocation and tuning it could need done before).
        const int m = k + na + np, q = 2 * m;
        ensure_gram(m, q);
        const int vj = variant(ST_JAC, na);
        variant(ST_HAM, np > 0 ? na + np : na);
        if (nl_rows_ > 0) ensure_nl(na + np);
        T *W0 = c.W.template as<T>(), *T0 = c.TMP.template as<T>();
        double *dG = dG_.as<double>();
        auto segment = [&] {
            residual_precond<T>(c.R.template as<T>(), W0, c.HX.template as<T>(), X(), dth_cur, v, dshift, diag_t, k, N_,
                                stream_);
            StencilParams jp{n_, na, h_, 1.0, omega, diag_t};
            T *const wl = W0 + (long)locked * N_, *const tl = T0 + (long)locked * N_;
            const T *const rl = c.R.template as<T>() + (long)locked * N_;
            stencil<T>(ST_JAC, jp, wl, tl, v, rl, dshift + locked, stream_, vj);
            stencil<T>(ST_JAC, jp, tl, wl, v, rl, dshift + locked, stream_, vj);
            stencil<T>(ST_JAC, jp, wl, tl, v, rl, dshift + locked, stream_, vj);
            hamiltonian2(v, tl, col(c.HW, locked), na, col(c.P, locked), col(c.HP, locked), np);
            Cols S{{{X(), k}, {tl, na}, {col(c.P, locked), np}}, np > 0 ? 3 : 2};
            Cols U{{{X(), k}, {tl, na}, {col(c.P, locked), np}, {c.R.ptr, k}, {col(c.HW, locked), na},
                    {col(c.HP, locked), np}}, 6};
            if (np == 0) U = Cols{{{X(), k}, {tl, na}, {c.R.ptr, k}, {col(c.HW, locked), na}}, 4};
            gram<T>(S, U, N_, dv_, gpart_.as<double>(), gpart_cap_, dG, stream_);
            RedJobs<T> jr{};
            for (int i = 0; i < k; i++) {
                jr.a[i % MAX_JOBS] = col(c.R, i);
                jr.b[i % MAX_JOBS] = jr.a[i % MAX_JOBS];
                jr.kind[i % MAX_JOBS] = RED_DOT;
                jr.count = i % MAX_JOBS + 1;
                if (jr.count == MAX_JOBS || i == k - 1) {
                    reduce_jobs<T>(jr, n_, dv_, partials_.as<double>(), dG + (size_t)m * q + (i / MAX_JOBS) * MAX_JOBS,
                                   stream_);
                    jr.count = 0;
                }
            }
        };
        const std::vector<uintptr_t> seg_key{
            (uintptr_t)X(), (uintptr_t)W0, (uintptr_t)T0, (uintptr_t)c.R.ptr, (uintptr_t)c.HX.ptr, (uintptr_t)c.HW.ptr,
            (uintptr_t)c.HP.ptr, (uintptr_t)c.P.ptr, (uintptr_t)v, (uintptr_t)dth_cur, (uintptr_t)dshift, (uintptr_t)dG,
            (uintptr_t)gpart_.ptr, (uintptr_t)partials_.ptr, (uintptr_t)nl_C_.ptr, (uintptr_t)nl_P_.ptr,
            (uintptr_t)beta_.ptr, (uintptr_t)k, (uintptr_t)na, (uintptr_t)np,
            (uintptr_t)locked, (uintptr_t)vj, (uintptr_t)nl_rows_, (uintptr_t)ionic_version_};
        run_graph(lobpcg_graph_[W0 < T0 ? 0 : 1], seg_key, segment);
        std::swap(c.W, c.TMP);  // the sweeps' result sits in TMP's buffer
        const double *g = readback(dG, (size_t)m * q + k);
        // A's X columns come through R: A_ij = S_i^T R_j + theta_j B_ij
        std::vector<double> Bm((size_t)m * m), Am((size_t)m * m);
        for (int i = 0; i < m; i++)
            for (int j = 0; j <= i; j++) {
                const double bij = g[(size_t)i * q + j];
                double aij = g[(size_t)i * q + m + j];
                if (j < k) aij += theta[j] * bij;
                Bm[(size_t)i * m + j] = Bm[(size_t)j * m + i] = bij;
                Am[(size_t)i * m + j] = Am[(size_t)j * m + i] = aij;
            }
        double head_res = 0.0, tail_res = 0.0, lock_res = 0.0;
        bool finite = true;
        const double tail_scale = tol / tail_tol;
        for (int i = 0; i < k; i++) {
            const double xx = Bm[(size_t)i * m + i];
            lambda[i] = Am[(size_t)i * m + i] / xx;  // the Rayleigh quotient of the stored x_i
            res[i] = std::sqrt(std::max(g[(size_t)m * q + i], 0.0) / xx);
            if (!std::isfinite(lambda[i]) || !std::isfinite(res[i])) finite = false;
            if (i >= checked) continue;  // an untested buffer column (it only speeds up the columns below it)
            if (i < head) head_res = std::max(head_res, res[i]);
            else {
                tail_res = std::max(tail_res, res[i]);
                lock_res = std::max(lock_res, res[i] * tail_scale);
            }
        }
        for (double x : Am)
            if (!std::isfinite(x)) finite = false;
        if (!finite) throw NumericalError("lobpcg: non-finite Gram entry or residual (NaN in a vector or in V)");
        out.max_residual = head_res;
        out.guard_residual = tail_res;
        const double split_res = std::max(head_res, tail_res * tail_scale);
        // at the guess, a head under its guess tolerance whose tested tail is not is locked
        if (lock && out.iterations == 0 && locked == 0 && head > 0 && head < checked && head_res < guess_tol &&
            !(split_res < guess_tol))
            locked = head;
        const double stop_res = locked > 0 ? lock_res : split_res;
        if (stop_res < tol || (out.iterations == 0 && stop_res < guess_tol) || out.iterations >= max_iter) break;
        if (inner_abort_ && inner_abort_(abort_user_)) {
            out.aborted = true;
            break;
        }
        out.iterations++;
        const int nm = k - locked;  // the columns this step moves (the lock may have been set above)
        // 3. Rayleigh–Ritz with the conditioning guard: a step whose coefficient amplification exceeds the budget drops P
        int mm = m;
        std::vector<double> th, Cc;
        double alpha;
        const double budget_res = locked > 0 ? tail_res : stop_res;
        const double alpha_budget = sizeof(T) == 4
                                        ? std::max(ALPHA_MAX, ALPHA_RESIDUAL_FRACTION * budget_res / std::max(floor, 1e-300))
                                        : 1e300;
        bool rr_failed = false;  // the subspace with P lost rank: retry without P (in every precision)
        try {
            alpha = rayleigh_ritz_locked(m, Am, Bm, locked, nm, RR_DROP, th, Cc);
        } catch (const NumericalError &) {
            if (np == 0) throw;
            rr_failed = true;
            alpha = 1e300;
        }
        if ((rr_failed || alpha > alpha_budget) && np > 0) {
            out.fallbacks++;
            mm = k + na;
            std::vector<double> As((size_t)mm * mm), Bs((size_t)mm * mm);
            for (int i = 0; i < mm; i++)
                for (int j = 0; j < mm; j++) {
                    As[(size_t)i * mm + j] = Am[(size_t)i * m + j];
                    Bs[(size_t)i * mm + j] = Bm[(size_t)i * m + j];
                }
            rayleigh_ritz_locked(mm, As, Bs, locked, nm, RR_DROP, th, Cc);
            Am.swap(As);
            Bm.swap(Bs);
        }
        if (Cc.size() != (size_t)mm * nm || th.size() != (size_t)nm)
            throw NumericalError("lobpcg: Rayleigh-Ritz returned no coefficients for the step");
        // 4. X_new = S C and P_new = X_new - X (X^T X)^-1 X^T X_new = S C_p (C_p = C - [Bxx^-1 Bxs C ; 0])
        std::vector<double> M((size_t)mm * 2 * nm, 0.0);
        for (int i = 0; i < mm; i++)
            for (int j = 0; j < nm; j++) M[(size_t)i * 2 * nm + j] = Cc[(size_t)i * nm + j];
        {
            std::vector<double> bxx((size_t)k * k), rhs((size_t)k * nm, 0.0);
            for (int i = 0; i < k; i++)
                for (int j = 0; j < k; j++) bxx[(size_t)i * k + j] = Bm[(size_t)i * mm + j];
            for (int i = 0; i < k; i++)
                for (int j = 0; j < nm; j++) {
                    double sacc = 0.0;
                    for (int a = 0; a < mm; a++) sacc += Bm[(size_t)i * mm + a] * Cc[(size_t)a * nm + j];
                    rhs[(size_t)i * nm + j] = sacc;
                }
            if (!cholesky(k, bxx)) throw NumericalError("lobpcg: Cholesky of X^T X failed (X lost rank)");
            chol_solve(k, bxx, rhs, nm);
            for (int i = 0; i < mm; i++)
                for (int j = 0; j < nm; j++)
                    M[(size_t)i * 2 * nm + nm + j] = Cc[(size_t)i * nm + j] - (i < k ? rhs[(size_t)i * nm + j] : 0.0);
        }
        for (int a = 0; a < nm; a++) theta[locked + a] = th[a];
        const int ldc = upload_coeffs(M, mm, 2 * nm, &theta, &dth_cur);
        const int np_used = mm > k + na ? np : 0;
        Cols Sm{{{X(), k}, {col(c.W, lk0), na}, {col(c.P, lk0), np_used}}, np_used > 0 ? 3 : 2};
        lincomb<T>(Sm, dC_.as<double>(), ldc, 2 * nm, col(c.XN, locked), N_, stream_, nm, col(c.PN, locked));
        if (locked > 0)  // the locked head is not written: it moves to the new block as it is
            LIRA_HIP(hipMemcpyAsync(c.XN.ptr, c.X.ptr, (size_t)N_ * locked * sizeof(T), hipMemcpyDeviceToDevice, stream_));
        std::swap(c.X, c.XN);
        std::swap(c.P, c.PN);
        np = nm;
        // 5. HX for the new columns (the locked head's is unchanged)
        hamiltonian(v, col(c.X, locked), col(c.HX, locked), nm);
    }
    out.head_kept = out.iterations == 0 || locked > 0;
    if (out.head_kept && out.iterations > 0) out.locked_calls = 1;
    c.evals.assign(lambda.begin(), lambda.end());
    c.res.assign(res.begin(), res.end());
    lobpcg_total_ += out.iterations;
    return out;
}

// Rearranges the first k columns of a channel's block (X, evals, res) so that new column j is old column order[j]
// (order.size() <= k columns are kept); device copies through the TMP block.
template <typename T> void Solver<T>::permute_columns(Channel<T> &c, const std::vector<int> &order) {
    const size_t bytes = (size_t)N_ * sizeof(T);
    for (size_t j = 0; j < order.size(); j++)
        LIRA_HIP(hipMemcpyAsync(c.TMP.template as<T>() + (long)j * N_, c.X.template as<T>() + (long)order[j] * N_, bytes,
                                hipMemcpyDeviceToDevice, stream_));
    std::swap(c.X, c.TMP);
    std::vector<double> ev(order.size()), rs(order.size());
    for (size_t j = 0; j < order.size(); j++) {
        ev[j] = c.evals[order[j]];
        rs[j] = c.res[order[j]];
    }
    c.evals = ev;
    c.res = rs;
    c.k = (int)order.size();
}

// Moves the block's highest Ritz value to the last position, the buffer's, when a lock or a 0-iteration accept left
// it elsewhere; returns the position the former last column moved to, or -1 (crates/qm-gpu/src/scf.rs:1031-1042).
template <typename T> int Solver<T>::buffer_last(Channel<T> &c) {
    const int last = c.k - 1;
    int top = last;
    for (int i = 0; i < c.k; i++)
        if (c.evals[i] > c.evals[top]) top = i;
    if (top == last) return -1;
    std::vector<int> order(c.k);
    std::iota(order.begin(), order.end(), 0);
    std::swap(order[top], order[last]);
    permute_columns(c, order);
    return top;
}

// One spin channel under Aufbau or Fermi–Dirac (qm/src/scf.rs:774-877; crates/qm-gpu/src/scf.rs:872-1033). Under
// Fermi–Dirac the block carries, above the tested levels, one untested buffer column (it speeds up the column below
// it, f = 0); the witness columns whose occupation is negligible against the SCF tolerance are held to the guard
// tolerance; on a loosened step the witness tail shrinks with the previous step's delta_rho (the step-accuracy tail);
// on a tight step a top tested witness with r above 1.5 guard yields to the lowest tested level that is a witness at
// r <= 1.5 guard, else the block grows by one; a buffer that clears both witness arms decides at r <= guard (one fresh
// buffer then joins), else the block grows by one. guess_band: a warm guess within 1.5 tol is kept (else only within
// tol). loosened_delta: the previous step's delta_rho on a loosened step, NaN on a tight one. Returns whether every
// call kept the tight columns bit for bit (the step's "frozen"); eo.aborted when the inner abort stopped a call.
template <typename T>
bool Solver<T>::solve_channel(int spin, const T *v, int nel, bool fd, double kt, double tol, double scf_tol,
                              bool guess_band, double loosened_delta, EigenOut &eo, Fd &occ) {
    Channel<T> &c = ch_[spin];
    if (c.k == 0) {
        occ = Fd{};
        return true;
    }
    auto account = [&](const EigenOut &e, int head, int checked) {
        eo.iterations += e.iterations;
        eo.fallbacks += e.fallbacks;
        eo.locked_calls += e.locked_calls;
        for (int i = 0; i < std::min(head, c.k); i++) eo.max_residual = std::max(eo.max_residual, c.res[i]);
        for (int i = head; i < std::min(checked, c.k); i++) eo.guard_residual = std::max(eo.guard_residual, c.res[i]);
    };
    if (!fd) {
        profile_.peak_block = std::max(profile_.peak_block, c.k);
        EigenOut e = lobpcg(c, v, c.k, NAN, tol, 200, tol, c.k, c.k, EIG_GUESS_FACTOR * tol, false);
        if (trace_on())
            std::fprintf(stderr, "lira:   spin %d k %d tol %.1e: %d iterations, max residual %.2e, fallbacks %d%s\n", spin,
                         c.k, tol, e.iterations, e.max_residual, e.fallbacks, e.head_kept ? " (kept)" : "");
        account(e, c.k, c.k);
        if (e.aborted) eo.aborted = true;
        occ.f.assign(c.k, 1.0);
        occ.mu = NAN;
        occ.entropy = 0.0;
        c.occ = occ.f;
        return e.head_kept;
    }
    const double guard = std::max(WITNESS_GUARD_TOL, tol);
    const double guess_tol = (guess_band ? EIG_GUESS_FACTOR : 1.0) * tol;
    const bool loosened = !std::isnan(loosened_delta);
    const int cap = witness_cap_;
    bool kept = true;
    auto buffered = [&](int len) { return len >= nel + 2 && len - nel < cap; };
    auto add_column = [&]() {
        ensure_channel(c, c.k + 1);
        seed_column(c.X.template as<T>() + (long)c.k * N_, c.k, spin);
        c.evals.push_back(c.evals.back());
        c.res.push_back(0.0);
        c.k++;
    };
    if (c.k == nel + 1 && cap >= 2) add_column();
    int swap_retries = 0;
    for (;;) {
        const int checked = buffered(c.k) ? c.k - 1 : c.k;
        const int head = std::min(std::max(c.tight, 0), checked);
        const double tail_kt = loosened ? step_tail_kt(c.k, loosened_delta) : WITNESS_TAIL_KT;
        profile_.peak_block = std::max(profile_.peak_block, c.k);
        EigenOut e = lobpcg(c, v, c.k, NAN, tol, 200, guard, head, checked, guess_tol, true);
        if (trace_on())
            std::fprintf(stderr,
                         "lira:   spin %d k %d head %d tested %d tol %.1e: %d iterations, residual %.2e / %.2e, "
                         "fallbacks %d%s\n",
                         spin, c.k, head, checked, tol, e.iterations, e.max_residual, e.guard_residual, e.fallbacks,
                         e.head_kept ? " (head kept)" : "");
        account(e, head, checked);
        if (e.aborted) {
            eo.aborted = true;
            return false;
        }
        if (checked < c.k) {
            // an untested column moved into the tested range is no witness until it meets the guard tolerance:
            // re-solve (it is tested now), at most twice per step
            const int moved = buffer_last(c);
            if (moved >= 0 && c.res[moved] > guard && swap_retries < 2) {
                swap_retries++;
                kept &= e.head_kept;
                continue;
            }
        }
        kept &= e.head_kept;
        // the tested levels are the first `checked` columns (the buffer, when there is one, is last)
        std::vector<double> ev(c.evals.begin(), c.evals.begin() + checked), rs(c.res.begin(), c.res.begin() + checked);
        occ = fermi_dirac(ev, nel, kt);
        const int need = tight_columns(ev, rs, occ.mu, kt, scf_tol);
        c.tight = need;
        if (need > head) continue;  // a loose column needs the eigen tolerance: re-solve with it tight
        step_tail_ = std::max(step_tail_, tail_kt);
        const bool witness = witness_bound_reached(ev, rs, nel, occ.mu, kt, tail_kt);
        auto o = ascending(ev);
        const double top_r = rs[o[checked - 1]];
        const double r_w = WITNESS_RESIDUAL_FACTOR * WITNESS_GUARD_TOL;
        int keep = -1;
        if (witness && !loosened && top_r > r_w && c.k - nel < cap) {
            keep = witness_keep_within(ev, rs, nel, occ.mu, kt, r_w, tail_kt);
            if (keep < 0) {
                add_column();
                profile_.witness_growths++;
                continue;
            }
        } else if (witness) {
            keep = witness_keep(ev, rs, nel, occ.mu, kt, tail_kt);
        }
        if (keep >= 0) {
            // the trim: keep the lowest `keep` tested levels (by value) plus the next one up as the buffer when one fits
            const int len = std::min(c.k, buffered(keep + 1) ? keep + 1 : keep);
            int chk = checked;
            if (len < c.k) {
                profile_.witness_trims++;
                const std::vector<int> all = ascending(c.evals);
                std::vector<bool> drop(c.k, false);
                for (int r = len; r < c.k; r++) drop[all[r]] = true;
                std::vector<int> order;
                for (int i = 0; i < c.k; i++)
                    if (!drop[i]) order.push_back(i);
                permute_columns(c, order);
                if (buffered(len)) buffer_last(c);
                chk = buffered(len) ? len - 1 : len;
                std::vector<double> ev2(c.evals.begin(), c.evals.begin() + chk);
                occ = fermi_dirac(ev2, nel, kt);
            }
            occ.f.resize(c.k, 0.0);  // the buffer holds no charge
            c.occ = occ.f;
            return kept;
        }
        // the buffer as the witness: it decides the step only at r <= guard; a buffer that clears both arms at a
        // larger residual grows the block by one (it is tested on the re-solve)
        if (checked < c.k && witness_bound_reached(c.evals, c.res, nel, occ.mu, kt, tail_kt)) {
            if (c.res[c.k - 1] <= WITNESS_GUARD_TOL) {
                occ.f.resize(c.k, 0.0);
                if (c.k - nel < cap) {
                    add_column();
                    occ.f.push_back(0.0);
                }
                c.occ = occ.f;
                return kept;
            }
            if (c.k - nel < cap) {
                add_column();
                profile_.witness_growths++;
                continue;
            }
        }
        if (c.k - nel >= cap)
            throw NumericalError("no gap witness within " + std::to_string(cap) + " extra levels of spin " +
                                 std::to_string(spin));
        const int extras = c.k - nel;
        const int target = std::min(std::max(2 * extras, extras + 1), cap);
        while (c.k - nel < target) add_column();
        profile_.witness_growths++;
        doublings_++;
    }
}

template <typename T> void Solver<T>::upload_nuclei(const lira_nucleus *nuc, int nat) {
    nat_ = nat;
    pos_h_.assign(3 * nat, 0.0);
    z_h_.assign(nat, 0.0);
    for (int i = 0; i < nat; i++) {
        for (int a = 0; a < 3; a++) pos_h_[3 * i + a] = nuc[i].pos[a];
        z_h_[i] = nuc[i].z;
    }
    if (pos_.bytes < (size_t)(3 * nat + 1) * sizeof(double)) pos_.alloc((size_t)(3 * nat + 1) * sizeof(double));
    if (zf_.bytes < (size_t)(nat + 1) * sizeof(double)) zf_.alloc((size_t)(nat + 1) * sizeof(double));
    if (nat > 0) {
        upload(pos_.as<double>(), pos_h_.data(), 3 * nat);
        upload(zf_.as<double>(), z_h_.data(), nat);
    }
}

// ------------------------------------------------------------------------------------ pseudised ions ----
template <typename T> void Solver<T>::set_ionic(const lira_species *sp, int nsp) {
    if (nsp < 0 || (nsp > 0 && !sp)) throw InvalidError("lira_set_ionic: bad species list");
    std::vector<SpeciesHost> out;
    std::vector<double> tab;
    auto put = [&](const double *f, int len, const char *what) {
        if (len < 0 || (len > 0 && !f)) throw InvalidError(std::string("lira_set_ionic: bad table ") + what);
        for (int i = 0; i < len; i++)
            if (!std::isfinite(f[i])) throw InvalidError(std::string("lira_set_ionic: non-finite table ") + what);
        const int off = (int)tab.size();
        tab.insert(tab.end(), f, f + len);
        return off;
    };
    for (int i = 0; i < nsp; i++) {
        const lira_species &s = sp[i];
        if (s.z < 1 || s.z > 118 || !(s.z_v > 0.0) || !(s.sigma_lr > 0.0) || !(s.du > 0.0) || !(s.r_sr >= 0.0))
            throw InvalidError("lira_set_ionic: bad species scalars");
        for (const auto &o : out)
            if (o.z == s.z) throw InvalidError("lira_set_ionic: a Z listed twice");
        if (s.n_projectors < 0 || s.n_projectors > NL_MAX_PROJ || (s.n_projectors > 0 && !s.projectors))
            throw InvalidError("lira_set_ionic: at most 3 projectors per species");
        SpeciesHost h{};
        h.z = s.z;
        h.z_v = s.z_v;
        h.sigma_lr = s.sigma_lr;
        h.du = s.du;
        h.r_sr = s.r_sr;
        if (!(s.r_core >= 0.0)) throw InvalidError("lira_set_ionic: r_core must be >= 0");
        h.r_core = s.r_core;
        h.sr_len = s.n_v_sr;
        h.sr_off = put(s.v_sr_u, s.n_v_sr, "v_sr");
        const bool core = s.rho_pc_u && s.n_rho_pc > 0 && s.pc_radius > 0.0;
        h.pc_r = core ? s.pc_radius : 0.0;
        h.pc_len = core ? s.n_rho_pc : 0;
        h.pc_off = core ? put(s.rho_pc_u, s.n_rho_pc, "rho_pc") : 0;
        h.nproj = s.n_projectors;
        int rows = 0;
        for (int j = 0; j < s.n_projectors; j++) {
            const lira_projector &pr = s.projectors[j];
            if (pr.l < 0 || pr.l > 2 || !(pr.radius >= 0.0) || !std::isfinite(pr.e_kb))
                throw InvalidError("lira_set_ionic: projectors need 0 <= l <= 2, radius >= 0 and a finite E_KB");
            h.l[j] = pr.l;
            h.e_kb[j] = pr.e_kb;
            h.radius[j] = pr.radius;
            h.t_len[j] = pr.n_t;
            h.t_off[j] = put(pr.t_u, pr.n_t, "t_u");
            rows += 2 * pr.l + 1;
        }
        if (rows > NL_MAX_ROWS) throw InvalidError("lira_set_ionic: more than 9 projector rows on one species");
        out.push_back(h);
    }
    LIRA_HIP(hipStreamSynchronize(stream_));
    species_ = std::move(out);
    tab_h_ = std::move(tab);
    if (!tab_h_.empty()) {
        if (tab_.bytes < tab_h_.size() * sizeof(double)) tab_.alloc(tab_h_.size() * sizeof(double));
        copy_now(tab_.ptr, tab_h_.data(), tab_h_.size() * sizeof(double), hipMemcpyHostToDevice);
    }
    ionic_version_++;
    nl_rows_ = 0;
    has_pc_ = false;
}

// The ionic terms of the uploaded nuclei (qm/src/scf.rs:720-745): V_loc into vnuc_, rho_pc into pc_, the projector
// boxes and beta rows, and the colouring of the boxes into disjoint sets for the scatter.
template <typename T> void Solver<T>::ionic_geometry(double sigma) {
    const int nat = nat_;
    q_h_.assign(nat, 0.0);
    std::vector<IonicCentre> cs(std::max(nat, 1));
    has_pc_ = false;
    nl_h_.clear();
    nl_atom_.clear();
    nl_e_h_.clear();
    colours_.clear();
    nl_rows_ = 0;
    long beta_total = 0;
    const double o = -((double)n_ - 1.0) * h_ / 2.0;
    for (int a = 0; a < nat; a++) {
        const SpeciesHost *sp = species_of((int)z_h_[a]);
        IonicCentre c{};
        for (int d = 0; d < 3; d++) c.pos[d] = pos_h_[3 * a + d];
        c.q = sp ? sp->z_v : z_h_[a];
        c.sigma = sp ? sp->sigma_lr : sigma;
        c.du = sp ? sp->du : 1.0;
        if (sp) {
            c.sr_off = sp->sr_off;
            c.sr_len = sp->sr_len;
            c.r_sr = sp->r_sr;
            c.pc_off = sp->pc_off;
            c.pc_len = sp->pc_len;
            c.pc_r = sp->pc_r;
            has_pc_ = has_pc_ || sp->pc_len > 0;
        }
        q_h_[a] = c.q;
        cs[a] = c;
        if (!sp || sp->nproj == 0) continue;
        // the box of the largest projector ball, clipped to the grid (qm/src/nuclear.rs:322-337)
        double rmax = 0.0;
        for (int j = 0; j < sp->nproj; j++) rmax = std::max(rmax, sp->radius[j]);
        NlIon ion{};
        bool empty = false;
        for (int d = 0; d < 3; d++) {
            double lo = std::floor((c.pos[d] - rmax - o) / h_), hi = std::ceil((c.pos[d] + rmax - o) / h_);
            lo = std::max(lo, 0.0);
            hi = std::min(hi, (double)(n_ - 1));
            if (!(hi >= lo)) empty = true;
            ion.lo[d] = empty ? 0 : (int)lo;
            ion.dim[d] = empty ? 0 : (int)hi - (int)lo + 1;
        }
        if (empty) continue;
        for (int d = 0; d < 3; d++) ion.pos[d] = c.pos[d];
        ion.du = sp->du;
        ion.pts = (long)ion.dim[0] * ion.dim[1] * ion.dim[2];
        ion.beta_off = beta_total;
        ion.nproj = sp->nproj;
        ion.rows = 0;
        for (int j = 0; j < sp->nproj; j++) {
            ion.l[j] = sp->l[j];
            ion.t_off[j] = sp->t_off[j];
            ion.t_len[j] = sp->t_len[j];
            ion.radius[j] = sp->radius[j];
            ion.e_kb[j] = sp->e_kb[j];
            for (int m = 0; m < 2 * sp->l[j] + 1; m++) nl_e_h_.push_back(sp->e_kb[j]);
            ion.rows += 2 * sp->l[j] + 1;
        }
        ion.row_off = nl_rows_;
        nl_rows_ += ion.rows;
        beta_total += (long)ion.rows * ion.pts;
        nl_h_.push_back(ion);
        nl_atom_.push_back(a);
    }
    LIRA_HIP(hipStreamSynchronize(stream_));
    if (centres_.bytes < cs.size() * sizeof(IonicCentre)) centres_.alloc(cs.size() * sizeof(IonicCentre));
    copy_now(centres_.ptr, cs.data(), cs.size() * sizeof(IonicCentre), hipMemcpyHostToDevice);
    if (qf_.bytes < (size_t)(nat + 1) * sizeof(double)) qf_.alloc((size_t)(nat + 1) * sizeof(double));
    if (nat > 0) copy_now(qf_.ptr, q_h_.data(), nat * sizeof(double), hipMemcpyHostToDevice);
    alloc_field(vnuc_);
    if (has_pc_) alloc_field(pc_);
    ionic_fields<T>(vnuc_.template as<T>(), has_pc_ ? pc_.template as<T>() : nullptr, n_, h_,
                    centres_.as<IonicCentre>(), nat, tab_.as<double>(), stream_);
    if (nl_rows_ == 0) return;
    // colours: greedy, an ion joins the first colour none of whose boxes its box overlaps
    std::vector<std::vector<int>> groups;
    auto overlap = [&](const NlIon &a, const NlIon &b) {
        for (int d = 0; d < 3; d++)
            if (a.lo[d] + a.dim[d] <= b.lo[d] || b.lo[d] + b.dim[d] <= a.lo[d]) return false;
        return true;
    };
    for (int i = 0; i < (int)nl_h_.size(); i++) {
        bool placed = false;
        for (auto &g : groups) {
            bool ok = true;
            for (int j : g) ok = ok && !overlap(nl_h_[i], nl_h_[j]);
            if (ok) {
                g.push_back(i);
                placed = true;
                break;
            }
        }
        if (!placed) groups.push_back({i});
    }
    std::vector<int> members;
    for (auto &g : groups) {
        Colour c{(int)members.size(), (int)g.size(), 0};
        for (int i : g) {
            members.push_back(i);
            c.max_pts = std::max(c.max_pts, nl_h_[i].pts);
        }
        colours_.push_back(c);
    }
    if (nl_ions_.bytes < nl_h_.size() * sizeof(NlIon)) nl_ions_.alloc(nl_h_.size() * sizeof(NlIon));
    copy_now(nl_ions_.ptr, nl_h_.data(), nl_h_.size() * sizeof(NlIon), hipMemcpyHostToDevice);
    if (nl_members_.bytes < members.size() * sizeof(int)) nl_members_.alloc(members.size() * sizeof(int));
    copy_now(nl_members_.ptr, members.data(), members.size() * sizeof(int), hipMemcpyHostToDevice);
    if (nl_erow_.bytes < nl_e_h_.size() * sizeof(double)) nl_erow_.alloc(nl_e_h_.size() * sizeof(double));
    copy_now(nl_erow_.ptr, nl_e_h_.data(), nl_e_h_.size() * sizeof(double), hipMemcpyHostToDevice);
    if (beta_.bytes < (size_t)beta_total * sizeof(T)) beta_.alloc((size_t)beta_total * sizeof(T));
    for (const NlIon &ion : nl_h_) nl_beta<T>(beta_.template as<T>(), ion, n_, h_, tab_.as<double>(), stream_);
}

// out_b += V_NL x_b for a block: the projections C (f64, rows x nb), then the scatter colour by colour
template <typename T> void Solver<T>::ensure_nl(int nb) {
    const size_t need = (size_t)nl_rows_ * nb * 4;
    if (need > nl_C_cap_) {
        nl_C_.alloc(need * sizeof(double));
        nl_C_cap_ = need;
    }
    long max_pts = 0;
    for (const NlIon &ion : nl_h_) max_pts = std::max(max_pts, ion.pts);
    const size_t work = (size_t)nl_rows_ * nb * nl_project_chunks(max_pts);
    if (work > nl_P_cap_) {
        nl_P_.alloc(work * sizeof(double));
        nl_P_cap_ = work;
    }
}

template <typename T> void Solver<T>::nl_projections(const T *x, int n1, const T *x2, int n2) {
    ensure_nl(n1 + n2);
    long max_pts = 0;
    for (const NlIon &ion : nl_h_) max_pts = std::max(max_pts, ion.pts);
    nl_project<T>(x, n1, x2, n2, N_, n_, beta_.template as<T>(), nl_ions_.as<NlIon>(), (int)nl_h_.size(), max_pts,
                  nl_rows_, dv_, nl_C_.as<double>(), nl_P_.as<double>(), stream_);
}

template <typename T> void Solver<T>::nl_apply_block(const T *x, T *out, int nb, const T *x2, T *out2, int nb2) {
    nl_projections(x, nb, x2, nb2);
    for (const Colour &c : colours_)
        nl_apply<T>(out, nb, out2, nb2, N_, n_, beta_.template as<T>(), nl_ions_.as<NlIon>(),
                    nl_members_.as<int>() + c.off, c.count, c.max_pts, nl_C_.as<double>(), nl_erow_.as<double>(), stream_);
}

// E_NL terms f_b E_r c_rb^2 of a channel in (orbital, sphere, m) order; Aufbau: every orbital at f = 1
// (qm/src/nonlocal.rs:148-172)
template <typename T> std::vector<double> Solver<T>::nl_energy_terms(Channel<T> &c, bool fd) {
    std::vector<double> terms;
    if (nl_rows_ == 0 || c.k == 0) return terms;
    nl_projections(c.X.template as<T>(), c.k);
    const double *C = readback(nl_C_.as<double>(), (size_t)nl_rows_ * c.k);
    for (int b = 0; b < c.k; b++) {
        const double f = fd ? c.occ[b] : 1.0;
        if (fd && f == 0.0) continue;
        for (int r = 0; r < nl_rows_; r++) {
            const double cr = C[(long)r * c.k + b];
            terms.push_back(fd ? f * nl_e_h_[r] * cr * cr : nl_e_h_[r] * cr * cr);
        }
    }
    return terms;
}

template <typename T>
void Solver<T>::ionic_fields_host(const lira_nucleus *nuc, int nat, double sigma, double *vloc, double *pc) {
    if (nat < 0 || (nat > 0 && !nuc) || !(sigma > 0.0)) throw InvalidError("lira_ionic_fields: bad arguments");
    upload_nuclei(nuc, nat);
    if (species_.empty()) throw InvalidError("lira_ionic_fields: no ionic setup");
    ionic_geometry(sigma);
    if (rho64_.bytes < N_ * sizeof(double)) rho64_.alloc(N_ * sizeof(double));
    for (int w = 0; w < 2; w++) {
        double *dst = w == 0 ? vloc : pc;
        if (!dst) continue;
        if (w == 1 && !has_pc_) {
            std::memset(dst, 0, N_ * sizeof(double));
            continue;
        }
        convert<double, T>(rho64_.as<double>(), (w == 0 ? vnuc_ : pc_).template as<T>(), N_, stream_);
        LIRA_HIP(hipMemcpyAsync(dst, rho64_.ptr, N_ * sizeof(double), hipMemcpyDeviceToHost, stream_));
        LIRA_HIP(hipStreamSynchronize(stream_));
    }
}

template <typename T>
void Solver<T>::nonlocal_apply(const double *x, int nb, double *out, double *c, int c_cap, int *rows) {
    if (nb < 1 || !x || !out) throw InvalidError("lira_nonlocal_apply: bad arguments");
    if (rows) *rows = nl_rows_;
    DevBuf x64((size_t)N_ * nb * sizeof(double)), xt((size_t)N_ * nb * sizeof(T)), ot((size_t)N_ * nb * sizeof(T));
    LIRA_HIP(hipMemcpyAsync(x64.ptr, x, (size_t)N_ * nb * sizeof(double), hipMemcpyHostToDevice, stream_));
    convert<T, double>(xt.template as<T>(), x64.as<double>(), N_ * nb, stream_);
    fill<T>(ot.template as<T>(), N_ * nb, 0.0, stream_);
    if (nl_rows_ > 0) nl_apply_block(xt.template as<T>(), ot.template as<T>(), nb);
    convert<double, T>(x64.as<double>(), ot.template as<T>(), N_ * nb, stream_);
    LIRA_HIP(hipMemcpyAsync(out, x64.ptr, (size_t)N_ * nb * sizeof(double), hipMemcpyDeviceToHost, stream_));
    LIRA_HIP(hipStreamSynchronize(stream_));
    if (c && nl_rows_ > 0) {
        const int cnt = std::min(c_cap, nl_rows_ * nb);
        copy_now(c, nl_C_.ptr, cnt * sizeof(double), hipMemcpyDeviceToHost);
    }
}

// The cold density: per centre a free atom (all-electron centres, free_atoms) rescaled so that its grid sum is its
// charge, else a sigma = 1 blob of its charge; rho_in_[s] = scale_s * that sum (qm/src/nuclear.rs:153-231)
template <typename T>
void Solver<T>::cold_density(bool free_atoms, double sigma, const double *charges, double scale0, double scale1) {
    const int nat = nat_;
    std::vector<AtomProfile> atoms;
    std::vector<double> tab, blobs;
    for (int a = 0; a < nat; a++) {
        const bool ae = species_of((int)z_h_[a]) == nullptr;
        RadialProfile p;
        if (free_atoms && ae) p = free_atom_profile((int)z_h_[a], sigma);
        if (p.ok) {
            AtomProfile A{};
            for (int d = 0; d < 3; d++) A.pos[d] = pos_h_[3 * a + d];
            A.x0 = p.x0;
            A.dx = p.dx;
            A.w = charges[a];  // replaced by charge / (grid sum) below
            A.off = (int)tab.size();
            A.len = (int)p.rho.size();
            tab.insert(tab.end(), p.rho.begin(), p.rho.end());
            atoms.push_back(A);
        } else {
            for (int d = 0; d < 3; d++) blobs.push_back(pos_h_[3 * a + d]);
            blobs.push_back(charges[a]);
        }
    }
    LIRA_HIP(hipStreamSynchronize(stream_));
    if (!tab.empty()) {
        if (atom_tab_.bytes < tab.size() * sizeof(double)) atom_tab_.alloc(tab.size() * sizeof(double));
        copy_now(atom_tab_.ptr, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice);
    }
    if (atoms_dev_.bytes < (atoms.size() + 1) * sizeof(AtomProfile))
        atoms_dev_.alloc((atoms.size() + 1) * sizeof(AtomProfile));
    // per-atom grid sums with unit weight, then w = charge / (dv * sum); an atom whose samples sum to nothing
    // (outside the box) keeps its blob
    std::vector<AtomProfile> kept;
    for (AtomProfile A : atoms) {
        const double q = A.w;
        A.w = 1.0;
        copy_now(atoms_dev_.ptr, &A, sizeof(AtomProfile), hipMemcpyHostToDevice);
        atom_density<T>(rho_t_.template as<T>(), n_, h_, atoms_dev_.as<AtomProfile>(), 1, atom_tab_.as<double>(),
                        nullptr, 0, 1.0, stream_);
        RedJobs<T> j{};
        j.a[0] = j.b[0] = rho_t_.template as<T>();
        j.kind[0] = RED_SUM;
        j.count = 1;
        const double sum = reduce(j, dv_)[0];
        if (std::isfinite(sum) && sum > 0.0) {
            A.w = q / sum;
            kept.push_back(A);
        } else {
            for (int d = 0; d < 3; d++) blobs.push_back(A.pos[d]);
            blobs.push_back(q);
        }
    }
    if (!kept.empty())
        copy_now(atoms_dev_.ptr, kept.data(), kept.size() * sizeof(AtomProfile), hipMemcpyHostToDevice);
    if (blobs_dev_.bytes < (blobs.size() + 4) * sizeof(double)) blobs_dev_.alloc((blobs.size() + 4) * sizeof(double));
    if (!blobs.empty())
        copy_now(blobs_dev_.ptr, blobs.data(), blobs.size() * sizeof(double), hipMemcpyHostToDevice);
    for (int s = 0; s < 2; s++)
        atom_density<T>(rho_in_[s].template as<T>(), n_, h_, atoms_dev_.as<AtomProfile>(), (int)kept.size(),
                        atom_tab_.as<double>(), blobs_dev_.as<double>(), (int)blobs.size() / 4, s == 0 ? scale0 : scale1,
                        stream_);
}

// The first columns of each channel's cold block: the bound free-atom levels of every centre (a pseudised centre's
// above its core), lowest first (the LCG columns beyond them stay). Only the starting subspace changes: LOBPCG
// converges to the same levels from any block that is not deficient in them.
template <typename T> void Solver<T>::seed_atomic_orbitals(double sigma) {
    struct Ao {
        int atom, orb, m;
        double eps;
    };
    ao_cols_.clear();
    ao_lm_.clear();
    std::vector<RadialProfile> profs(nat_);
    std::vector<Ao> list;
    for (int a = 0; a < nat_; a++) {
        // a pseudised centre starts from its all-electron atom's levels above the Z - Z_v core electrons
        const SpeciesHost *sp = species_of((int)z_h_[a]);
        profs[a] = free_atom_profile((int)z_h_[a], sigma);
        if (!profs[a].ok) continue;
        double core = sp ? z_h_[a] - sp->z_v : 0.0;
        for (int i = 0; i < (int)profs[a].orbitals.size(); i++) {
            const auto &o = profs[a].orbitals[i];
            const double shell = 2.0 * (2 * o.l + 1);
            if (core > 0.5) {
                core -= shell;
                continue;
            }
            for (int m = -o.l; m <= o.l; m++) list.push_back({a, i, m, o.eps});
        }
    }
    if (list.empty()) return;
    std::stable_sort(list.begin(), list.end(), [](const Ao &x, const Ao &y) { return x.eps < y.eps; });
    std::vector<double> tab;
    std::map<std::pair<int, int>, int> placed;
    for (const Ao &e : list) {
        auto key = std::make_pair(e.atom, e.orb);
        auto it = placed.find(key);
        if (it == placed.end()) {
            const auto &R = profs[e.atom].orbitals[e.orb].R;
            it = placed.emplace(key, (int)tab.size()).first;
            tab.insert(tab.end(), R.begin(), R.end());
        }
        const RadialProfile &p = profs[e.atom];
        AtomProfile A{};
        for (int d = 0; d < 3; d++) A.pos[d] = pos_h_[3 * e.atom + d];
        A.x0 = p.x0;
        A.dx = p.dx;
        A.w = 1.0;
        A.off = it->second;
        A.len = (int)p.orbitals[e.orb].R.size();
        ao_cols_.push_back(A);
        ao_lm_.push_back({p.orbitals[e.orb].l, e.m});
    }
    LIRA_HIP(hipStreamSynchronize(stream_));
    if (ao_tab_.bytes < tab.size() * sizeof(double)) ao_tab_.alloc(tab.size() * sizeof(double));
    copy_now(ao_tab_.ptr, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice);
    for (int s = 0; s < 2; s++) {
        Channel<T> &c = ch_[s];
        for (int col = 0; col < std::min(c.k, (int)ao_cols_.size()); col++)
            seed_column(c.X.template as<T>() + (long)col * N_, col, s);
    }
}

// column col of a block: the col-th free-atom orbital while there is one, else the reference's seeded LCG vector
template <typename T> void Solver<T>::seed_column(T *dst, int col, int spin) {
    if (col < (int)ao_cols_.size())
        atom_orbital<T>(dst, n_, h_, ao_cols_[col], ao_lm_[col].first, ao_lm_[col].second, ao_tab_.as<double>(),
                        stream_);
    else
        initial_orbitals<T>(dst, n_, h_, pos_.as<double>(), nat_, 1, 1000ULL + 2ULL * (uint64_t)col + (uint64_t)spin,
                            stream_);
}

template <typename T> void Solver<T>::sample(const double *pts, int count, double *out) {
    if (count < 0 || (count > 0 && (!pts || !out))) throw InvalidError("lira_sample_density: bad arguments");
    if (!valid_) throw InvalidError("lira_sample_density: no converged state");
    if (count == 0) return;
    if (pts_dev_.bytes < (size_t)count * 3 * sizeof(double)) pts_dev_.alloc((size_t)count * 3 * sizeof(double));
    if (samp_dev_.bytes < (size_t)count * sizeof(double)) samp_dev_.alloc((size_t)count * sizeof(double));
    LIRA_HIP(hipMemcpyAsync(pts_dev_.ptr, pts, (size_t)count * 3 * sizeof(double), hipMemcpyHostToDevice, stream_));
    sample_density<T>(rho_out_[0].template as<T>(), rho_out_[1].template as<T>(), n_, h_, pts_dev_.as<double>(), count,
                      samp_dev_.as<double>(), stream_);
    LIRA_HIP(hipMemcpyAsync(out, samp_dev_.ptr, (size_t)count * sizeof(double), hipMemcpyDeviceToHost, stream_));
    LIRA_HIP(hipStreamSynchronize(stream_));
}

// ----------------------------------------------------------------------------------------------- SCF ----
template <typename T>
lira_status Solver<T>::solve(const lira_system *sys, const lira_settings *st, int warm, lira_abort_fn abort,
                             lira_abort_fn inner_abort, lira_observe_fn observe, void *user, lira_scf_stats *out) {
    auto t0 = Clock::now();
    std::memset(out, 0, sizeof *out);
    readbacks_ = 0;
    cg_total_ = 0;
    lobpcg_total_ = 0;
    double t_poisson = 0.0, t_eigen = 0.0;
    profile_ = lira_profile{};
    drho_steps_.clear();
    etol_steps_.clear();
    const bool had_prev = have_prev_;
    have_prev_ = false;  // re-established only by a converged warm solve
    const int nat = sys->n_nuclei;
    if (nat < 0 || (nat > 0 && !sys->nuclei) || sys->n_up < 0 || sys->n_dn < 0)
        throw InvalidError("lira_solve: bad system");
    if (!(st->sigma > 0.0) || !(st->tol > 0.0) || st->max_iter < 1 || st->diis_history > LIRA_MAX_DIIS_HISTORY)
        throw InvalidError("lira_solve: bad settings");
    // nucleus-nucleus repulsion and the coincidence check (qm/src/nuclear.rs:277-289; with valence charges under an
    // ionic setup, nuclear.rs:371-382): the warm state is kept
    const bool ionic = !species_.empty();
    auto charge = [&](int z) {
        const SpeciesHost *sp = species_of(z);
        return sp ? sp->z_v : (double)z;
    };
    double e_nn = 0.0;
    for (int i = 0; i < nat; i++)
        for (int j = i + 1; j < nat; j++) {
            double d2 = 0.0;
            for (int a = 0; a < 3; a++) d2 += std::pow(sys->nuclei[i].pos[a] - sys->nuclei[j].pos[a], 2);
            const double d = std::sqrt(d2);
            if (d < 1e-8) throw std::pair<int, int>(i, j);
            e_nn += charge(sys->nuclei[i].z) * charge(sys->nuclei[j].z) / d;
        }
    // cores that overlap (qm/src/ionic.rs:175-185): reported, the solve goes on
    std::vector<lira_core_overlap> overlaps;
    if (ionic)
        for (int i = 0; i < nat; i++)
            for (int j = i + 1; j < nat; j++) {
                const SpeciesHost *a = species_of(sys->nuclei[i].z), *b = species_of(sys->nuclei[j].z);
                const double rc = (a ? a->r_core : 0.0) + (b ? b->r_core : 0.0);
                double d2 = 0.0;
                for (int x = 0; x < 3; x++) d2 += std::pow(sys->nuclei[i].pos[x] - sys->nuclei[j].pos[x], 2);
                if (std::sqrt(d2) < rc) overlaps.push_back(lira_core_overlap{i, j, std::sqrt(d2)});
            }
    const bool fd = st->occupations == LIRA_FERMI_DIRAC;
    const double kt = fd ? st->kt : 0.0;
    // the warm guard: a converged state, or the one an aborted solve parked, with the same counts and scheme
    const bool reuse = warm && (valid_ || parked_) && nel_[0] == sys->n_up && nel_[1] == sys->n_dn &&
                       occ_scheme_ == st->occupations && (!fd || kt_ == st->kt) && nat_ == nat &&
                       state_ionic_version_ == ionic_version_;
    const bool prev_converged = valid_;
    upload_nuclei(sys->nuclei, nat);
    sigma_ = st->sigma;
    const int nel[2] = {sys->n_up, sys->n_dn};
    const bool restricted = sys->n_up == sys->n_dn;
    valid_ = false;
    parked_ = false;
    // fields
    for (DevBuf *b : {&vnuc_, &vh_t_, &rho_in_[0], &rho_in_[1], &rho_out_[0], &rho_out_[1], &vxc_[0], &vxc_[1], &exc_,
                      &veff_[0], &veff_[1], &rho_t_})
        alloc_field(*b);
    if (vh_.bytes < N_ * sizeof(double)) {
        vh_.alloc(N_ * sizeof(double));
        fill<double>(vh_.as<double>(), N_, 0.0, stream_);
    }
    if (rho64_.bytes < N_ * sizeof(double)) rho64_.alloc(N_ * sizeof(double));
    if (ionic) {
        ionic_geometry(st->sigma);  // V_loc, rho_pc, projectors (qm/src/scf.rs:733-745)
    } else {
        nuclear_potential<T>(vnuc_.template as<T>(), n_, h_, pos_.as<double>(), zf_.as<double>(), nat, st->sigma, stream_);
        nl_rows_ = 0;
        has_pc_ = false;
    }
    const T *pcp = has_pc_ ? pc_.template as<T>() : nullptr;
    ao_cols_.clear();
    ao_lm_.clear();
    if (!reuse) {
        // cold start (qm/src/scf.rs:948-965): sigma = 1 blobs of each centre's charge scaled to the electron count,
        // seeded orbitals, V_H = 0
        double ztot = 0.0;
        for (int i = 0; i < nat; i++) ztot += charge(sys->nuclei[i].z);
        const double total = nel[0] + nel[1];
        const double scale = ztot > 0.0 ? total / ztot : 0.0;
        for (int s = 0; s < 2; s++) {
            const double f = total > 0.0 ? nel[s] / total : 0.0;
            ch_[s].limit = nel[s] + witness_cap_ + 1;
            if (!fd)
                deposit_charge<T>(rho_in_[s].template as<T>(), n_, h_, pos_.as<double>(),
                                  (ionic ? qf_ : zf_).template as<double>(), nat, 1.0, scale * f, stream_);
            Channel<T> &c = ch_[s];
            const int k = nel[s] + (fd && nel[s] > 0 ? 1 : 0);
            c.k = 0;
            ensure_channel(c, std::max(k, 1));
            c.k = k;
            if (k > 0) initial_orbitals<T>(c.X.template as<T>(), n_, h_, pos_.as<double>(), nat, k, (uint64_t)(s + 1), stream_);
            c.evals.assign(k, -1.0);
            c.res.assign(k, 0.0);
            c.occ.assign(k, 1.0);
            c.tight = nel[s];
        }
        if (fd) {
            // Fermi-Dirac: all-electron centres start from their sigma-smeared free atoms, pseudised ones (and any
            // atom that does not solve) from sigma = 1 blobs of their charge (qm/src/nuclear.rs:198-231)
            std::vector<double> q(nat);
            for (int i = 0; i < nat; i++) q[i] = charge(sys->nuclei[i].z);
            cold_density(true, st->sigma, q.data(), scale * (total > 0.0 ? nel[0] / total : 0.0),
                         scale * (total > 0.0 ? nel[1] / total : 0.0));
            seed_atomic_orbitals(st->sigma);
        }
        fill<double>(vh_.as<double>(), N_, 0.0, stream_);
    } else {
        for (int s = 0; s < 2; s++) {
            ch_[s].limit = nel[s] + witness_cap_ + 1;
            // the tight count starts from the warm state's occupations (qm/src/occupations.rs:352-366)
            if (fd) ch_[s].tight = tight_start(ch_[s].occ, nel[s], st->tol);
            if (extrapolate_) {  // keep rho_t: it becomes rho_{t-1} once this solve converges
                alloc_field(save_rho_[s]);
                copy<T>(save_rho_[s].template as<T>(), rho_out_[s].template as<T>(), N_, stream_);
            }
            if (extrapolate_ && had_prev && prev_converged)
                extrapolate<T>(rho_in_[s].template as<T>(), rho_out_[s].template as<T>(), prev_rho_[s].template as<T>(),
                               N_, stream_);
            else
                copy<T>(rho_in_[s].template as<T>(), rho_out_[s].template as<T>(), N_, stream_);
        }
        profile_.extrapolated = extrapolate_ && had_prev && prev_converged ? 1 : 0;
    }
    tune_stencil(std::max(ch_[0].k, 1));
    profile_.start_wall = secs(t0);
    auto record = [&](int steps, const EigenOut &e) {
        profile_.steps = steps;
        profile_.readbacks = readbacks_;
        profile_.poisson_iterations = cg_total_;
        profile_.eigen_iterations = lobpcg_total_;
        profile_.eigen_fallbacks = e.fallbacks;
        profile_.locked_calls = e.locked_calls;
        profile_.poisson_wall = t_poisson;
        profile_.eigen_wall = t_eigen;
        profile_.eig_tol_used = etol_steps_.empty() ? 0.0 : etol_steps_.back();
        profile_.wall = secs(t0);
    };
    const double ptol = std::min(1e-10, 1e-3 * st->tol);  // qm/src/scf.rs:519
    const double lo = effective_eig_tol(st->eig_tol), hi = std::max(EIG_TOL_MAX, lo);
    // the partial-level stop of the f32 solve (crates/qm-gpu/src/scf.rs:1519-1521); the f64 solve has no
    // occupation-noise floor and keeps settings.tol, as the CPU solver does
    const double partial_tol = sizeof(T) == 4 ? GPU_PARTIAL_TOL : 0.0;
    // DIIS (crates/qm-gpu/src/scf.rs:563-646): H entries, one more slot for the entry being pushed
    const int H = std::max(st->diis_history, 1), H1 = H + 1;
    if ((int)slots_.size() < H1) slots_.resize(H1);
    for (auto &sl : slots_)
        for (int s = 0; s < 2; s++) {
            alloc_field(sl.rho[s]);
            alloc_field(sl.res[s]);
        }
    std::deque<int> order;                           // DIIS history, oldest first (slot indices)
    std::vector<double> gram((size_t)H1 * H1, 0.0);  // <r_i, r_j> by slot
    double delta_rho = INFINITY, test_tol = st->tol;
    // the stall watch (qm/src/scf.rs:442-480)
    double anchor = INFINITY, best = INFINITY;
    int anchor_at = 0;
    const double level = STALL_LEVEL_PER_ELECTRON * (nel[0] + nel[1]);
    Fd occ[2];
    EigenOut eo;
    double max_pres = 0.0;
    int max_pit = 0;
    int it = 0;
    bool converged = false, stalled = false, abandoned = false, aborted_in_step = false;
    bool frozen = false;  // the last step's eigensolves kept their tight columns bit for bit
    int consistent = 0;   // DIIS entries measured on the current frozen orbitals' output
    double prev_tail = INFINITY;
    std::vector<lira_partial_level> partial;
    double q_out[2] = {0.0, 0.0};
    inner_abort_ = fd ? inner_abort : nullptr;
    abort_user_ = user;
    struct InnerReset {
        Solver *s;
        ~InnerReset() { s->inner_abort_ = nullptr; }
    } inner_reset{this};
    while (it < st->max_iter) {
        if (abort && abort(user)) {
            abandoned = true;
            break;
        }
        it++;
        // the step's eigen tolerance from the previous step's delta_rho (crates/qm-gpu/src/scf.rs:383-386, 1417-1419): a
        // delta_rho already below the tolerance it was tested against buys a tight step
        const double step_tol = delta_rho < test_tol ? lo
                                : std::isfinite(delta_rho) ? std::min(std::max(EIG_TOL_FRACTION * delta_rho, lo), hi)
                                                           : hi;
        const double delta_prev = delta_rho, test_prev = test_tol;
        const bool tight = step_tol <= lo;
        const int growths_before = doublings_;
        step_tail_ = 0.0;
        add2<T>(rho_t_.template as<T>(), rho_in_[0].template as<T>(), rho_in_[1].template as<T>(), N_, stream_);
        PoissonOut po = poisson(rho_t_.template as<T>(), vh_.as<double>(), ptol, 5000);
        t_poisson += po.seconds;
        max_pres = std::max(max_pres, po.residual);
        max_pit = std::max(max_pit, po.iterations);
        convert<T, double>(vh_t_.template as<T>(), vh_.as<double>(), N_, stream_);
        xc_lsda<T>(rho_in_[0].template as<T>(), rho_in_[1].template as<T>(), exc_.template as<T>(), vxc_[0].template as<T>(), vxc_[1].template as<T>(), N_, stream_, pcp);
        sum3<T>(veff_[0].template as<T>(), vnuc_.template as<T>(), vh_t_.template as<T>(), vxc_[0].template as<T>(), N_, stream_);
        if (!restricted) sum3<T>(veff_[1].template as<T>(), vnuc_.template as<T>(), vh_t_.template as<T>(), vxc_[1].template as<T>(), N_, stream_);
        auto te = Clock::now();
        // Fermi-Dirac: a warm guess within 1.5 tol is kept only when the previous step's delta_rho was below its stop;
        // on a loosened step the witness tail follows the previous delta_rho (crates/qm-gpu/src/scf.rs:1433-1437)
        const bool guess_band = delta_prev < test_prev;
        const double loosened = step_tol > lo ? delta_prev : NAN;
        bool step_frozen = solve_channel(0, veff_[0].template as<T>(), nel[0], fd, kt, step_tol, st->tol, guess_band,
                                         loosened, eo, occ[0]);
        if (!eo.aborted) {
            if (restricted) {
                Channel<T> &u = ch_[0], &d = ch_[1];
                ensure_channel(d, std::max(u.k, 1));
                if (u.k > 0)
                    LIRA_HIP(hipMemcpyAsync(d.X.ptr, u.X.ptr, (size_t)N_ * u.k * sizeof(T), hipMemcpyDeviceToDevice, stream_));
                d.k = u.k;
                d.evals = u.evals;
                d.res = u.res;
                d.occ = u.occ;
                d.tight = u.tight;
                occ[1] = occ[0];
            } else {
                step_frozen &= solve_channel(1, veff_[1].template as<T>(), nel[1], fd, kt, step_tol, st->tol, guess_band,
                                             loosened, eo, occ[1]);
            }
        }
        t_eigen += secs(te);
        if (eo.aborted) {
            // stopped inside an eigensolve: the iterate is parked with occupations 1 on each channel's lowest n
            // levels, and the step does not count (crates/qm-gpu/src/scf.rs:1445-1456)
            for (int s = 0; s < 2; s++) {
                Channel<T> &c = ch_[s];
                const std::vector<int> o = ascending(c.evals);
                c.occ.assign(c.evals.size(), 0.0);
                for (int i = 0; i < std::min((int)o.size(), nel[s]); i++) c.occ[o[i]] = 1.0;
            }
            it--;
            abandoned = true;
            aborted_in_step = true;
            break;
        }
        // output densities
        double *df = dsmall_.as<double>() + 1024;
        for (int s = 0; s < 2; s++) {
            Channel<T> &c = ch_[s];
            if (c.k > 0) {
                upload(df + 256 * s, c.occ.data(), c.k);
                density<T>(rho_out_[s].template as<T>(), c.X.template as<T>(), df + 256 * s, c.k, N_, stream_);
            } else {
                fill<T>(rho_out_[s].template as<T>(), N_, 0.0, stream_);
            }
        }
        // the residual of the entry this step would push goes to a free slot (one always is: H + 1 slots)
        int slot = 0;
        {
            std::vector<bool> used(H1, false);
            for (int o : order) used[o] = true;
            while (used[slot]) slot++;
        }
        for (int s = 0; s < 2; s++)
            sub2<T>(slots_[slot].res[s].template as<T>(), rho_out_[s].template as<T>(), rho_in_[s].template as<T>(), N_, stream_);
        // delta_rho, the electron counts, the new DIIS Gram row and (observed solves) the Harris-Foulkes double
        // counting in one readback (crates/qm-gpu/src/scf.rs:1488-1510)
        std::vector<Job> jobs;
        jobs.push_back({rho_out_[0].template as<T>(), rho_in_[0].template as<T>(), RED_ABSDIFF});
        jobs.push_back({rho_out_[1].template as<T>(), rho_in_[1].template as<T>(), RED_ABSDIFF});
        jobs.push_back({rho_out_[0].template as<T>(), rho_out_[0].template as<T>(), RED_SUM});
        jobs.push_back({rho_out_[1].template as<T>(), rho_out_[1].template as<T>(), RED_SUM});
        for (int o : order)
            for (int s = 0; s < 2; s++)
                jobs.push_back({slots_[slot].res[s].template as<T>(), slots_[o].res[s].template as<T>(), RED_DOT});
        for (int s = 0; s < 2; s++)
            jobs.push_back({slots_[slot].res[s].template as<T>(), slots_[slot].res[s].template as<T>(), RED_DOT});
        const size_t row_end = jobs.size();
        if (observe) {
            jobs.push_back({rho_t_.template as<T>(), vh_t_.template as<T>(), RED_DOT});
            jobs.push_back({rho_in_[0].template as<T>(), vxc_[0].template as<T>(), RED_DOT});
            jobs.push_back({rho_in_[1].template as<T>(), vxc_[1].template as<T>(), RED_DOT});
            jobs.push_back({exc_.template as<T>(), exc_.template as<T>(), RED_SUM});
        }
        const std::vector<double> vals = reduce_list(jobs, dv_);
        delta_rho = vals[0] + vals[1];
        q_out[0] = vals[2];
        q_out[1] = vals[3];
        std::vector<double> row;
        for (size_t i = 4; i < row_end; i += 2) row.push_back(vals[i] + vals[i + 1]);
        drho_steps_.push_back(delta_rho);
        etol_steps_.push_back(step_tol);
        // the tolerance this step is tested against: max(tol, partial_tol) when its occupations have a partly occupied
        // level (crates/qm/src/occupations.rs:427-446)
        partial.clear();
        if (fd)
            for (int s = 0; s < 2; s++) {
                const Channel<T> &c = ch_[s];
                const std::vector<int> o = ascending(c.evals);
                for (int r = 0; r < (int)o.size(); r++) {
                    const double f = c.occ[o[r]];
                    if (std::min(f, 1.0 - f) >= PARTIAL_OCCUPATION_MIN)
                        partial.push_back(lira_partial_level{s, r, c.evals[o[r]], f});
                }
            }
        test_tol = fd && !partial.empty() ? std::max(st->tol, partial_tol) : st->tol;
        if (trace_on())
            std::fprintf(stderr, "lira: SCF step %d delta_rho %.3e (test %.0e) eig tol %.1e%s, stored %d/%d, DIIS %d, %.1f ms\n",
                         it, delta_rho, test_tol, step_tol, tight ? " (tight)" : "", ch_[0].k, ch_[1].k,
                         (int)order.size(), 1e3 * secs(t0));
        // the stall watch (qm/src/scf.rs:455-466)
        if (delta_rho < best) best = delta_rho;
        if (std::isfinite(delta_rho) && (!std::isfinite(anchor) || delta_rho <= 0.5 * anchor)) {
            anchor = delta_rho;
            anchor_at = it;
        }
        const int since = it - anchor_at;
        if (observe) {
            // Harris-Foulkes energy of the step's input density: sum f eps + double counting - kT S
            std::vector<double> band;
            for (int s = 0; s < 2; s++)
                for (int i = 0; i < ch_[s].k; i++) band.push_back(ch_[s].evals[i] * ch_[s].occ[i]);
            const double dc = neumaier({-0.5 * vals[row_end], -vals[row_end + 1], -vals[row_end + 2], vals[row_end + 3], e_nn});
            const double ent = fd ? neumaier({occ[0].entropy, occ[1].entropy}) : 0.0;
            const lira_progress pr{it, st->max_iter, delta_rho, st->tol, neumaier(band) + dc - kt * ent, best, since,
                                   st->stall_window, level};
            observe(&pr, user);
        }
        if (!std::isfinite(delta_rho)) break;
        if (delta_rho < test_tol && tight) {
            converged = true;
            break;
        }
        if (!(delta_rho < test_tol) && st->stall_window > 0 && since >= st->stall_window && best >= level) {
            stalled = true;
            break;
        }
        // a frozen orbital set re-solved changes the map rho_in -> rho_out, and so does a witness window that widened
        // and grew the block: the history restarts (crates/qm-gpu/src/scf.rs:1538-1552)
        // (the f32 solve's rules: its eigen tolerance sits at the f32 residual floor, where a converged head comes back
        // frozen and re-solved by turns; the f64 solve keeps its history through such steps, as the CPU solver does)
        const bool f32_rules = sizeof(T) == 4;
        const bool widened = step_tail_ > prev_tail && doublings_ > growths_before;
        prev_tail = step_tail_;
        const bool map_changed = f32_rules && ((frozen && !step_frozen) || widened) && !order.empty();
        if (map_changed) profile_.diis_resets++;
        // a frozen phase that does not contract the residual by (1 - alpha) per step in the Gram norm drops the
        // entries older than the freeze (crates/qm-gpu/src/scf.rs:1553-1566)
        consistent = step_frozen && fd ? (frozen ? consistent + 1 : 2) : 0;
        const bool not_contracting =
            !order.empty() && row.back() > (1.0 - st->mixing) * (1.0 - st->mixing) * gram[(size_t)order.back() * H1 + order.back()];
        const bool freeze_trim = f32_rules && fd && frozen && step_frozen && !map_changed && not_contracting &&
                                 (int)order.size() + 1 > consistent;
        if (freeze_trim) profile_.diis_freeze_trims++;
        if (step_frozen && fd) profile_.frozen_steps++;
        frozen = step_frozen;
        // push (rho_in, r): the Gram row against every entry, then the restart, the trim or the oldest dropped
        for (size_t i = 0; i < order.size(); i++)
            gram[(size_t)slot * H1 + order[i]] = gram[(size_t)order[i] * H1 + slot] = row[i];
        gram[(size_t)slot * H1 + slot] = row.back();
        if (map_changed) order.clear();
        else if (freeze_trim)
            while ((int)order.size() > consistent - 1) order.pop_front();
        for (int s = 0; s < 2; s++) std::swap(slots_[slot].rho[s], rho_in_[s]);  // the entry keeps this rho_in
        order.push_back(slot);
        while ((int)order.size() > H) order.pop_front();
        // Pulay coefficients: the bordered system on the Gram scaled by its largest diagonal, +1e-12 on the diagonal,
        // LU; the oldest entries dropped while the solve fails (crates/qm-gpu/src/scf.rs:606-646)
        std::vector<double> cvec;
        for (;;) {
            const int m = (int)order.size();
            if (m == 1) {
                cvec.assign(1, 1.0);
                break;
            }
            const int mm = m + 1;
            std::vector<double> Bm((size_t)mm * mm, 0.0), rhs(mm, 0.0);
            double scale = 0.0;
            for (int i = 0; i < m; i++) scale = std::max(scale, gram[(size_t)order[i] * H1 + order[i]]);
            bool ok = false;
            if (scale > 0.0) {
                for (int i = 0; i < m; i++) {
                    for (int j = 0; j < m; j++) Bm[(size_t)i * mm + j] = gram[(size_t)order[i] * H1 + order[j]] / scale;
                    Bm[(size_t)i * mm + i] += DIIS_REG;
                    Bm[(size_t)i * mm + m] = Bm[(size_t)m * mm + i] = 1.0;
                }
                rhs[m] = 1.0;
                if (lu_solve(mm, Bm, rhs)) {
                    ok = true;
                    for (int i = 0; i < m; i++)
                        if (!std::isfinite(rhs[i])) ok = false;
                }
            }
            if (ok) {
                cvec.assign(rhs.begin(), rhs.begin() + m);
                break;
            }
            order.pop_front();
        }
        double *dcv = dsmall_.as<double>() + 2048;
        upload(dcv, cvec.data(), cvec.size());
        for (int s = 0; s < 2; s++) {
            std::vector<const T *> rp, rr;
            for (int o : order) {
                rp.push_back(slots_[o].rho[s].template as<T>());
                rr.push_back(slots_[o].res[s].template as<T>());
            }
            diis_mix<T>(rho_in_[s].template as<T>(), rp.data(), rr.data(), dcv, (int)order.size(), st->mixing, N_, stream_);
        }
        // each channel's charge renormalised after the clamp to its output's integral, on the device (no readback;
        // crates/qm-gpu/src/scf.rs:1574-1582)
        {
            RedJobs<T> jq{};
            for (int s = 0; s < 2; s++) {
                jq.a[s] = jq.b[s] = rho_in_[s].template as<T>();
                jq.kind[s] = RED_SUM;
            }
            jq.count = 2;
            reduce_jobs<T>(jq, n_, dv_, partials_.as<double>(), dres_.as<double>() + 1000, stream_);
            for (int s = 0; s < 2; s++)
                if (nel[s] > 0)
                    scale_ratio<T>(rho_in_[s].template as<T>(), N_, dres_.as<double>() + 2 + s, dres_.as<double>() + 1000 + s,
                                   stream_);
        }
    }
    out->iterations = it;
    out->delta_rho = delta_rho;
    out->max_eigen_residual = eo.max_residual;
    out->max_guard_residual = fd ? eo.guard_residual : 0.0;
    out->lobpcg_iterations = lobpcg_total_;
    out->restricted = restricted ? 1 : 0;
    out->tol_used = test_tol;
    record(it, eo);
    if (abandoned) {
        // the in-progress state is parked as the next warm start: its input densities stand in for the output ones
        // (crates/qm-gpu/src/scf.rs:1588-1606)
        for (int s = 0; s < 2; s++) copy<T>(rho_out_[s].template as<T>(), rho_in_[s].template as<T>(), N_, stream_);
        LIRA_HIP(hipStreamSynchronize(stream_));
        parked_ = true;
        state_ionic_version_ = ionic_version_;
        nel_[0] = nel[0];
        nel_[1] = nel[1];
        occ_scheme_ = st->occupations;
        kt_ = st->kt;
        profile_.abandoned_in_step = aborted_in_step ? it + 1 : 0;
        profile_.abandoned_before_step = aborted_in_step ? 0 : it + 1;
        out->seconds_total = secs(t0);
        return LIRA_ABORTED;
    }
    if (stalled) {
        out->seconds_total = secs(t0);
        return LIRA_STALLED;
    }
    if (!converged) {
        out->seconds_total = secs(t0);
        return LIRA_NOT_CONVERGED;
    }
    // energy from the output density and orbitals (qm/src/scf.rs:1048-1081)
    add2<T>(rho_t_.template as<T>(), rho_out_[0].template as<T>(), rho_out_[1].template as<T>(), N_, stream_);
    PoissonOut po = poisson(rho_t_.template as<T>(), vh_.as<double>(), ptol, 5000);
    t_poisson += po.seconds;
    max_pres = std::max(max_pres, po.residual);
    max_pit = std::max(max_pit, po.iterations);
    xc_lsda<T>(rho_out_[0].template as<T>(), rho_out_[1].template as<T>(), exc_.template as<T>(), vxc_[0].template as<T>(), vxc_[1].template as<T>(), N_, stream_, pcp);
    double e_nl = 0.0;
    if (nl_rows_ > 0)  // E_NL per spin channel, f-weighted under Fermi-Dirac (qm/src/scf.rs:1066-1072)
        e_nl = neumaier(nl_energy_terms(ch_[0], fd)) + neumaier(nl_energy_terms(ch_[1], fd));
    double kinetic = 0.0;
    {
        std::vector<double> terms;
        StencilParams sp{n_, 0, h_, -0.5, 0.0, 0.0};
        for (int s = 0; s < 2; s++) {
            Channel<T> &c = ch_[s];
            if (c.k == 0) continue;
            sp.nb = c.k;
            stencil<T>(ST_LAP, sp, c.X.template as<T>(), c.TMP.template as<T>(), nullptr, nullptr, nullptr, stream_);
            for (int b0 = 0; b0 < c.k; b0 += MAX_JOBS) {
                RedJobs<T> jk{};
                for (int b = b0; b < std::min(c.k, b0 + MAX_JOBS); b++) {
                    jk.a[b - b0] = c.X.template as<T>() + (long)b * N_;
                    jk.b[b - b0] = c.TMP.template as<T>() + (long)b * N_;
                    jk.kind[b - b0] = RED_DOT;
                    jk.count = b - b0 + 1;
                }
                auto tv = reduce(jk, dv_);
                for (int b = b0; b < std::min(c.k, b0 + MAX_JOBS); b++) {
                    const double f = fd ? c.occ[b] : 1.0;
                    if (fd && f == 0.0) continue;
                    terms.push_back(fd ? f * tv[b - b0] : tv[b - b0]);
                }
            }
        }
        kinetic = neumaier(terms);
    }
    convert<double, T>(rho64_.as<double>(), rho_t_.template as<T>(), N_, stream_);
    RedJobs<double> je{};
    je.a[0] = rho64_.as<double>();
    je.b[0] = vh_.as<double>();
    je.kind[0] = RED_DOT;
    je.count = 1;
    reduce_jobs<double>(je, n_, dv_, partials_.as<double>(), dres_.as<double>() + 32, stream_);
    RedJobs<T> jt{};
    jt.a[0] = rho_t_.template as<T>();
    jt.b[0] = vnuc_.template as<T>();
    jt.kind[0] = RED_DOT;
    jt.a[1] = exc_.template as<T>();
    jt.kind[1] = RED_SUM;
    jt.a[2] = rho_out_[0].template as<T>();
    jt.kind[2] = RED_SUM;
    jt.a[3] = rho_out_[1].template as<T>();
    jt.kind[3] = RED_SUM;
    jt.a[4] = rho_t_.template as<T>();
    jt.kind[4] = RED_FACEMAX;
    jt.count = 5;
    for (int i = 1; i < 5; i++) jt.b[i] = jt.a[i];
    reduce_jobs<T>(jt, n_, dv_, partials_.as<double>(), dres_.as<double>(), stream_);
    const double *ev = readback(dres_.as<double>(), 33);
    const double external_ = ev[0] + e_nl, xc_e = ev[1], hartree = 0.5 * ev[32];
    out->electrons[0] = ev[2];
    out->electrons[1] = ev[3];
    out->boundary_density = ev[4];
    const double internal = kinetic + external_ + hartree + xc_e + e_nn;
    const double entropy = neumaier({occ[0].entropy, occ[1].entropy});
    const double ts = kt * entropy;
    out->energy = lira_energy{kinetic, external_, hartree, xc_e, e_nn, internal, ts, internal - ts};
    out->entropy = entropy;
    out->mu[0] = occ[0].mu;
    out->mu[1] = occ[1].mu;
    out->poisson_residual = max_pres;
    out->poisson_iterations = max_pit;
    out->n_stored[0] = ch_[0].k;
    out->n_stored[1] = ch_[1].k;
    out->cg_iterations = cg_total_;
    out->readbacks = readbacks_;
    // the partly occupied levels, the core overlaps and the warnings (qm/src/scf.rs:1083-1090)
    out->n_partial = std::min((int)partial.size(), LIRA_MAX_PARTIAL);
    for (int i = 0; i < out->n_partial; i++) out->partial[i] = partial[i];
    out->n_core_overlaps = std::min((int)overlaps.size(), LIRA_MAX_OVERLAPS);
    for (int i = 0; i < out->n_core_overlaps; i++) out->core_overlaps[i] = overlaps[i];
    {
        std::string w;
        int nw = 0;
        char line[160];
        if (out->boundary_density > BOUNDARY_DENSITY_WARN) {
            std::snprintf(line, sizeof line, "density on box faces is %.2e: box too small\n", out->boundary_density);
            w += line;
            nw++;
        }
        for (const auto &ov : overlaps) {
            std::snprintf(line, sizeof line, "cores overlap: %d-%d at R = %.3f Bohr\n", ov.i, ov.j, ov.r);
            w += line;
            nw++;
        }
        if (w.size() >= sizeof out->warnings) w.resize(sizeof out->warnings - 1);
        std::memcpy(out->warnings, w.c_str(), w.size() + 1);
        out->n_warnings = nw;
    }
    out->seconds_poisson = t_poisson;
    out->seconds_eigen = t_eigen;
    out->seconds_total = secs(t0);
    out->seconds_other = out->seconds_total - t_poisson - t_eigen;
    valid_ = true;
    state_ionic_version_ = ionic_version_;
    last_tol_ = st->tol;
    last_eig_tol_ = st->eig_tol;
    if (extrapolate_ && reuse && prev_converged) {
        for (int s = 0; s < 2; s++) std::swap(prev_rho_[s], save_rho_[s]);
        have_prev_ = true;
    }
    record(it, eo);
    nel_[0] = nel[0];
    nel_[1] = nel[1];
    occ_scheme_ = st->occupations;
    kt_ = st->kt;
    return LIRA_OK;
}

// The five pseudised force terms (qm/src/forces.rs:259-350): [ion-ion, local LR, local SR, nonlocal, partial core]
template <typename T> void Solver<T>::force_terms(const lira_nucleus *nuc, int nat, double sigma, double *out) {
    if (!valid_) throw InvalidError("lira_force_terms: no converged state");
    if (species_.empty() || state_ionic_version_ != ionic_version_)
        throw InvalidError("lira_force_terms: the state was not converged under the current ionic setup");
    if (std::fabs(sigma - sigma_) > 1e-15) throw InvalidError("lira_force_terms: sigma differs from the solve's");
    if (nat < 0 || (nat > 0 && !nuc)) throw InvalidError("lira_force_terms: bad nuclei");
    upload_nuclei(nuc, nat);
    ionic_geometry(sigma);
    std::vector<double> t((size_t)15 * nat, 0.0);
    // ion-ion with valence charges (forces.rs:335-350)
    for (int i = 0; i < nat; i++)
        for (int j = 0; j < nat; j++) {
            if (i == j) continue;
            double d[3], d2 = 0.0;
            for (int a = 0; a < 3; a++) {
                d[a] = nuc[i].pos[a] - nuc[j].pos[a];
                d2 += d[a] * d[a];
            }
            const double dd = std::sqrt(d2);
            if (dd < 1e-8) throw std::pair<int, int>(std::min(i, j), std::max(i, j));
            const double sc = q_h_[i] * q_h_[j] / (dd * dd * dd);
            for (int a = 0; a < 3; a++) t[3 * i + a] += sc * d[a];
        }
    // local LR, local SR and partial core: radial sums on the device
    add2<T>(rho_t_.template as<T>(), rho_out_[0].template as<T>(), rho_out_[1].template as<T>(), N_, stream_);
    if (has_pc_)  // v_xc on rho_sigma + rho_pc / 2, the SCF's own XC (forces.rs:320-327)
        xc_lsda<T>(rho_out_[0].template as<T>(), rho_out_[1].template as<T>(), exc_.template as<T>(),
                   vxc_[0].template as<T>(), vxc_[1].template as<T>(), N_, stream_, pc_.template as<T>());
    const size_t np = (size_t)9 * std::max(nat, 1);
    if (fpart_.bytes < np * RED_BLOCKS * sizeof(double)) fpart_.alloc(np * RED_BLOCKS * sizeof(double));
    if (fout_.bytes < np * sizeof(double)) fout_.alloc(np * sizeof(double));
    ionic_forces<T>(rho_t_.template as<T>(), vxc_[0].template as<T>(), vxc_[1].template as<T>(), n_, h_,
                    centres_.as<IonicCentre>(), tab_.as<double>(), nat, fpart_.as<double>(), fout_.as<double>(),
                    stream_);
    if (nat > 0) {
        const double *r = readback(fout_.as<double>(), (size_t)9 * nat);
        for (int term = 0; term < 3; term++)
            for (int i = 0; i < 3 * nat; i++) t[(size_t)(term == 2 ? 4 : term + 1) * 3 * nat + i] = r[(size_t)term * 3 * nat + i];
    }
    // nonlocal: 2 sum_i f_i sum_lm E_KB c_lm <grad beta_lm | psi_i> (forces.rs:299-318)
    if (nl_rows_ > 0) {
        std::vector<double> cg[2];
        for (int s = 0; s < 2; s++) {
            Channel<T> &c = ch_[s];
            if (c.k == 0) continue;
            const size_t need = (size_t)nl_rows_ * 4 * c.k;
            if (need > nl_C_cap_) {
                nl_C_.alloc(need * sizeof(double));
                nl_C_cap_ = need;
            }
            nl_force<T>(c.X.template as<T>(), c.k, N_, n_, h_, nl_ions_.as<NlIon>(), (int)nl_h_.size(),
                        tab_.as<double>(), nl_C_.as<double>(), stream_);
            const double *r = readback(nl_C_.as<double>(), need);
            cg[s].assign(r, r + need);
        }
        const bool fdocc = occ_scheme_ == LIRA_FERMI_DIRAC;
        std::vector<std::vector<double>> per((size_t)3 * nat);
        for (size_t ii = 0; ii < nl_h_.size(); ii++) {
            const NlIon &ion = nl_h_[ii];
            const int atom = nl_atom_[ii];
            int row = ion.row_off;
            for (int j = 0; j < ion.nproj; j++) {  // one sphere per projector
                const int nm = 2 * ion.l[j] + 1;
                for (int s = 0; s < 2; s++) {
                    Channel<T> &c = ch_[s];
                    for (int b = 0; b < c.k; b++) {
                        const double f = fdocc ? c.occ[b] : 1.0;
                        if (f == 0.0) continue;
                        for (int m = 0; m < nm; m++) {
                            const long rr = row + m;
                            const double cv = cg[s][(rr * 4 + 0) * c.k + b];
                            for (int a = 0; a < 3; a++)
                                per[(size_t)atom * 3 + a].push_back(2.0 * f * ion.e_kb[j] * cv *
                                                                    cg[s][(rr * 4 + 1 + a) * c.k + b]);
                        }
                    }
                }
                row += nm;
            }
        }
        for (int i = 0; i < 3 * nat; i++) t[(size_t)9 * nat + i] = neumaier(per[i]);
    }
    std::memcpy(out, t.data(), t.size() * sizeof(double));
}

template <typename T> void Solver<T>::forces(const lira_nucleus *nuc, int nat, double sigma, double *out) {
    if (!valid_) throw InvalidError("lira_forces: no converged state");
    if (std::fabs(sigma - sigma_) > 1e-15) throw InvalidError("lira_forces: sigma differs from the solve's");
    if (state_ionic_version_ == ionic_version_ && !species_.empty()) {
        // pseudised: the per-component compensated sum of the five terms, in order (qm/src/forces.rs:249-257)
        std::vector<double> t((size_t)15 * nat);
        force_terms(nuc, nat, sigma, t.data());
        for (int i = 0; i < 3 * nat; i++) {
            std::vector<double> five(5);
            for (int term = 0; term < 5; term++) five[term] = t[(size_t)term * 3 * nat + i];
            out[i] = neumaier(five);
        }
        return;
    }
    upload_nuclei(nuc, nat);
    add2<T>(rho_t_.template as<T>(), rho_out_[0].template as<T>(), rho_out_[1].template as<T>(), N_, stream_);
    // 3 RED_BLOCKS partials and 3 sums per nucleus: sized by nat, like the ionic terms' buffers
    const size_t np = (size_t)3 * std::max(nat, 1);
    if (fpart_.bytes < np * RED_BLOCKS * sizeof(double)) fpart_.alloc(np * RED_BLOCKS * sizeof(double));
    if (fout_.bytes < np * sizeof(double)) fout_.alloc(np * sizeof(double));
    hf_forces<T>(rho_t_.template as<T>(), n_, h_, pos_.as<double>(), zf_.as<double>(), nat, sigma, fpart_.as<double>(),
                 fout_.as<double>(), stream_);
    const double *f = readback(fout_.as<double>(), (size_t)3 * nat);
    std::vector<double> fo(f, f + 3 * nat);
    // nucleus-nucleus part (qm/src/forces.rs:224-234)
    for (int i = 0; i < nat; i++)
        for (int j = 0; j < nat; j++) {
            if (i == j) continue;
            double d[3], d2 = 0.0;
            for (int a = 0; a < 3; a++) {
                d[a] = nuc[i].pos[a] - nuc[j].pos[a];
                d2 += d[a] * d[a];
            }
            const double dd = std::sqrt(d2);
            if (dd < 1e-8) throw std::pair<int, int>(std::min(i, j), std::max(i, j));
            const double s = (double)nuc[i].z * nuc[j].z / (dd * dd * dd);
            for (int a = 0; a < 3; a++) fo[3 * i + a] += s * d[a];
        }
    std::memcpy(out, fo.data(), fo.size() * sizeof(double));
}

template <typename T> void Solver<T>::get_field(int which, double *out) {
    if (!valid_ && which != 4 && which != 5) throw InvalidError("lira_get_field: no converged state");
    if (rho64_.bytes < N_ * sizeof(double)) rho64_.alloc(N_ * sizeof(double));
    double *d = rho64_.as<double>();
    switch (which) {
        case 0: convert<double, T>(d, rho_out_[0].template as<T>(), N_, stream_); break;
        case 1: convert<double, T>(d, rho_out_[1].template as<T>(), N_, stream_); break;
        case 2:
            add2<T>(rho_t_.template as<T>(), rho_out_[0].template as<T>(), rho_out_[1].template as<T>(), N_, stream_);
            convert<double, T>(d, rho_t_.template as<T>(), N_, stream_);
            break;
        case 3: copy<double>(d, vh_.as<double>(), N_, stream_); break;
        case 4:
            if (!vnuc_.ptr) throw InvalidError("lira_get_field: no potential yet");
            convert<double, T>(d, vnuc_.template as<T>(), N_, stream_);
            break;
        case 5:
            if (!has_pc_) fill<double>(d, N_, 0.0, stream_);
            else convert<double, T>(d, pc_.template as<T>(), N_, stream_);
            break;
        default: throw InvalidError("lira_get_field: unknown field");
    }
    LIRA_HIP(hipMemcpyAsync(out, d, N_ * sizeof(double), hipMemcpyDeviceToHost, stream_));
    LIRA_HIP(hipStreamSynchronize(stream_));
}

template <typename T> void Solver<T>::get_levels(int spin, double *evals, double *occ, int cap, int *count) {
    if (spin < 0 || spin > 1) throw InvalidError("lira_get_levels: spin must be 0 or 1");
    if (!valid_) throw InvalidError("lira_get_levels: no converged state");
    Channel<T> &c = ch_[spin];
    *count = c.k;
    for (int i = 0; i < std::min(cap, c.k); i++) {
        if (evals) evals[i] = c.evals[i];
        if (occ) occ[i] = c.occ[i];
    }
}

template <typename T> void Solver<T>::get_orbitals(int spin, double *out, int cap, int *count) {
    if (spin < 0 || spin > 1) throw InvalidError("lira_get_orbitals: spin must be 0 or 1");
    if (!valid_) throw InvalidError("lira_get_orbitals: no converged state");
    Channel<T> &c = ch_[spin];
    *count = c.k;
    if (rho64_.bytes < N_ * sizeof(double)) rho64_.alloc(N_ * sizeof(double));
    for (int b = 0; b < std::min(cap, c.k); b++) {
        convert<double, T>(rho64_.as<double>(), c.X.template as<T>() + (long)b * N_, N_, stream_);
        LIRA_HIP(hipMemcpyAsync(out + (long)b * N_, rho64_.ptr, N_ * sizeof(double), hipMemcpyDeviceToHost, stream_));
        LIRA_HIP(hipStreamSynchronize(stream_));
    }
}

template <typename T> PoissonOut Solver<T>::poisson_host(const double *rho, double *v, double tol, int max_iter) {
    alloc_field(rho_t_);
    if (vh_.bytes < N_ * sizeof(double)) vh_.alloc(N_ * sizeof(double));
    if (rho64_.bytes < N_ * sizeof(double)) rho64_.alloc(N_ * sizeof(double));
    LIRA_HIP(hipMemcpyAsync(rho64_.ptr, rho, N_ * sizeof(double), hipMemcpyHostToDevice, stream_));
    convert<T, double>(rho_t_.template as<T>(), rho64_.as<double>(), N_, stream_);
    LIRA_HIP(hipMemcpyAsync(vh_.ptr, v, N_ * sizeof(double), hipMemcpyHostToDevice, stream_));
    LIRA_HIP(hipStreamSynchronize(stream_));  // the reported solve time excludes the host transfers
    cg_total_ = 0;
    PoissonOut po = poisson(rho_t_.template as<T>(), vh_.as<double>(), tol, max_iter);
    LIRA_HIP(hipMemcpyAsync(v, vh_.ptr, N_ * sizeof(double), hipMemcpyDeviceToHost, stream_));
    LIRA_HIP(hipStreamSynchronize(stream_));
    valid_ = false;  // V_H is no longer the state's
    parked_ = false;
    return po;
}

// z = M r: one V-cycle of the Poisson preconditioner on host f64 arrays (Multigrid::apply_in,
// crates/qm-gpu/src/multigrid.rs:107-197), through the solve's own levels and kernels in precision T
template <typename T> void Solver<T>::mg_vcycle_host(const double *r, double *z) {
    if (!r || !z) throw InvalidError("lira_mg_vcycle: null array");
    poisson_setup();
    if (rho64_.bytes < N_ * sizeof(double)) rho64_.alloc(N_ * sizeof(double));
    LIRA_HIP(hipMemcpyAsync(rho64_.ptr, r, N_ * sizeof(double), hipMemcpyHostToDevice, stream_));
    convert<T, double>(pq_.template as<T>(), rho64_.as<double>(), N_, stream_);
    vcycle(0, pe_.template as<T>(), pq_.template as<T>());
    convert<double, T>(rho64_.as<double>(), pe_.template as<T>(), N_, stream_);
    LIRA_HIP(hipMemcpyAsync(z, rho64_.ptr, N_ * sizeof(double), hipMemcpyDeviceToHost, stream_));
    LIRA_HIP(hipStreamSynchronize(stream_));
}

template <typename T>
EigenOut Solver<T>::lobpcg_host(const double *v, int k, double *x, double *evals, double lam, double tol,
                                int max_iter) {
    if (k <= 0) throw InvalidError("lira_lobpcg: k must be positive");
    alloc_field(veff_[0]);
    if (rho64_.bytes < N_ * sizeof(double)) rho64_.alloc(N_ * sizeof(double));
    LIRA_HIP(hipMemcpyAsync(rho64_.ptr, v, N_ * sizeof(double), hipMemcpyHostToDevice, stream_));
    convert<T, double>(veff_[0].template as<T>(), rho64_.as<double>(), N_, stream_);
    Channel<T> &c = ch_[0];
    c.k = 0;
    ensure_channel(c, k);
    for (int b = 0; b < k; b++) {
        LIRA_HIP(hipMemcpyAsync(rho64_.ptr, x + (long)b * N_, N_ * sizeof(double), hipMemcpyHostToDevice, stream_));
        convert<T, double>(c.X.template as<T>() + (long)b * N_, rho64_.as<double>(), N_, stream_);
    }
    c.k = k;
    valid_ = false;  // the block is no longer the state's
    parked_ = false;
    EigenOut eo = lobpcg(c, veff_[0].template as<T>(), k, lam, tol, max_iter, tol, k, k, tol, false);
    for (int b = 0; b < k; b++) {
        convert<double, T>(rho64_.as<double>(), c.X.template as<T>() + (long)b * N_, N_, stream_);
        LIRA_HIP(hipMemcpyAsync(x + (long)b * N_, rho64_.ptr, N_ * sizeof(double), hipMemcpyDeviceToHost, stream_));
        LIRA_HIP(hipStreamSynchronize(stream_));
    }
    for (int i = 0; i < k; i++) evals[i] = c.evals[i];
    return eo;
}

// ------------------------------------------------------------------------------- display products ----
// One spin channel's display eigensolve on a given potential: the block's columns ascending after the solve.
template <typename T> void Solver<T>::sort_block(Channel<T> &c) {
    const std::vector<int> o = ascending(c.evals);
    bool sorted = true;
    for (int i = 0; i < (int)o.size(); i++)
        if (o[i] != i) sorted = false;
    if (!sorted) permute_columns(c, o);
}

// The unoccupied levels of the converged state (OrbitalView::unoccupied, crates/qm-gpu/src/orbitals.rs:997-1148): in
// the state's own output potentials V_nuc + V_H[rho_out] + v_xc[rho_out] (V_NL inside H under an ionic setup), one
// LOBPCG of the shown occupied columns plus m + UNOCC_GUARD more at the SCF's effective eigen tolerance, guessed from
// copies of the occupied orbitals, the previous call's unoccupied vectors, the witness columns and fresh LCG ones (seed
// 3 up / 4 down). The occupied pairs are checked against the SCF's (within tau) and dropped; the others are projected
// onto the complement of the SCF's occupied orbitals, re-measured (Rayleigh quotient and residual) and sorted, and kept
// as the next call's guess.
template <typename T>
void Solver<T>::unoccupied(int m, lira_unoccupied_result *out, double *orbitals) {
    if (!valid_) throw InvalidError("lira_unoccupied: no converged state");
    if (m < 0 || m + UNOCC_GUARD > LIRA_UNOCC_MAX) throw InvalidError("lira_unoccupied: m out of range");
    std::memset(out, 0, sizeof *out);
    const int ku = m + UNOCC_GUARD;
    const double tol = effective_eig_tol(last_eig_tol_);
    const double tau = last_tol_ <= TAU_MIN / TAU_PER_TOL ? TAU_MIN : std::max(TAU_PER_TOL * last_tol_, TAU_MIN);
    const bool fd = occ_scheme_ == LIRA_FERMI_DIRAC;
    const T *pcp = has_pc_ ? pc_.template as<T>() : nullptr;
    // the output potentials of the state (the SCF's kernels and term order)
    xc_lsda<T>(rho_out_[0].template as<T>(), rho_out_[1].template as<T>(), exc_.template as<T>(),
               vxc_[0].template as<T>(), vxc_[1].template as<T>(), N_, stream_, pcp);
    convert<T, double>(vh_t_.template as<T>(), vh_.as<double>(), N_, stream_);
    for (int s = 0; s < 2; s++)
        sum3<T>(veff_[s].template as<T>(), vnuc_.template as<T>(), vh_t_.template as<T>(), vxc_[s].template as<T>(), N_,
                stream_);
    const bool mirror = nel_[0] == nel_[1];
    const size_t fb = (size_t)N_ * sizeof(T);
    std::string failure;
    for (int s = 0; s < 2; s++) {
        if (mirror && s == 1) {  // a singlet's down channel is the up channel
            out->count[1] = out->count[0];
            out->iterations[1] = out->iterations[0];
            out->occupied_error[1] = out->occupied_error[0];
            for (int i = 0; i < out->count[0]; i++) {
                out->evals[1][i] = out->evals[0][i];
                out->residuals[1][i] = out->residuals[0][i];
            }
            if (orbitals)
                std::memcpy(orbitals + (size_t)ku * N_, orbitals, (size_t)ku * N_ * sizeof(double));
            continue;
        }
        Channel<T> &c = ch_[s];
        // the shown columns: every stored one, or under Fermi-Dirac the f-bearing ones (qm/src/orbitals.rs:1518-1526)
        std::vector<int> shown, witness;
        for (int i = 0; i < c.k; i++)
            ((!fd || c.occ[i] >= PARTIAL_OCCUPATION_MIN) ? shown : witness).push_back(i);
        const int nc = (int)shown.size();
        std::vector<double> ev_scf;
        for (int i : shown) ev_scf.push_back(c.evals[i]);
        std::sort(ev_scf.begin(), ev_scf.end());
        Channel<T> &d = disp_;
        d.k = 0;
        ensure_channel(d, nc + ku);
        int col = 0;
        auto copy_col = [&](const T *src) {
            LIRA_HIP(hipMemcpyAsync(d.X.template as<T>() + (long)col * N_, src, fb, hipMemcpyDeviceToDevice, stream_));
            col++;
        };
        for (int i : shown) copy_col(c.X.template as<T>() + (long)i * N_);
        const int prev = std::min(unocc_k_[s], ku);
        for (int i = 0; i < prev; i++) copy_col(unocc_[s].template as<T>() + (long)i * N_);
        for (int i = 0; i < (int)witness.size() && col < nc + ku; i++) copy_col(c.X.template as<T>() + (long)witness[i] * N_);
        if (col < nc + ku)
            initial_orbitals<T>(d.X.template as<T>() + (long)col * N_, n_, h_, pos_.as<double>(), nat_, nc + ku - col,
                                3ULL + (uint64_t)s, stream_);
        d.k = nc + ku;
        d.evals.assign(d.k, 0.0);
        d.res.assign(d.k, 0.0);
        const T *v = veff_[s].template as<T>();
        EigenOut e = lobpcg(d, v, d.k, NAN, tol, 200, tol, d.k, d.k, tol, false);
        out->iterations[s] = e.iterations;
        sort_block(d);
        double err = 0.0;
        for (int i = 0; i < nc; i++) err = std::max(err, std::fabs(d.evals[i] - ev_scf[i]));
        out->occupied_error[s] = err;
        if (!(err <= tau) && failure.empty()) {
            char msg[200];
            std::snprintf(msg, sizeof msg,
                          "the display eigensolve's occupied eigenvalues disagree with the SCF's by %.2e Ha (spin %d) > "
                          "tau = %.0e",
                          err, s, tau);
            failure = msg;
        }
        // the unoccupied vectors, projected onto the complement of the SCF's occupied orbitals
        T *phi = d.X.template as<T>() + (long)nc * N_;
        if (nc > 0) {
            for (int i = 0; i < nc; i++)
                LIRA_HIP(hipMemcpyAsync(d.TMP.template as<T>() + (long)i * N_, c.X.template as<T>() + (long)shown[i] * N_,
                                        fb, hipMemcpyDeviceToDevice, stream_));
            ensure_gram(ku, nc);
            Cols S{{{phi, ku}}, 1}, U{{{d.TMP.ptr, nc}}, 1};
            gram<T>(S, U, N_, dv_, gpart_.as<double>(), gpart_cap_, dG_.as<double>(), stream_);
            const double *g = readback(dG_.as<double>(), (size_t)ku * nc);
            std::vector<double> C((size_t)(ku + nc) * ku, 0.0);
            for (int i = 0; i < ku; i++) {
                C[(size_t)i * ku + i] = 1.0;
                for (int j = 0; j < nc; j++) C[(size_t)(ku + j) * ku + i] = -g[(size_t)i * nc + j];
            }
            const int ldc = upload_coeffs(C, ku + nc, ku);
            Cols Sm{{{phi, ku}, {d.TMP.ptr, nc}}, 2};
            lincomb<T>(Sm, dC_.as<double>(), ldc, ku, d.XN.template as<T>(), N_, stream_);
            LIRA_HIP(hipMemcpyAsync(phi, d.XN.ptr, (size_t)ku * fb, hipMemcpyDeviceToDevice, stream_));
        }
        // re-measured: the Rayleigh quotient and residual of each projected vector
        hamiltonian(v, phi, d.HX.template as<T>(), ku);
        ensure_gram(ku, 2 * ku);
        {
            Cols S{{{phi, ku}}, 1}, U{{{phi, ku}, {d.HX.ptr, ku}}, 2};
            gram<T>(S, U, N_, dv_, gpart_.as<double>(), gpart_cap_, dG_.as<double>(), stream_);
        }
        const double *g = readback(dG_.as<double>(), (size_t)ku * 2 * ku);
        std::vector<double> theta(ku), xx(ku);
        for (int i = 0; i < ku; i++) {
            xx[i] = g[(size_t)i * 2 * ku + i];
            theta[i] = g[(size_t)i * 2 * ku + ku + i] / xx[i];
        }
        upload(dsmall_.as<double>() + 256, theta.data(), ku);
        residual<T>(d.R.template as<T>(), d.HX.template as<T>(), phi, dsmall_.as<double>() + 256, ku, N_, stream_);
        std::vector<Job> jobs;
        for (int i = 0; i < ku; i++)
            jobs.push_back({d.R.template as<T>() + (long)i * N_, d.R.template as<T>() + (long)i * N_, RED_DOT});
        const std::vector<double> rr = reduce_list(jobs, dv_);
        std::vector<int> o(ku);
        std::iota(o.begin(), o.end(), 0);
        std::stable_sort(o.begin(), o.end(), [&](int a, int b) { return theta[a] < theta[b]; });
        // kept in order as the next call's guess
        if ((size_t)unocc_[s].bytes < (size_t)ku * fb) unocc_[s].alloc((size_t)ku * fb);
        for (int i = 0; i < ku; i++)
            LIRA_HIP(hipMemcpyAsync(unocc_[s].template as<T>() + (long)i * N_, phi + (long)o[i] * N_, fb,
                                    hipMemcpyDeviceToDevice, stream_));
        unocc_k_[s] = ku;
        out->count[s] = ku;
        for (int i = 0; i < ku; i++) {
            out->evals[s][i] = theta[o[i]];
            out->residuals[s][i] = std::sqrt(std::max(rr[o[i]], 0.0) / xx[o[i]]);
        }
        if (orbitals)
            for (int i = 0; i < ku; i++) {
                convert<double, T>(rho64_.as<double>(), unocc_[s].template as<T>() + (long)i * N_, N_, stream_);
                copy_now(orbitals + ((size_t)s * ku + i) * N_, rho64_.ptr, (size_t)N_ * sizeof(double),
                         hipMemcpyDeviceToHost);
            }
    }
    LIRA_HIP(hipStreamSynchronize(stream_));
    if (!failure.empty()) {
        unocc_k_[0] = unocc_k_[1] = 0;
        throw DisplayRejected(failure);
    }
}

// The Harris display solve (crates/qm-gpu/src/harris.rs:477-): one non-self-consistent diagonalisation in the potential
// of a given input density, V_sigma = v_nuc + V_H[rho_in] + v_xc,sigma[rho_in (+ rho_pc / 2)], under the solver's ionic
// setup at the given geometry. Per channel a block of n_sigma + m + UNOCC_GUARD levels plus one untested buffer column:
// the lowest n_sigma + m held to the SCF's effective eigen tolerance, the guards to GUARD_TOL; the occupations applied
// once (Aufbau: the n lowest; Fermi-Dirac: fermi_dirac over the tested levels, the block grown while no tested level is
// a witness, as the SCF grows its block); rho_out,sigma = sum f psi^2. The solver's converged state is left as it was.
template <typename T>
void Solver<T>::harris(const lira_system *sys, const lira_settings *st, const double *rho_in_up, const double *rho_in_dn,
                       int m, lira_harris_result *out, double *rho_out_up, double *rho_out_dn) {
    auto t0 = Clock::now();
    const int nat = sys->n_nuclei;
    if (nat < 0 || (nat > 0 && !sys->nuclei) || sys->n_up < 0 || sys->n_dn < 0 || !rho_in_up || !rho_in_dn || m < 0)
        throw InvalidError("lira_harris: bad arguments");
    if (!(st->sigma > 0.0) || !(st->tol > 0.0)) throw InvalidError("lira_harris: bad settings");
    const int nel[2] = {sys->n_up, sys->n_dn};
    for (int s = 0; s < 2; s++)
        if (nel[s] + m + UNOCC_GUARD + witness_cap_ + 1 > LIRA_HARRIS_MAX_LEVELS)
            throw InvalidError("lira_harris: more levels than a result holds");
    std::memset(out, 0, sizeof *out);
    const bool fd = st->occupations == LIRA_FERMI_DIRAC;
    const double kt = st->kt;
    const double tol = effective_eig_tol(st->eig_tol);
    // the state's geometry is restored afterwards: Harris has its own
    const std::vector<double> keep_pos = pos_h_, keep_z = z_h_;
    const int keep_nat = nat_;
    const double keep_sigma = sigma_;
    auto restore = [&]() {
        std::vector<lira_nucleus> nk(keep_nat);
        for (int i = 0; i < keep_nat; i++) {
            for (int a = 0; a < 3; a++) nk[i].pos[a] = keep_pos[3 * i + a];
            nk[i].z = (int)keep_z[i];
        }
        upload_nuclei(nk.data(), keep_nat);
        if (!species_.empty() && (valid_ || parked_)) {
            ionic_geometry(keep_sigma);
            LIRA_HIP(hipMemcpyAsync(vnuc_.ptr, hvnuc_save_.ptr, (size_t)N_ * sizeof(T), hipMemcpyDeviceToDevice, stream_));
        }
    };
    for (DevBuf *b : {&hr_[0], &hr_[1], &hrho_, &hvht_, &hexc_, &hxc_[0], &hxc_[1], &hv_[0], &hv_[1], &hvnuc_})
        alloc_field(*b);
    if (hvh_.bytes < N_ * sizeof(double)) {
        hvh_.alloc(N_ * sizeof(double));
        fill<double>(hvh_.as<double>(), N_, 0.0, stream_);  // the first call's Poisson starts cold
    }
    if (rho64_.bytes < N_ * sizeof(double)) rho64_.alloc(N_ * sizeof(double));
    const double *rin[2] = {rho_in_up, rho_in_dn};
    for (int s = 0; s < 2; s++) {
        LIRA_HIP(hipMemcpyAsync(rho64_.ptr, rin[s], N_ * sizeof(double), hipMemcpyHostToDevice, stream_));
        convert<T, double>(hr_[s].template as<T>(), rho64_.as<double>(), N_, stream_);
    }
    const bool ionic = !species_.empty();
    if (ionic && (valid_ || parked_)) {
        alloc_field(hvnuc_save_);
        LIRA_HIP(hipMemcpyAsync(hvnuc_save_.ptr, vnuc_.ptr, (size_t)N_ * sizeof(T), hipMemcpyDeviceToDevice, stream_));
    }
    upload_nuclei(sys->nuclei, nat);
    try {
        if (ionic) {
            ionic_geometry(st->sigma);
            LIRA_HIP(hipMemcpyAsync(hvnuc_.ptr, vnuc_.ptr, (size_t)N_ * sizeof(T), hipMemcpyDeviceToDevice, stream_));
        } else {
            nuclear_potential<T>(hvnuc_.template as<T>(), n_, h_, pos_.as<double>(), zf_.as<double>(), nat, st->sigma,
                                 stream_);
            nl_rows_ = 0;
            has_pc_ = false;
        }
        const T *pcp = has_pc_ ? pc_.template as<T>() : nullptr;
        add2<T>(hrho_.template as<T>(), hr_[0].template as<T>(), hr_[1].template as<T>(), N_, stream_);
        const double ptol = std::min(1e-10, 1e-3 * st->tol);
        PoissonOut po = poisson(hrho_.template as<T>(), hvh_.as<double>(), ptol, 5000);
        (void)po;
        convert<T, double>(hvht_.template as<T>(), hvh_.as<double>(), N_, stream_);
        xc_lsda<T>(hr_[0].template as<T>(), hr_[1].template as<T>(), hexc_.template as<T>(), hxc_[0].template as<T>(),
                   hxc_[1].template as<T>(), N_, stream_, pcp);
        for (int s = 0; s < 2; s++)
            sum3<T>(hv_[s].template as<T>(), hvnuc_.template as<T>(), hvht_.template as<T>(), hxc_[s].template as<T>(), N_,
                    stream_);
        for (int s = 0; s < 2; s++) {
            Channel<T> &c = hch_[s];
            const int base = nel[s] + m + UNOCC_GUARD;
            // the guess: the previous call's vectors (trimmed or topped up with fresh LCG columns, seed 5 up / 6 down)
            const int keep = std::min(c.k, base + 1);
            ensure_channel(c, base + 1);
            if (keep < base + 1)
                initial_orbitals<T>(c.X.template as<T>() + (long)keep * N_, n_, h_, pos_.as<double>(), nat, base + 1 - keep,
                                    5ULL + (uint64_t)s + 2ULL * (uint64_t)keep, stream_);
            if (c.k == 0) out->cold = 1;
            c.k = base + 1;
            c.evals.resize(c.k, 0.0);
            c.res.resize(c.k, 0.0);
            for (;;) {
                const int checked = c.k - 1;  // the top column is the untested buffer
                const int head = std::min(nel[s] + m, checked);
                EigenOut e = lobpcg(c, hv_[s].template as<T>(), c.k, NAN, tol, 200, std::max(GUARD_TOL, tol), head,
                                    checked, tol, false);
                out->iterations[s] += e.iterations;
                sort_block(c);
                if (!fd || nel[s] == 0) break;
                std::vector<double> ev(c.evals.begin(), c.evals.begin() + checked);
                const Fd o = fermi_dirac(ev, nel[s], kt);
                // a short window (no witness among the tested levels, qm::witness_reached on theta): the block grows
                // as the SCF's does, up to the witness cap
                auto ordr = ascending(ev);
                const double top = ev[ordr.back()], eps_n = ev[ordr[nel[s] - 1]];
                const bool reached = top >= eps_n + WITNESS_GAP && top - o.mu >= WITNESS_TAIL_KT * kt;
                if (reached || c.k - nel[s] >= witness_cap_ + 1) {
                    out->window_short[s] = reached ? 0 : 1;
                    break;
                }
                const int extras = c.k - nel[s];
                const int target = std::min(std::max(2 * extras, extras + 1), witness_cap_ + 1);
                ensure_channel(c, nel[s] + target);
                initial_orbitals<T>(c.X.template as<T>() + (long)c.k * N_, n_, h_, pos_.as<double>(), nat,
                                    nel[s] + target - c.k, 5ULL + (uint64_t)s + 2ULL * (uint64_t)c.k, stream_);
                c.evals.resize(nel[s] + target, c.evals.back());
                c.res.resize(nel[s] + target, 0.0);
                c.k = nel[s] + target;
            }
            // the occupations, applied once (qm::one_shot_occupations, crates/qm/src/scf.rs:304-325), over the tested
            // levels; the buffer holds no charge
            const int checked = c.k - 1;
            std::vector<double> f(c.k, 0.0);
            if (nel[s] > 0) {
                if (fd) {
                    std::vector<double> ev(c.evals.begin(), c.evals.begin() + checked);
                    const Fd o = fermi_dirac(ev, nel[s], kt);
                    for (int i = 0; i < checked; i++) f[i] = o.f[i];
                } else {
                    for (int i = 0; i < nel[s]; i++) f[i] = 1.0;
                }
            }
            int nocc = 0;
            for (int i = 0; i < c.k; i++)
                if (f[i] >= (fd ? PARTIAL_OCCUPATION_MIN : 0.5)) nocc++;
            out->n_occ[s] = nocc;
            out->n_levels[s] = checked;
            for (int i = 0; i < checked; i++) {
                out->evals[s][i] = c.evals[i];
                out->occ[s][i] = f[i];
            }
            c.occ = f;
            // rho_out = sum f psi^2
            double *df = dsmall_.as<double>() + 1024 + 256 * s;
            upload(df, f.data(), c.k);
            density<T>(hr_[s].template as<T>(), c.X.template as<T>(), df, c.k, N_, stream_);
        }
        RedJobs<T> jq{};
        for (int s = 0; s < 2; s++) {
            jq.a[s] = jq.b[s] = hr_[s].template as<T>();
            jq.kind[s] = RED_SUM;
        }
        jq.count = 2;
        const std::vector<double> q = reduce(jq, dv_);
        out->electrons[0] = q[0];
        out->electrons[1] = q[1];
        double *ro[2] = {rho_out_up, rho_out_dn};
        for (int s = 0; s < 2; s++)
            if (ro[s]) {
                convert<double, T>(rho64_.as<double>(), hr_[s].template as<T>(), N_, stream_);
                copy_now(ro[s], rho64_.ptr, (size_t)N_ * sizeof(double), hipMemcpyDeviceToHost);
            }
    } catch (...) {
        for (auto &c : hch_) c.k = 0;  // the next call is cold
        restore();
        throw;
    }
    restore();
    LIRA_HIP(hipStreamSynchronize(stream_));
    out->seconds = secs(t0);
}

}  // namespace lira

// ========================================================================================== C ABI ====
using namespace lira;

struct lira_solver {
    std::unique_ptr<SolverIface> impl;
    std::string err;
    int n = 0;
    double h = 0.0;
};

static thread_local std::string g_create_error;

template <typename F> static lira_status guarded(lira_solver *s, F &&f) {
    try {
        f();
        if (s) s->err.clear();
        return LIRA_OK;
    } catch (const std::pair<int, int> &p) {
        if (s) s->err = "nuclei " + std::to_string(p.first) + " and " + std::to_string(p.second) + " coincide";
        return LIRA_COINCIDENT_NUCLEI;
    } catch (const NumericalError &e) {
        if (s) {
            s->err = std::string("numerical failure in the solver: ") + e.what();
            if (s->impl) s->impl->reset();
        }
        return LIRA_NUMERICAL;
    } catch (const DeviceError &e) {
        if (s) {
            s->err = std::string("device failure: ") + e.what();
            if (s->impl) s->impl->reset();
        }
        return LIRA_DEVICE;
    } catch (const InvalidError &e) {
        if (s) s->err = e.what();
        return LIRA_INVALID;
    } catch (const std::exception &e) {
        if (s) {
            s->err = e.what();
            if (s->impl) s->impl->reset();
        }
        return LIRA_DEVICE;
    } catch (...) {  // nothing crosses the C ABI
        if (s) {
            s->err = "unknown failure";
            if (s->impl) s->impl->reset();
        }
        return LIRA_DEVICE;
    }
}

static hipStream_t as_stream(void *s) { return static_cast<hipStream_t>(s); }

extern "C" {

const char *lira_version(void) { return "liblira 0.1 (gfx950)"; }

void lira_default_settings(double h, lira_settings *out) {
    out->sigma = h;
    out->tol = 1e-6;
    out->max_iter = 200;
    out->mixing = 0.3;
    out->eig_tol = 1e-6;
    out->diis_history = 6;
    out->occupations = LIRA_FERMI_DIRAC;
    out->kt = 5e-3;
    out->stall_window = 50;
}

lira_status lira_solver_create(int n, double h, int precision, lira_solver **out) {
    if (!out) return LIRA_INVALID;
    *out = nullptr;
    // the f64 Poisson fields bound the grid in both precisions
    if (n < 3 || !stencil_buffer_ok(n, 8) || !(h > 0.0) || (precision != LIRA_FP64 && precision != LIRA_MIXED)) {
        g_create_error = "lira_solver_create: bad grid or precision";
        return LIRA_INVALID;
    }
    std::unique_ptr<lira_solver> s;
    try {
        s = std::make_unique<lira_solver>();
    } catch (const std::exception &e) {
        g_create_error = std::string("lira_solver_create: ") + e.what();
        return LIRA_DEVICE;
    }
    s->n = n;
    s->h = h;
    lira_status st = guarded(s.get(), [&] {
        if (precision == LIRA_FP64) s->impl = std::make_unique<Solver<double>>(n, h);
        else s->impl = std::make_unique<Solver<float>>(n, h);
    });
    if (st != LIRA_OK) {
        g_create_error = s->err;
        return st;
    }
    *out = s.release();
    return LIRA_OK;
}

void lira_solver_destroy(lira_solver *s) { delete s; }

const char *lira_last_error(const lira_solver *s) { return s ? s->err.c_str() : g_create_error.c_str(); }

lira_status lira_solve(lira_solver *s, const lira_system *sys, const lira_settings *settings, int warm,
                       lira_abort_fn abort, void *user, lira_scf_stats *stats) {
    return lira_solve_polled(s, sys, settings, warm, abort, nullptr, nullptr, user, stats);
}

lira_status lira_solve_polled(lira_solver *s, const lira_system *sys, const lira_settings *settings, int warm,
                              lira_abort_fn abort, lira_abort_fn inner_abort, lira_observe_fn observe, void *user,
                              lira_scf_stats *stats) {
    if (!s || !sys || !settings || !stats) return LIRA_INVALID;
    lira_status result = LIRA_OK;
    lira_status g =
        guarded(s, [&] { result = s->impl->solve(sys, settings, warm, abort, inner_abort, observe, user, stats); });
    if (g != LIRA_OK) return g;
    if (result == LIRA_NOT_CONVERGED)
        s->err = "SCF did not converge after " + std::to_string(stats->iterations) + " iterations";
    else if (result == LIRA_STALLED)
        s->err = "SCF stalled after " + std::to_string(stats->iterations) + " iterations";
    else if (result == LIRA_ABORTED)
        s->err = "aborted";
    return result;
}

lira_status lira_forces(lira_solver *s, const lira_nucleus *nuclei, int n_nuclei, double sigma, double *out) {
    if (!s || !nuclei || !out || n_nuclei <= 0) return LIRA_INVALID;
    return guarded(s, [&] { s->impl->forces(nuclei, n_nuclei, sigma, out); });
}

double lira_effective_gpu_tol(double tol, int n_valence) { return effective_gpu_tol(tol, n_valence); }

unsigned long long lira_solve_footprint_bytes(int n, int precision, int n_up, int n_dn, int witness,
                                              int diis_history) {
    if (n < 3 || n_up < 0 || n_dn < 0 || witness < 0 || (precision != LIRA_FP64 && precision != LIRA_MIXED)) return 0;
    return solve_footprint(n, precision == LIRA_FP64 ? 8 : 4, n_up, n_dn, witness, diis_history);
}

int lira_witness_cap_for_budget(int n, int precision, int n_up, int n_dn, int diis_history,
                                unsigned long long budget) {
    const int esize = precision == LIRA_FP64 ? 8 : 4;
    return witness_cap_for_budget([&](int w) { return solve_footprint(n, esize, n_up, n_dn, w, diis_history); },
                                  budget, WITNESS_GROWTH_COUNTED, MAX_WITNESS_GROWTH);
}

// One attempt of the provider's evaluation (GpuKs::evaluate_once, crates/providers/src/ks.rs:1573-1686): the
// floor-aware tolerance, solve_polled with both polls and the observer, the R15 cold retry of a failed warm solve, then
// the forces.
static lira_status evaluate_once(lira_solver *s, const lira_system *sys, const lira_settings *settings, int warm,
                                 lira_abort_fn abort, lira_abort_fn inner_abort, lira_observe_fn observe, void *user,
                                 lira_scf_stats *stats, double *forces, int *flags, double *tol_used) {
    lira_settings set = *settings;
    int fl = 0;
    // the floor-aware stop: under an ionic setup the tolerance is raised to what the f32 density can reach
    if (s->impl->ionic()) {
        const double te = effective_gpu_tol(set.tol, (long)sys->n_up + sys->n_dn);
        if (te > set.tol) {
            set.tol = te;
            fl |= LIRA_EVAL_TOL_RAISED;
        }
    }
    // R15: a warm attempt that did not converge or stalled is retried once cold, density extrapolation off for that
    // solve and back as it was afterwards; the failed attempt's steps and readbacks count in the result
    const bool extrapolate = s->impl->density_extrapolation();
    lira_scf_stats failed{};
    bool retried = false;
    lira_status first = LIRA_OK;
    const lira_status r = retry_cold_on_stall(
        warm != 0, LIRA_NOT_CONVERGED, LIRA_STALLED,
        [&](bool w) {
            if (warm && !w) {
                failed = *stats;
                s->impl->set_density_extrapolation(false);
            }
            return lira_solve_polled(s, sys, &set, w ? 1 : 0, abort, inner_abort, observe, user, stats);
        },
        &retried, &first);
    if (retried) {
        s->impl->set_density_extrapolation(extrapolate);
        fl |= LIRA_EVAL_RETRIED_COLD | (first == LIRA_STALLED ? LIRA_EVAL_WARM_STALLED : 0);
        stats->iterations += failed.iterations;
        stats->readbacks += failed.readbacks;
    }
    *flags = fl;
    *tol_used = set.tol;
    if (r != LIRA_OK) return r;
    lira_profile prof{};
    const lira_status pg = guarded(s, [&] { s->impl->get_profile(&prof, nullptr, nullptr, 0, nullptr); });
    if (pg != LIRA_OK) return pg;
    if (prof.extrapolated) *flags |= LIRA_EVAL_EXTRAPOLATED;
    return lira_forces(s, sys->nuclei, sys->n_nuclei, set.sigma, forces);
}

lira_status lira_evaluate(lira_solver *s, const lira_system *sys, const lira_settings *settings, int warm,
                          lira_abort_fn abort, lira_abort_fn inner_abort, lira_observe_fn observe, void *user,
                          unsigned long long vram_budget, lira_scf_stats *stats, double *forces, int *flags,
                          double *tol_used) {
    if (!s || !sys || !settings || !stats || !forces || !flags || !tol_used || sys->n_nuclei <= 0) return LIRA_INVALID;
    // the witness cap a memory budget leaves this solve (GpuKs::apply_witness_cap, crates/providers/src/ks.rs:400-428)
    int cap = s->impl->witness_cap();
    const bool budgeted = vram_budget > 0;
    if (budgeted) {
        cap = lira_witness_cap_for_budget(s->n, s->impl->precision(), sys->n_up, sys->n_dn, settings->diis_history,
                                          vram_budget);
        s->impl->set_witness_cap(cap);
    }
    // the "GPU busy" policy: a device failure is retried after a pause (ks.rs:317-350, 1688-1706)
    int fl = 0, device_failed = 0;
    const lira_status r = retry_on_device_fault(
        LIRA_DEVICE, LIRA_ABORTED,
        [&] { return evaluate_once(s, sys, settings, warm, abort, inner_abort, observe, user, stats, forces, &fl,
                                   tol_used); },
        [&](int ms) {  // sleeps in 20 ms slices, false once the caller asks to abort
            const auto end = std::chrono::steady_clock::now() + std::chrono::milliseconds(ms);
            while (std::chrono::steady_clock::now() < end) {
                if (abort && abort(user)) return false;
                std::this_thread::sleep_for(std::chrono::milliseconds(20));
            }
            return !(abort && abort(user));
        },
        &device_failed);
    if (device_failed > 0) fl |= LIRA_EVAL_DEVICE_RETRIED;
    if (budgeted) fl |= LIRA_EVAL_BUDGET_CAP;
    *flags = fl;
    // a witness search stopped by the budget's cap names the way out (with_witness_hint, ks.rs:289-297)
    const std::string witness_msg = "numerical failure in the solver: no gap witness within";
    if (r == LIRA_NUMERICAL && budgeted && cap < MAX_WITNESS_GROWTH && s->err.compare(0, witness_msg.size(), witness_msg) == 0)
        s->err += " - the device memory budget sets this cap: raise the budget or use a coarser grid";
    return r;
}

lira_status lira_set_witness_cap(lira_solver *s, int cap) {
    if (!s) return LIRA_INVALID;
    return guarded(s, [&] { s->impl->set_witness_cap(cap); });
}

int lira_witness_cap(const lira_solver *s) { return s ? s->impl->witness_cap() : 0; }

lira_status lira_set_density_extrapolation(lira_solver *s, int on) {
    if (!s) return LIRA_INVALID;
    return guarded(s, [&] { s->impl->set_density_extrapolation(on != 0); });
}

lira_status lira_get_profile(lira_solver *s, lira_profile *out, double *delta_rho_steps, double *eig_tol_steps,
                             int cap, int *n_steps) {
    if (!s || !out) return LIRA_INVALID;
    return guarded(s, [&] { s->impl->get_profile(out, delta_rho_steps, eig_tol_steps, cap, n_steps); });
}

lira_status lira_sample_density(lira_solver *s, const double *points, int count, double *out) {
    if (!s) return LIRA_INVALID;
    return guarded(s, [&] { s->impl->sample(points, count, out); });
}

lira_status lira_state_summary(lira_solver *s, double *spin_abs_integral, double *density_max) {
    if (!s || !spin_abs_integral || !density_max) return LIRA_INVALID;
    return guarded(s, [&] { s->impl->state_summary(spin_abs_integral, density_max); });
}

lira_status lira_set_graphs(lira_solver *s, int on) {
    if (!s) return LIRA_INVALID;
    return guarded(s, [&] { s->impl->set_graphs(on != 0); });
}

lira_status lira_set_device_rr(lira_solver *s, int on) {
    if (!s) return LIRA_INVALID;
    return guarded(s, [&] { s->impl->set_device_rr(on != 0); });
}

lira_status lira_free_atom_levels(int z, double *eps, double *electrons) {
    if (z < 1 || z > 36 || !eps || !electrons) return LIRA_INVALID;
    std::vector<double> e;
    double el = 0.0;
    if (!free_atom_levels(z, e, el)) return LIRA_NUMERICAL;
    for (size_t i = 0; i < e.size() && i < 5; i++) eps[i] = e[i];
    *electrons = el;
    return LIRA_OK;
}

lira_status lira_set_ionic(lira_solver *s, const lira_species *species, int n_species) {
    if (!s) return LIRA_INVALID;
    return guarded(s, [&] { s->impl->set_ionic(species, n_species); });
}

lira_status lira_force_terms(lira_solver *s, const lira_nucleus *nuclei, int n_nuclei, double sigma, double *out) {
    if (!s || !out) return LIRA_INVALID;
    return guarded(s, [&] { s->impl->force_terms(nuclei, n_nuclei, sigma, out); });
}

lira_status lira_ionic_fields(lira_solver *s, const lira_nucleus *nuclei, int n_nuclei, double sigma, double *v_loc,
                              double *rho_pc) {
    if (!s) return LIRA_INVALID;
    return guarded(s, [&] { s->impl->ionic_fields_host(nuclei, n_nuclei, sigma, v_loc, rho_pc); });
}

lira_status lira_nonlocal_apply(lira_solver *s, const double *x, int nb, double *out, double *c, int c_cap,
                                int *n_rows) {
    if (!s) return LIRA_INVALID;
    return guarded(s, [&] { s->impl->nonlocal_apply(x, nb, out, c, c_cap, n_rows); });
}

lira_status lira_get_field(lira_solver *s, int which, double *out) {
    if (!s || !out) return LIRA_INVALID;
    return guarded(s, [&] { s->impl->get_field(which, out); });
}

lira_status lira_get_levels(lira_solver *s, int spin, double *evals, double *occ, int cap, int *count) {
    if (!s || !count) return LIRA_INVALID;
    return guarded(s, [&] { s->impl->get_levels(spin, evals, occ, cap, count); });
}

lira_status lira_get_orbitals(lira_solver *s, int spin, double *out, int cap, int *count) {
    if (!s || !out || !count) return LIRA_INVALID;
    return guarded(s, [&] { s->impl->get_orbitals(spin, out, cap, count); });
}

double lira_effective_eig_tol(const lira_solver *s, double requested) {
    return s ? s->impl->effective_eig_tol(requested) : requested;
}

void lira_reset(lira_solver *s) {
    if (s) s->impl->reset();
}

lira_status lira_inject_readback_fault(lira_solver *s, int after) {
    if (!s) return LIRA_INVALID;
    s->impl->inject_readback_fault(after);
    return LIRA_OK;
}

lira_status lira_poisson(lira_solver *s, const double *rho, double *v, double tol, int max_iter,
                         lira_poisson_stats *stats) {
    if (!s || !rho || !v || !(tol > 0.0) || max_iter < 1) return LIRA_INVALID;
    return guarded(s, [&] {
        PoissonOut po = s->impl->poisson_host(rho, v, tol, max_iter);
        if (stats) *stats = lira_poisson_stats{po.iterations, po.outer, po.residual, po.seconds};
    });
}

lira_status lira_unoccupied(lira_solver *s, int m, lira_unoccupied_result *out, double *orbitals) {
    if (!s || !out) return LIRA_INVALID;
    lira_status result = LIRA_OK;
    const lira_status g = guarded(s, [&] {
        try {
            s->impl->unoccupied(m, out, orbitals);
        } catch (const DisplayRejected &e) {  // the converged state stays
            s->err = e.what();
            result = LIRA_NUMERICAL;
        }
    });
    return g != LIRA_OK ? g : result;
}

lira_status lira_harris(lira_solver *s, const lira_system *sys, const lira_settings *settings, const double *rho_in_up,
                        const double *rho_in_dn, int m, lira_harris_result *out, double *rho_out_up,
                        double *rho_out_dn) {
    if (!s || !sys || !settings || !out || !rho_in_up || !rho_in_dn) return LIRA_INVALID;
    lira_status result = LIRA_OK;
    const lira_status g = guarded(s, [&] {
        try {
            s->impl->harris(sys, settings, rho_in_up, rho_in_dn, m, out, rho_out_up, rho_out_dn);
        } catch (const NumericalError &e) {  // a display solve's numerical failure leaves the SCF state alone
            s->err = std::string("the Harris display solve failed: ") + e.what();
            result = LIRA_NUMERICAL;
        }
    });
    return g != LIRA_OK ? g : result;
}

lira_status lira_mg_vcycle(lira_solver *s, const double *r, double *z) {
    if (!s || !r || !z) return LIRA_INVALID;
    return guarded(s, [&] { s->impl->mg_vcycle_host(r, z); });
}

lira_status lira_mg_level_op(int dtype, int op, int n, double h, double d_low, double d_high, const void *x,
                             const void *b, void *out, void *stream) {
    if ((dtype != 4 && dtype != 8) || n < 1 || !(h > 0.0) || !(d_low >= 1.0) || !(d_high >= 1.0) || !x || !out ||
        x == out)
        return LIRA_INVALID;
    if ((op == LIRA_MG_JACOBI || op == LIRA_MG_RESIDUAL) && (!b || b == out)) return LIRA_INVALID;
    if ((op == LIRA_MG_RESTRICT || op == LIRA_MG_PROLONG_ADD) && n < 2) return LIRA_INVALID;
    if (op < LIRA_MG_JACOBI || op > LIRA_MG_PROLONG_ADD) return LIRA_INVALID;
    return guarded(nullptr, [&] {
        const double h2 = h * h, gl = 1.0 - 1.0 / d_low, gh = 1.0 - 1.0 / d_high;
        hipStream_t st = as_stream(stream);
        auto run = [&](auto *tag) {
            using U = std::remove_pointer_t<decltype(tag)>;
            const U *xx = static_cast<const U *>(x), *bb = static_cast<const U *>(b);
            U *o = static_cast<U *>(out);
            if (op == LIRA_MG_JACOBI) mg_jacobi<U>(o, xx, bb, n, h2, 6.0 / 7.0, gl, gh, st);
            else if (op == LIRA_MG_RESIDUAL) mg_residual<U>(o, xx, bb, n, h2, gl, gh, st);
            else if (op == LIRA_MG_RESTRICT) mg_restrict<U>(o, xx, n / 2, n, st);
            else mg_prolong_add<U>(o, xx, n, n / 2, st);
        };
        if (dtype == 4) run((float *)nullptr);
        else run((double *)nullptr);
        LIRA_HIP(hipStreamSynchronize(st));
    });
}

lira_status lira_lobpcg(lira_solver *s, const double *v, int k, double *x, double *evals, double precond_lambda,
                        double tol, int max_iter, lira_eigen_stats *stats) {
    if (!s || !v || !x || !evals || k < 1 || !(tol > 0.0)) return LIRA_INVALID;
    return guarded(s, [&] {
        EigenOut eo = s->impl->lobpcg_host(v, k, x, evals, precond_lambda, tol, max_iter);
        if (stats) *stats = lira_eigen_stats{eo.iterations, eo.max_residual, eo.fallbacks};
    });
}

lira_status lira_rayleigh_ritz(int m, const double *a, const double *b, int nk, double *theta, double *c) {
    if (m < 1 || nk < 1 || nk > m || !a || !b || !theta || !c) return LIRA_INVALID;
    try {
        std::vector<double> A(a, a + (size_t)m * m), B(b, b + (size_t)m * m), th, C;
        rayleigh_ritz(m, A, B, nk, 1e-6, th, C);
        std::memcpy(theta, th.data(), nk * sizeof(double));
        std::memcpy(c, C.data(), (size_t)m * nk * sizeof(double));
        return LIRA_OK;
    } catch (const NumericalError &) {
        return LIRA_NUMERICAL;
    } catch (const std::bad_alloc &) {
        return LIRA_INVALID;  // m too large for the host
    } catch (...) {
        return LIRA_NUMERICAL;
    }
}

lira_status lira_rayleigh_ritz_device(int m, const double *a, const double *b, int nk, double *theta, double *c,
                                      double *c_p) {
    if (m < 1 || m > RR_DEV_MAX || nk < 1 || nk > m || !a || !b || !theta || !c) return LIRA_INVALID;
    lira_status status = LIRA_OK;
    const lira_status g = guarded(nullptr, [&] {
        // the Gram layout of a LOBPCG step with k = nk columns of X at theta = 0 (so A comes through unchanged), no
        // residual tested and a step that never stops
        const int q = 2 * m, mp = lincomb_rows(m), ldc = lincomb_ldc(2 * nk);
        std::vector<double> G((size_t)m * q + m, 0.0);
        for (int i = 0; i < m; i++)
            for (int j = 0; j < m; j++) {
                G[(size_t)i * q + j] = b[(size_t)i * m + j];
                G[(size_t)i * q + m + j] = a[(size_t)i * m + j];
            }
        DevBuf dg(G.size() * sizeof(double)), dth(RR_DEV_MAX * sizeof(double)),
            dc((size_t)mp * ldc * sizeof(double)), dw(rr_step_work_doubles() * sizeof(double)),
            drec(RR_REC_STRIDE * sizeof(double));
        LIRA_HIP(hipMemcpy(dg.ptr, G.data(), G.size() * sizeof(double), hipMemcpyHostToDevice));
        LIRA_HIP(hipMemset(dth.ptr, 0, RR_DEV_MAX * sizeof(double)));
        RrStepArgs ra{};
        ra.G = dg.as<double>();
        ra.theta = dth.as<double>();
        ra.C = dc.as<double>();
        ra.work = dw.as<double>();
        ra.record = drec.as<double>();
        ra.m = m, ra.q = q, ra.k = nk, ra.na = m - nk, ra.np = 0, ra.nm = nk, ra.locked = 0, ra.head = 0;
        ra.checked = 0, ra.iter = 1, ra.max_iter = 1 << 30, ra.is_f32 = 0, ra.mp = mp, ra.ldc = ldc;
        ra.tol = -1.0, ra.guess_tol = -1.0, ra.tail_scale = 1.0, ra.floor = 0.0, ra.alpha_max = 1e300;
        ra.alpha_fraction = 0.0, ra.drop = 1e-6;
        rr_step(ra, nullptr);
        std::vector<double> C((size_t)mp * ldc), rec(RR_REC_STRIDE);
        LIRA_HIP(hipMemcpy(C.data(), dc.ptr, C.size() * sizeof(double), hipMemcpyDeviceToHost));
        LIRA_HIP(hipMemcpy(rec.data(), drec.ptr, rec.size() * sizeof(double), hipMemcpyDeviceToHost));
        LIRA_HIP(hipMemcpy(theta, dth.ptr, nk * sizeof(double), hipMemcpyDeviceToHost));
        if (rec[1] != 0.0) {
            status = LIRA_NUMERICAL;
            return;
        }
        for (int i = 0; i < m; i++)
            for (int j = 0; j < nk; j++) {
                c[(size_t)i * nk + j] = C[(size_t)i * ldc + j];
                if (c_p) c_p[(size_t)i * nk + j] = C[(size_t)i * ldc + nk + j];
            }
    });
    return g != LIRA_OK ? g : status;
}

lira_status lira_hamiltonian(int dtype, int n, double h, int nbands, const void *v, const void *x, void *hx,
                             void *stream) {
    if ((dtype != 4 && dtype != 8) || n < 1 || nbands < 0 || !v || !x || !hx || x == hx) return LIRA_INVALID;
    return guarded(nullptr, [&] {
        StencilParams sp{n, nbands, h, 1.0, 0.0, 0.0};
        if (dtype == 4)
            stencil<float>(ST_HAM, sp, (const float *)x, (float *)hx, (const float *)v, nullptr, nullptr, as_stream(stream));
        else
            stencil<double>(ST_HAM, sp, (const double *)x, (double *)hx, (const double *)v, nullptr, nullptr,
                            as_stream(stream));
    });
}

lira_status lira_laplacian(int dtype, int n, double h, int nbands, double scale, const void *x, void *out,
                           void *stream) {
    if ((dtype != 4 && dtype != 8) || n < 1 || nbands < 0 || !x || !out || x == out) return LIRA_INVALID;
    return guarded(nullptr, [&] {
        StencilParams sp{n, nbands, h, scale, 0.0, 0.0};
        if (dtype == 4)
            stencil<float>(ST_LAP, sp, (const float *)x, (float *)out, nullptr, nullptr, nullptr, as_stream(stream));
        else
            stencil<double>(ST_LAP, sp, (const double *)x, (double *)out, nullptr, nullptr, nullptr, as_stream(stream));
    });
}

lira_status lira_jacobi_sweep(int dtype, int n, double h, int nbands, const void *v, const void *w, const void *r,
                              const double *lam, void *out, void *stream) {
    if ((dtype != 4 && dtype != 8) || n < 1 || nbands < 0 || !v || !w || !r || !lam || !out || w == out)
        return LIRA_INVALID;
    return guarded(nullptr, [&] {
        StencilParams sp{n, nbands, h, 1.0, jacobi_omega(), -1.5 * COEFF0 / (h * h)};
        if (dtype == 4)
            stencil<float>(ST_JAC, sp, (const float *)w, (float *)out, (const float *)v, (const float *)r, lam,
                           as_stream(stream));
        else
            stencil<double>(ST_JAC, sp, (const double *)w, (double *)out, (const double *)v, (const double *)r, lam,
                            as_stream(stream));
    });
}

lira_status lira_xc_lsda(int dtype, long count, const void *nu, const void *nd, void *e, void *vu, void *vd,
                         void *stream) {
    if ((dtype != 4 && dtype != 8) || count < 0 || !nu || !nd || !e || !vu || !vd) return LIRA_INVALID;
    return guarded(nullptr, [&] {
        if (dtype == 4)
            xc_lsda<float>((const float *)nu, (const float *)nd, (float *)e, (float *)vu, (float *)vd, count,
                           as_stream(stream));
        else
            xc_lsda<double>((const double *)nu, (const double *)nd, (double *)e, (double *)vu, (double *)vd, count,
                            as_stream(stream));
    });
}

lira_status lira_nuclear_potential(int dtype, int n, double h, const lira_nucleus *nuclei, int n_nuclei, double sigma,
                                   void *out, void *stream) {
    if ((dtype != 4 && dtype != 8) || n < 1 || n_nuclei < 0 || !out || !(sigma > 0.0)) return LIRA_INVALID;
    return guarded(nullptr, [&] {
        std::vector<double> ph(3 * n_nuclei + 1), zh(n_nuclei + 1);
        for (int i = 0; i < n_nuclei; i++) {
            for (int a = 0; a < 3; a++) ph[3 * i + a] = nuclei[i].pos[a];
            zh[i] = nuclei[i].z;
        }
        DevBuf dp(ph.size() * sizeof(double)), dz(zh.size() * sizeof(double));
        LIRA_HIP(hipMemcpy(dp.ptr, ph.data(), ph.size() * sizeof(double), hipMemcpyHostToDevice));
        LIRA_HIP(hipMemcpy(dz.ptr, zh.data(), zh.size() * sizeof(double), hipMemcpyHostToDevice));
        if (dtype == 4)
            nuclear_potential<float>((float *)out, n, h, dp.as<double>(), dz.as<double>(), n_nuclei, sigma,
                                     as_stream(stream));
        else
            nuclear_potential<double>((double *)out, n, h, dp.as<double>(), dz.as<double>(), n_nuclei, sigma,
                                      as_stream(stream));
        LIRA_HIP(hipStreamSynchronize(as_stream(stream)));
    });
}

lira_status lira_gram(int dtype, int n, double h, int m, const void *s_block, int q, const void *u_block, double *g,
                      void *stream) {
    if ((dtype != 4 && dtype != 8) || n < 1 || m < 1 || q < 1 || !s_block || !u_block || !g) return LIRA_INVALID;
    return guarded(nullptr, [&] {
        const long N = (long)n * n * n;
        const size_t need = gram_partials_needed(m, q, N);
        DevBuf part(need * sizeof(double)), dg((size_t)m * q * sizeof(double));
        Cols S{{{s_block, m}}, 1}, U{{{u_block, q}}, 1};
        if (dtype == 4) gram<float>(S, U, N, h * h * h, part.as<double>(), need, dg.as<double>(), as_stream(stream));
        else gram<double>(S, U, N, h * h * h, part.as<double>(), need, dg.as<double>(), as_stream(stream));
        LIRA_HIP(hipMemcpyAsync(g, dg.ptr, (size_t)m * q * sizeof(double), hipMemcpyDeviceToHost, as_stream(stream)));
        LIRA_HIP(hipStreamSynchronize(as_stream(stream)));
    });
}

int lira_stencil_variant_count(void) { return stencil_variants(); }
const char *lira_stencil_variant_name(int v) { return stencil_variant_name(v); }
void lira_set_stencil_variant(int dtype, int variant) { stencil_set_default_variant(dtype, variant); }

int lira_autotune_stencil(int dtype, int n, int nbands, double *ms_out) {
    if ((dtype != 4 && dtype != 8) || n < 1 || nbands < 1) return -1;
    int best = -1;
    double best_ms = 1e300;
    lira_status st = guarded(nullptr, [&] {
        const long N = (long)n * n * n;
        const size_t es = dtype;
        DevBuf x(N * nbands * es), o(N * nbands * es), v(N * es);
        LIRA_HIP(hipMemset(x.ptr, 0, N * nbands * es));
        LIRA_HIP(hipMemset(v.ptr, 0, N * es));
        hipEvent_t a, b;
        LIRA_HIP(hipEventCreate(&a));
        LIRA_HIP(hipEventCreate(&b));
        StencilParams sp{n, nbands, 0.2, 1.0, 0.0, 0.0};
        for (int var = 0; var < stencil_variants(); var++) {
            if (!stencil_variant_usable(var, n, dtype)) continue;
            auto run = [&] {
                if (dtype == 4) stencil<float>(ST_HAM, sp, x.as<float>(), o.as<float>(), v.as<float>(), nullptr, nullptr, nullptr, var);
                else stencil<double>(ST_HAM, sp, x.as<double>(), o.as<double>(), v.as<double>(), nullptr, nullptr, nullptr, var);
            };
            run();
            LIRA_HIP(hipEventRecord(a, nullptr));
            for (int r = 0; r < 3; r++) run();
            LIRA_HIP(hipEventRecord(b, nullptr));
            LIRA_HIP(hipEventSynchronize(b));
            float ms = 0.f;
            LIRA_HIP(hipEventElapsedTime(&ms, a, b));
            if (ms / 3.0 < best_ms) {
                best_ms = ms / 3.0;
                best = var;
            }
        }
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        stencil_set_default_variant(dtype, best);
    });
    if (st != LIRA_OK) return -1;
    if (ms_out) *ms_out = best_ms;
    return best;
}

}  // extern "C"
