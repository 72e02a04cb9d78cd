/*
 * lira.h — C ABI of liblira, the MI355X-native real-space Kohn–Sham SCF hot path.
 *
 * It replaces the reference's `qm_gpu::GpuSolver` Rust API (crates/qm-gpu/src/scf.rs:1044-1842,
 * crates/qm-gpu/src/forces.rs:239-297) with plain pointers and sizes, so a Rust `-sys` crate (or ctypes) binds it
 * without any GPU types. Units: Bohr and Hartree. Grids are cubic n^3 with spacing h, centred at the origin
 * (origin = -(n-1)h/2, crates/qm/src/grid.rs:11-14); host arrays use the qm layout idx = i + n(j + n k).
 *
 * Threading: a solver is used by one thread at a time (the reference's GpuSolver is Send, not Sync,
 * crates/qm-gpu/src/scf.rs:1042-1043). Errors: every call returns a lira_status; lira_last_error() gives the message
 * (the reference's QmError variants, crates/qm/src/error.rs:1-33). After any error except LIRA_COINCIDENT_NUCLEI
 * the solver drops its warm state (the next solve is cold), as crates/qm-gpu/src/scf.rs:1177-1192.
 */
#ifndef LIRA_H
#define LIRA_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    LIRA_OK = 0,
    LIRA_COINCIDENT_NUCLEI = 1, /* QmError::CoincidentNuclei */
    LIRA_NOT_CONVERGED = 2,     /* QmError::ScfNotConverged */
    LIRA_STALLED = 3,           /* QmError::ScfStalled */
    LIRA_NUMERICAL = 4,         /* QmError::Numerical */
    LIRA_DEVICE = 5,            /* QmError::Device (a failed HIP call or readback) */
    LIRA_INVALID = 6,           /* bad arguments (a programming error in the reference: a panic) */
    LIRA_ABORTED = 7            /* the abort callback returned nonzero (the reference's Ok(None)) */
} lira_status;

typedef enum {
    LIRA_FP64 = 0,  /* every grid quantity in f64 */
    LIRA_MIXED = 1  /* f32 grid fields (stencil, V-cycles, block products), f64 Rayleigh–Ritz, DIIS, energies */
} lira_precision;

typedef enum { LIRA_AUFBAU = 0, LIRA_FERMI_DIRAC = 1 } lira_occupations;

typedef struct lira_solver lira_solver;

/* qm::Nucleus (crates/qm/src/nuclear.rs:100-104) */
typedef struct {
    double pos[3];
    int z;
} lira_nucleus;

/* qm::System (crates/qm/src/scf.rs:329) without an ionic setup (all-electron) */
typedef struct {
    const lira_nucleus *nuclei;
    int n_nuclei;
    int n_up, n_dn;
} lira_system;

/* qm::ScfSettings (crates/qm/src/scf.rs:375-385); lira_default_settings fills ScfSettings::for_grid */
typedef struct {
    double sigma;     /* nuclear smearing width (Bohr); for_grid: h */
    double tol;       /* on delta_rho = int |rho_out - rho_in|; 1e-6 */
    int max_iter;     /* 200 */
    double mixing;    /* DIIS step alpha; 0.3 */
    double eig_tol;   /* LOBPCG residual tolerance; 1e-6 */
    int diis_history; /* 6 */
    int occupations;  /* lira_occupations; for_grid: Fermi–Dirac */
    double kt;        /* Fermi–Dirac kT (Ha); 5e-3 */
    int stall_window; /* 50; 0 = off */
} lira_settings;

/* qm::Energy (crates/qm/src/scf.rs:546-548) */
typedef struct {
    double kinetic, external_, hartree, xc, nuclear, internal, ts, total;
} lira_energy;

/* qm::PartialLevel (crates/qm/src/occupations.rs:425-437): a level with min(f, 1 - f) >= 1e-6 */
typedef struct {
    int spin;   /* 0 up, 1 down */
    int index;  /* rank in ascending order of eps within its channel */
    double eps, f;
} lira_partial_level;
/* one entry of ScfStats::core_overlaps (crates/qm/src/ionic.rs:175-185): nuclei i < j whose core radii overlap */
typedef struct {
    int i, j;
    double r;
} lira_core_overlap;
#define LIRA_MAX_PARTIAL 16
#define LIRA_MAX_OVERLAPS 16
#define LIRA_WARNINGS_BYTES 512
#define LIRA_MAX_DIIS_HISTORY 15

/* qm::ScfStats (crates/qm/src/scf.rs:558-574), plus the solve's profile (SolveProfile, crates/qm-gpu/src/scf.rs:420-514) */
typedef struct {
    int iterations;
    double delta_rho;
    lira_energy energy;
    double boundary_density;
    double max_eigen_residual;
    double poisson_residual;
    int poisson_iterations;
    double entropy;
    double mu[2];
    double electrons[2];
    double tol_used;  /* the delta_rho tolerance the converged step was tested against: settings.tol, or in LIRA_MIXED
                         under Fermi-Dirac max(settings.tol, 1e-5) on a step with a partly occupied level */
    int n_stored[2];  /* orbitals stored per spin */
    int restricted;   /* 1 when the spin-down channel mirrored the spin-up one (n_up == n_dn) */
    /* profile */
    int lobpcg_iterations;
    int cg_iterations;
    int readbacks;
    double seconds_total, seconds_poisson, seconds_eigen, seconds_other;
    /* Fermi-Dirac: the worst exit residual of the loose (guard) witness columns; 0 under Aufbau */
    double max_guard_residual;
    /* the final step's partly occupied levels, ascending eps per spin, up first (at most LIRA_MAX_PARTIAL) */
    int n_partial;
    lira_partial_level partial[LIRA_MAX_PARTIAL];
    /* pseudised nuclei whose cores overlap (at most LIRA_MAX_OVERLAPS); each is also a warning */
    int n_core_overlaps;
    lira_core_overlap core_overlaps[LIRA_MAX_OVERLAPS];
    /* ScfStats::warnings: n_warnings lines, each ending in '\n', NUL-terminated (cut at LIRA_WARNINGS_BYTES) */
    int n_warnings;
    char warnings[LIRA_WARNINGS_BYTES];
} lira_scf_stats;

/* qm::ScfProgress (crates/qm/src/scf.rs:482-498): one SCF step as an observer sees it. energy is the Harris-Foulkes
 * energy of the step's input density (F_HF = sum f eps + double counting - kT S under Fermi-Dirac). */
typedef struct {
    int iteration, max_iter;
    double delta_rho, tol, energy, best_delta_rho;
    int since_gain, window;
    double stall_level;
} lira_progress;
typedef void (*lira_observe_fn)(const lira_progress *progress, void *user);

/* Poisson solve statistics (PoissonStats, crates/qm/src/poisson.rs:136) */
typedef struct {
    int iterations;      /* inner (preconditioned CG) iterations */
    int outer;           /* f64 residual-correction cycles */
    double residual;     /* ||b - A V|| / ||b|| in f64 at exit */
    double seconds;
} lira_poisson_stats;

/* LOBPCG statistics (EigenStats, crates/qm/src/eigen.rs:132-139) */
typedef struct {
    int iterations;
    double max_residual;
    int fallbacks;
} lira_eigen_stats;

typedef int (*lira_abort_fn)(void *user);

/* ------------------------------------------------------------------------------------------- solver ---- */

/* Fills ScfSettings::for_grid (crates/qm/src/scf.rs:380-385) for spacing h. */
void lira_default_settings(double h, lira_settings *out);

/* GpuSolver::new (crates/qm-gpu/src/scf.rs:1083-1101): device buffers for an n^3 grid at spacing h on the current
 * HIP device. precision: lira_precision. 3 <= n <= 644 (a field plus 4 ghost planes must stay below 2 GiB in f64, the
 * range of the kernels' 32-bit buffer offsets); LIRA_INVALID otherwise. */
lira_status lira_solver_create(int n, double h, int precision, lira_solver **out);
void lira_solver_destroy(lira_solver *s);
const char *lira_last_error(const lira_solver *s);

/* GpuSolver::solve_with (crates/qm-gpu/src/scf.rs:1195-1212): runs the SCF. warm != 0 reuses the previous converged
 * state (or the one an aborted solve parked) when it fits (same grid, counts and scheme). abort (may be NULL) is polled
 * before each SCF step. lira_solve_polled with no inner abort and no observer. */
lira_status lira_solve(lira_solver *s, const lira_system *sys, const lira_settings *settings, int warm,
                       lira_abort_fn abort, void *user, lira_scf_stats *stats);

/* GpuSolver::solve_polled (crates/qm-gpu/src/scf.rs:1213-1235; the caller providers::ks::GpuKs::evaluate_once,
 * crates/providers/src/ks.rs:1617-1621). abort is polled before each SCF step; inner_abort (may be NULL) also once per
 * LOBPCG iteration of the Fermi-Dirac eigensolves. When either returns nonzero the solve is abandoned: LIRA_ABORTED, no
 * converged state, and the in-progress iterate (orbitals, levels, the mixed input densities, V_H) is parked as the next
 * warm start. observe (may be NULL) gets one lira_progress per SCF step after its delta_rho readback; an observer
 * changes no number of the solve. user is passed to all three. */
lira_status lira_solve_polled(lira_solver *s, const lira_system *sys, const lira_settings *settings, int warm,
                              lira_abort_fn abort, lira_abort_fn inner_abort, lira_observe_fn observe, void *user,
                              lira_scf_stats *stats);

/* GpuSolver::forces (crates/qm-gpu/src/forces.rs:239-266): Hellmann–Feynman forces of the last converged state,
 * out[3 * n_nuclei]. sigma must be the solve's. */
lira_status lira_forces(lira_solver *s, const lira_nucleus *nuclei, int n_nuclei, double sigma, double *out);

/* ---- pseudised ions (qm::ionic::IonicSetup, crates/qm/src/ionic.rs:60-70; its tables qm::radial::filter::FilteredTable
 * and FilteredProjector, crates/qm/src/radial/filter.rs:320-338). Tables are uniform in d with spacing du from d = 0,
 * read by 4-point Lagrange with index -1 mirrored (filter.rs:302-318) and 0 beyond their radius. */
typedef struct {
    int l;             /* 0..2 */
    double e_kb;       /* Kleinman-Bylander energy, Ha */
    double radius;     /* t_u is 0 beyond (Bohr) */
    const double *t_u; /* t_l = F_l / d^l */
    int n_t;
} lira_projector;

typedef struct {
    int z;                /* the species' atomic number */
    double z_v;           /* valence charge */
    double sigma_lr;      /* width of the long-range erf (Bohr) */
    double du;            /* table spacing (Bohr) */
    const double *v_sr_u; /* short-range local table */
    int n_v_sr;
    double r_sr;
    const double *rho_pc_u; /* partial core (NULL: none) */
    int n_rho_pc;
    double pc_radius;
    const lira_projector *projectors; /* at most 3, one per non-local channel */
    int n_projectors;
    double r_core; /* the derived core radius (FilteredTable::r_core): two cores closer than the sum of theirs warn */
} lira_species;

/* The ionic setup of the next solves (System::ionic, crates/qm/src/scf.rs:329): species listed here are pseudised,
 * every other Z is all-electron at settings.sigma (Ionic::AllElectron). n_species = 0 clears it. The tables are
 * copied. A warm state converged under another setup is not reused (scf.rs:771-777). The caller passes the valence
 * electron counts in the lira_system it solves, as System::with_ionic does (scf.rs:362-365). */
lira_status lira_set_ionic(lira_solver *s, const lira_species *species, int n_species);

/* ---- the pseudopotential generator (qm::radial, crates/qm/src/radial/recipe.rs:140-176; host code, f64) ----
 * Z in, a norm-conserving Troullier-Martins / Kleinman-Bylander ion with the LFC partial core out, filtered for a grid
 * of spacing h (qm::radial::filter::filtered, filter.rs:384). The ladder tries f in {1.6, 1.4, 1.2, 1.0} (softest
 * first) and, per f, the highest generated channel as the local potential, then the polynomial local
 * (ledger.rs:181-182); the first entry that passes the ghost gates G1/G2 (kb.rs:611-663) and the transferability
 * gates T1-T7 (gates.rs:251-371) is kept. The dimer gate of Z <= 2 is not run: H and He come out as the reference's
 * ungated generate() makes them. f_index (0..3) and local (0, 1) pin one entry instead (-1: the ladder). Z in 1..18,
 * h in [0.02, 1] Bohr.
 * The accepted entry of each (Z, f_index, local) is kept for the process; later calls only filter it for h. */
typedef enum { LIRA_PSEUDO_LOCAL_HIGHEST = 0, LIRA_PSEUDO_LOCAL_POLYNOMIAL = 1 } lira_pseudo_local;

/* The generation record (qm::radial::record::GenerationRecord, crates/qm/src/radial/record.rs; the columns of the
 * reference's pinned records table): f, the local candidate, r_c / eps / T5 per generated channel, E_KB, GSS and the
 * support r_beta per projector, r_nlc (0: no core), r_core, r_cmp, the partial core's support, and h. */
typedef struct {
    int z;
    double f;
    int local;   /* lira_pseudo_local */
    int local_l; /* the channel used as V_loc (-1: the polynomial local) */
    double z_v;
    int n_core;
    int n_channels;
    int channel_l[4];
    double r_c[4];
    double eps[4];
    double t5[4]; /* max |delta E_equiv| over eps_l +- 0.5 Ha (Ha) */
    int n_projectors;
    int proj_l[3];
    double e_kb[3];
    double gss[3];
    double r_beta[3];
    double r_nlc;
    double r_core;
    double r_cmp;
    double r_pc_support;
    double h;
    int provisional;     /* Z <= 2 generated without the dimer gate (the reference's ungated generate: usable by no
                            setup there); 0 for every other table */
    int dimer;           /* the dimer gate's verdict: 0 not run, 1 pass, 2 not applicable (no all-electron force zero in
                            the bracket, the pseudised repulsive wall within 5 % of the all-electron one) */
    double dimer_pp[2];  /* the pseudised dimer's R_e (Bohr) and omega (a.u.) */
    double dimer_ae[2];  /* the all-electron dimer's at sigma -> 0; NaN when not located */
    int dimer_solves;    /* SCF solves the gate ran for this generation */
    double h_req;        /* T8: the coarsest spacing of {0.35, 0.30, 0.25, 0.20, 0.15} from which the filtered tables
                            pass (a)-(d) at every finer one (NaN: none) */
    double h_eb;         /* T8(e): the same for the egg-box force with the nonlinear partial-core term (reported) */
} lira_pseudo_record;

typedef struct lira_pseudo lira_pseudo;
/* qm::radial::generate + filtered (recipe.rs:116, filter.rs:384). LIRA_NUMERICAL with lira_pseudo_error()'s reason
 * when the element is refused (no ladder entry passes, two valence states of one l, a failed radial solve). */
lira_status lira_pseudo_generate(int z, double h, int f_index, int local, lira_pseudo **out);
/* qm::radial::generate_with and providers::pseudo::GridDimerGate (recipe.rs:116-176, crates/providers/src/pseudo.rs:
 * 798-1100): the ladder with the molecular gate on f for Z <= 2. For each entry that passes the atomic gates the
 * homonuclear dimer (two nuclei at +-R/2 on x, singlet) is solved on the current GPU on 81^3 x 0.15 (a 12 Bohr cube),
 * mixed precision, cold, Fermi-Dirac kT 5 mHa, tol 1e-6, eig_tol 1e-7: R_e where the bond force changes sign in
 * [1, 4] r_peak (bisection to 0.002 Bohr, then linear interpolation), k = -dF/dR by +-h, omega = sqrt(2k/m) with m the
 * most abundant isotope's mass; the all-electron dimer's force at each R extrapolated sigma -> 0 from sigma = h x {1,
 * 1.2, 1.4, 1.6}. An entry passes when |R_e - R_e,AE| < 0.02 Bohr and |omega/omega_AE - 1| < 0.05, or when the
 * all-electron dimer has no force zero in the bracket and the pseudised force at its outer end is within 5 % of the
 * all-electron one. For Z > 2 it is lira_pseudo_generate (the gate does not apply). */
lira_status lira_pseudo_generate_gated(int z, double h, lira_pseudo **out);
const char *lira_pseudo_error(void);
lira_status lira_pseudo_record_of(const lira_pseudo *p, lira_pseudo_record *rec);
/* qm::ionic::check, what IonicSetup::build applies to every pseudised species (crates/qm/src/ionic.rs:84-102): LIRA_OK
 * when the table may be used at the spacing it was filtered for: not provisional, h <= h_req (with no h_req on the
 * ladder: h finer than 0.15), and its filtered re-run at h passing T8 (a)-(d); else LIRA_NUMERICAL with the reason
 * (the reference's Provisional, TooCoarse, FilteredGate) in lira_pseudo_error(). lira_set_ionic itself takes tables
 * unchecked, as IonicSetup::from_tables does. */
lira_status lira_pseudo_check(const lira_pseudo *p);
/* The filtered tables as a lira_species for lira_set_ionic (IonicSetup::from_tables, crates/qm/src/ionic.rs); sp and
 * proj[0..n_projectors) point into p, which must outlive them (lira_set_ionic copies them). */
lira_status lira_pseudo_species(const lira_pseudo *p, lira_species *sp, lira_projector *proj, int proj_cap);
void lira_pseudo_free(lira_pseudo *p);
/* The five pseudised force terms of the converged state, out[(t * n_nuclei + I) * 3 + a] for t in [ion-ion, local LR,
 * local SR, nonlocal, partial core] (GpuSolver::force_terms, crates/qm-gpu/src/forces.rs:276; qm::forces::
 * pseudo_force_terms, crates/qm/src/forces.rs:259). lira_forces is their per-component sum. */
lira_status lira_force_terms(lira_solver *s, const lira_nucleus *nuclei, int n_nuclei, double sigma, double *out);
/* The ionic terms of a geometry under the current setup (qm/src/nuclear.rs:342-368): V_loc and rho_pc (n^3 each,
 * f64, either may be NULL); they also become the solver's operator for lira_nonlocal_apply. */
lira_status lira_ionic_fields(lira_solver *s, const lira_nucleus *nuclei, int n_nuclei, double sigma, double *v_loc,
                              double *rho_pc);
/* V_NL x for nb host vectors (band-major n^3, f64) and their projections c (rows x nb, rows = *n_rows) under the
 * geometry of the last lira_ionic_fields or solve (NonlocalOp::apply_add / project, crates/qm/src/nonlocal.rs:133-145). */
lira_status lira_nonlocal_apply(lira_solver *s, const double *x, int nb, double *out, double *c, int c_cap,
                                int *n_rows);

/* ---- provider-facing controls and products of GpuSolver (crates/qm-gpu/src/scf.rs, sample.rs) ---- */
/* GpuSolver::set_witness_cap (scf.rs:1126): at most cap extra vectors per spin under Fermi-Dirac, clamped to 1..32
 * (the provider derives it from its memory budget). lira_witness_cap returns the cap in force (scf.rs:1132). */
lira_status lira_set_witness_cap(lira_solver *s, int cap);
int lira_witness_cap(const lira_solver *s);
/* GpuSolver::set_density_extrapolation (scf.rs:1160): a warm solve that follows two converged solves starts from
 * rho_in = max(2 rho_t - rho_{t-1}, 0); off drops rho_{t-1}; a cold or failed solve drops it too. */
lira_status lira_set_density_extrapolation(lira_solver *s, int on);
/* SolveProfile of the last solve, kept after a failed one (scf.rs:439-514). */
typedef struct {
    int steps, readbacks, poisson_iterations, eigen_iterations, eigen_fallbacks;
    int witness_growths, witness_trims, peak_block, extrapolated;
    double wall, start_wall, poisson_wall, eigen_wall, eig_tol_used;
    /* DIIS restarts (a frozen orbital set re-solved, or a widened witness window that grew the block), freeze trims,
     * frozen Fermi-Dirac steps and eigensolves with a locked head (SolveProfile, crates/qm-gpu/src/scf.rs:420-514) */
    int diis_resets, diis_freeze_trims, frozen_steps, locked_calls;
    /* an abandoned solve: the step whose eigensolve inner_abort stopped, or the step abort came before (else 0) */
    int abandoned_in_step, abandoned_before_step;
} lira_profile;
/* the profile, plus delta_rho and the eigen tolerance of every step (up to cap each; *n_steps the count) */
lira_status lira_get_profile(lira_solver *s, lira_profile *out, double *delta_rho_steps, double *eig_tol_steps,
                             int cap, int *n_steps);
/* GpuSolver::sample_density (crates/qm-gpu/src/sample.rs:84): rho_up + rho_dn of the converged state at count points
 * (3 doubles each, Bohr) by trilinear interpolation with qm::Grid::trilinear's cell and blend order (x, then y, then
 * z; crates/qm/src/grid.rs:54-86); 0 outside [origin, origin + (n-1)h]^3. */
lira_status lira_sample_density(lira_solver *s, const double *points, int count, double *out);

/* GpuSolver::spin_abs_integral and density_max (crates/qm-gpu/src/scf.rs:1834-1845) of the converged state: the
 * integral of |rho_up - rho_dn| (electrons; the provider's spin guard) and the maximum of rho_up + rho_dn (1/Bohr^3). */
lira_status lira_state_summary(lira_solver *s, double *spin_abs_integral, double *density_max);

/* ---- display products of the converged state (crates/qm-gpu/src/orbitals.rs, harris.rs) ---- */
#define LIRA_UNOCC_MAX 34
typedef struct {
    int count[2];                               /* m + 2 guard levels per spin (ascending) */
    double evals[2][LIRA_UNOCC_MAX];            /* re-measured Rayleigh quotients (Ha) */
    double residuals[2][LIRA_UNOCC_MAX];        /* ||H phi - eps phi|| / ||phi|| */
    double occupied_error[2];                   /* max |display - SCF| over the occupied levels */
    int iterations[2];                          /* LOBPCG iterations per spin */
} lira_unoccupied_result;
/* OrbitalView::unoccupied (crates/qm-gpu/src/orbitals.rs:997-1148): the m lowest unoccupied levels (plus 2 guard
 * levels) of each spin channel of the converged state, in the state's own output potentials, from one LOBPCG of the
 * shown occupied columns + m + 2 at the SCF's effective eigen tolerance; the unoccupied vectors are projected onto the
 * complement of the SCF's occupied orbitals, re-measured and sorted, and kept as the next call's guess. orbitals (may
 * be NULL): 2 (m + 2) n^3 f64, spin-major, band-major. LIRA_NUMERICAL when the display's occupied levels disagree
 * with the SCF's by more than tau = max(100 tol, 1e-4) Ha (the state is kept); 0 <= m <= 32. */
lira_status lira_unoccupied(lira_solver *s, int m, lira_unoccupied_result *out, double *orbitals);

#define LIRA_HARRIS_MAX_LEVELS 96
typedef struct {
    int n_levels[2];    /* tested levels kept per spin (ascending): occupied, then unoccupied */
    int n_occ[2];       /* levels with f > 0 (Fermi-Dirac: f >= 1e-6) */
    double evals[2][LIRA_HARRIS_MAX_LEVELS], occ[2][LIRA_HARRIS_MAX_LEVELS];
    double electrons[2]; /* integral of rho_out per spin */
    int iterations[2];  /* LOBPCG iterations per spin */
    int window_short[2]; /* Fermi-Dirac: no witness level within the block at the witness cap */
    int cold;           /* some channel started from fresh vectors */
    double seconds;
} lira_harris_result;
/* GpuSolver::harris (crates/qm-gpu/src/harris.rs:477): one non-self-consistent Kohn-Sham diagonalisation in the
 * potential of the input density rho_in (host f64, n^3 per spin) at the geometry of sys under the solver's ionic setup
 * (V = v_nuc + V_H[rho_in] + v_xc[rho_in (+ rho_pc/2)]), for display: per spin the lowest n_sigma + m + 2 levels at the
 * SCF's effective eigen tolerance (an untested buffer column above them), the occupations of settings applied once
 * (Fermi-Dirac grows the block while no witness level is within it), rho_out = sum f psi^2 (host f64, may be NULL). No
 * energy, no forces, no SCF loop. The previous call's vectors and V_H start the next; the converged SCF state is left
 * as it was. */
lira_status lira_harris(lira_solver *s, const lira_system *sys, const lira_settings *settings, const double *rho_in_up,
                        const double *rho_in_dn, int m, lira_harris_result *out, double *rho_out_up,
                        double *rho_out_dn);

/* ---- the provider's evaluation around GpuSolver (providers::ks::GpuKs::evaluate_once) ---- */
/* max(tol, 2 * 5e-8 * n_valence): the SCF tolerance the GPU's f32 density can reach (effective_gpu_tol,
 * crates/providers/src/pseudo.rs:423-440). Host only. */
double lira_effective_gpu_tol(double tol, int n_valence);
typedef enum {
    LIRA_EVAL_TOL_RAISED = 1,      /* ionic setup: tol raised to lira_effective_gpu_tol */
    LIRA_EVAL_RETRIED_COLD = 2,    /* the warm attempt failed and was retried cold (R15) */
    LIRA_EVAL_WARM_STALLED = 4,    /* ... and that warm attempt had stalled (else it ran out of steps) */
    LIRA_EVAL_EXTRAPOLATED = 8,    /* the converged solve started from an extrapolated density */
    LIRA_EVAL_DEVICE_RETRIED = 16, /* a device failure was retried after a pause ("GPU busy") */
    LIRA_EVAL_BUDGET_CAP = 32      /* the witness cap came from the device memory budget */
} lira_eval_flag;
/* The device memory (bytes) one solve holds on an n^3 grid with n_up / n_dn electrons, at most `witness` witness vectors
 * per spin and a DIIS history of diis_history (the provider's footprint model, crates/providers/src/ks.rs:360-400,
 * for this library's buffers). 0 for bad arguments. Host only. */
unsigned long long lira_solve_footprint_bytes(int n, int precision, int n_up, int n_dn, int witness,
                                              int diis_history);
/* witness_cap_for (crates/providers/src/ks.rs:400-428): the largest witness cap w <= 32 whose solve footprint fits
 * budget bytes, never below 3 (the witness vectors a solve starts with). Host only. */
int lira_witness_cap_for_budget(int n, int precision, int n_up, int n_dn, int diis_history,
                                unsigned long long budget);
/* One evaluation as the reference's provider drives the solver (GpuKs::evaluate / evaluate_once,
 * crates/providers/src/ks.rs:1573-1706): with vram_budget > 0 the witness cap is set from the budget
 * (lira_witness_cap_for_budget); under an ionic setup the tolerance is raised to the GPU floor (raise_to_gpu_floor,
 * ks.rs:543-558); the solve runs through lira_solve_polled with abort before each step, inner_abort per Fermi-Dirac
 * LOBPCG iteration and observe per step; a warm solve that ends LIRA_NOT_CONVERGED or LIRA_STALLED is retried once cold
 * with density extrapolation off for that solve (retry_cold_on_stall, ks.rs:585-601), the failed attempt's steps and
 * readbacks added to *stats; then the forces (3 n doubles). A LIRA_DEVICE failure anywhere in the attempt is retried
 * after 100 ms, then 400 ms (retry_on_device_fault, ks.rs:317-350), abort polled every 20 ms meanwhile; a witness
 * search stopped at a budget's cap says so in lira_last_error. Returns the last attempt's status; *flags
 * (lira_eval_flag) and *tol_used are set in every case. */
lira_status lira_evaluate(lira_solver *s, const lira_system *sys, const lira_settings *settings, int warm,
                          lira_abort_fn abort, lira_abort_fn inner_abort, lira_observe_fn observe, void *user,
                          unsigned long long vram_budget, lira_scf_stats *stats, double *forces, int *flags,
                          double *tol_used);

/* ---- batched SCF (no counterpart in the reference, which evaluates a set of molecules one at a time:
 * crates/providers/src/sampling/mod.rs:1635) ---- */
#define LIRA_MAX_GRID 644
#define LIRA_BATCH_MAX_WORKERS 64
/* one all-electron molecule on its own grid */
typedef struct {
    int n;              /* grid points per edge, 3 .. LIRA_MAX_GRID */
    double h;           /* spacing (Bohr) */
    lira_system system; /* at least one nucleus */
} lira_batch_job;
typedef struct {
    lira_status status; /* of this molecule (LIRA_ABORTED: no worker took it) */
    lira_scf_stats stats;
    int worker;         /* the worker thread that ran it */
    double start, end;  /* seconds from the start of the batch: its wait in the queue ends at start */
    char message[128];  /* lira_last_error of a failed molecule */
} lira_batch_result;
/* Runs a cold SCF and the forces for each of n_jobs molecules on the current HIP device. workers host threads
 * (1 .. LIRA_BATCH_MAX_WORKERS) each own one solver (and so one HIP stream) and take molecules from a shared queue,
 * largest grids first; the solver is rebuilt when a worker's grid changes. Each molecule's result is the one a lone
 * lira_solve + lira_forces on a fresh solver gives, bit for bit. settings: n_jobs entries, or NULL for
 * lira_default_settings(h) per molecule. forces: 3 * n_nuclei doubles per molecule, concatenated in job order, or NULL.
 * Returns LIRA_INVALID for bad arguments (before any work), LIRA_DEVICE if some molecule found no worker able to select
 * the device, else LIRA_OK with each molecule's status in results. */
lira_status lira_solve_batch(const lira_batch_job *jobs, int n_jobs, int precision, const lira_settings *settings,
                             int workers, lira_batch_result *results, double *forces);

/* Dispatch control, no counterpart in the reference (its recorder batches dispatches instead,
 * crates/qm-gpu/src/batch.rs:55-233): on, the fixed kernel sequences between two host syncs (the Poisson MG-PCG
 * iteration, in its two forms, and its start; each LOBPCG iteration's segment from the residuals through the
 * preconditioner, H W, H P and the KB projections to the Gram product) are captured once per buffer arrangement and
 * block size as HIP graphs and replayed; off (the default), every
 * kernel is launched one by one. Results are bitwise identical either way. Turn it on only where no other host thread
 * allocates device memory or launches work while this solver captures (a capture is open during the first call of
 * each sequence); with solvers on several threads keep it off. */
lira_status lira_set_graphs(lira_solver *s, int on);
/* LOBPCG's Rayleigh-Ritz step (and its stop test) on the device from the second iteration of each eigensolve on, for
 * blocks whose subspace [X, W, P] has at most 36 columns (off by default; larger blocks keep the host step). The host
 * then reads each iteration's record one iteration late while the next one runs; an iteration queued past the stop
 * leaves the block bit for bit as it was. The results equal the host step's to rounding. */
lira_status lira_set_device_rr(lira_solver *s, int on);

/* The spherical spin-restricted LSDA free atom behind the Fermi-Dirac cold guess (qm::radial::atom::ground_state,
 * crates/qm/src/radial/atom.rs:474; used by free_atom_density, crates/qm/src/nuclear.rs:153): eps receives the
 * 1s, 2s, 3s, 2p, 3p eigenvalues (1.0 for an unbound level) and *electrons the density's electron count. Host only. */
lira_status lira_free_atom_levels(int z, double *eps, double *electrons);

/* state getters (GpuSolver::state / ks_state / density_f32, crates/qm-gpu/src/scf.rs:1172, 1763-1851) as f64 in the
 * qm layout. which: 0 = rho_up, 1 = rho_dn, 2 = rho, 3 = V_H, 4 = V_nuc (V_loc under an ionic setup), 5 = rho_pc. */
lira_status lira_get_field(lira_solver *s, int which, double *out);
/* eigenvalues / occupations of a spin channel (0 up, 1 down); *count receives the stored count */
lira_status lira_get_levels(lira_solver *s, int spin, double *evals, double *occ, int cap, int *count);
/* orbitals of a spin channel, cap * n^3 doubles (band-major) */
lira_status lira_get_orbitals(lira_solver *s, int spin, double *out, int cap, int *count);
/* GpuSolver::effective_eig_tol (crates/qm-gpu/src/scf.rs:1713): the eigen tolerance the precision can reach */
double lira_effective_eig_tol(const lira_solver *s, double requested);
/* drops the warm state (the next solve is cold) */
void lira_reset(lira_solver *s);
/* Fault injection for the device-failure path (the reference's readback-fault latch, crates/qm-gpu/src/grid.rs:113-157,
 * and its test hook, crates/gpu/src/readback.rs:37-69): the readback `after` readbacks from now (0 = the next one)
 * completes its copy, then reports a failure, which the call in progress returns as LIRA_DEVICE with no state kept.
 * Host side only: every kernel runs and exits normally. after < 0 disarms it. */
lira_status lira_inject_readback_fault(lira_solver *s, int after);

/* --------------------------------------------------------------------------------- sub-solvers ---- */

/* GpuPoisson::solve (crates/qm-gpu/src/poisson.rs:284-342): del^2 V = -4 pi rho with the multipole boundary
 * (l <= 6); rho and v are host f64 arrays (v: initial guess in, solution out). */
lira_status lira_poisson(lira_solver *s, const double *rho, double *v, double tol, int max_iter,
                         lira_poisson_stats *stats);

/* Multigrid::apply_in (crates/qm-gpu/src/multigrid.rs:107-197): z = M r, one V-cycle of the Poisson solve's
 * preconditioner (V(2,2), omega = 6/7, odd-index coarsening with boundary-distance ghosts, 12 coarsest sweeps) on the
 * solver's grid and in its precision; r and z are host f64 arrays (n^3). */
lira_status lira_mg_vcycle(lira_solver *s, const double *r, double *z);

/* One level operator of the V-cycle on device pointers (crates/qm-gpu/src/shaders/multigrid.wgsl:30-126). A2 is the
 * 7-point -lap at spacing h whose missing neighbours are ghosts (1 - 1/d) x0 at boundary distance d_low / d_high (1 on
 * the finest level); D its diagonal. LIRA_MG_JACOBI: out = x + (6/7) D^-1 (b - A2 x); LIRA_MG_RESIDUAL: out = b - A2 x;
 * LIRA_MG_RESTRICT: out ((n/2)^3) = full weighting of x (n^3) about the odd fine points; LIRA_MG_PROLONG_ADD: out
 * (n^3) += trilinear interpolation of x ((n/2)^3). Waits for the stream before returning. */
enum { LIRA_MG_JACOBI = 0, LIRA_MG_RESIDUAL = 1, LIRA_MG_RESTRICT = 2, LIRA_MG_PROLONG_ADD = 3 };
lira_status lira_mg_level_op(int dtype, int op, int n, double h, double d_low, double d_high, const void *x,
                             const void *b, void *out, void *stream);

/* lobpcg (crates/qm-gpu/src/eigen.rs:595-653) for H = -1/2 lap + V with the damped-Jacobi preconditioner shifted by
 * precond_lambda: V (n^3), x (k * n^3, guess in, orthonormal eigenvectors out), evals (k). Host f64 arrays. */
lira_status lira_lobpcg(lira_solver *s, const double *v, int k, double *x, double *evals, double precond_lambda,
                        double tol, int max_iter, lira_eigen_stats *stats);

/* Rayleigh–Ritz on a measured pair (A, B), m x m row-major: the lowest nk Ritz values theta and B-orthonormal
 * coefficient columns C (m x nk, row-major) (crates/qm-gpu/src/eigen.rs:515-542). Host-only. */
lira_status lira_rayleigh_ritz(int m, const double *a, const double *b, int nk, double *theta, double *c);
/* The same step on the device (the f64 single-workgroup kernel the solver's LOBPCG uses from its second iteration on,
 * for subspaces of at most 36 columns): theta and C as lira_rayleigh_ritz; also c_p (m x nk, may be NULL), the
 * coefficients of C's component B-orthogonal to the first nk columns, C - [B11^-1 B1: C; 0] (the P update of LOBPCG).
 * Host arrays; LIRA_INVALID for m > 36, LIRA_NUMERICAL when the subspace lost rank or B11 is not positive definite. */
lira_status lira_rayleigh_ritz_device(int m, const double *a, const double *b, int nk, double *theta, double *c,
                                      double *c_p);

/* ----------------------------------------------------------------- kernels on device pointers ---- */
/* dtype: 4 = f32, 8 = f64. Pointers are device pointers; stream is a hipStream_t (NULL = default stream).
 * Band-major blocks: band b of a block starts at b * n^3. Outputs must not alias inputs. */

/* hamiltonian/laplacian (crates/qm-gpu/src/stencil.rs:43-121): hx_b = -1/2 lap x_b + v x_b */
lira_status lira_hamiltonian(int dtype, int n, double h, int nbands, const void *v, const void *x, void *hx,
                             void *stream);
/* out_b = scale * lap x_b (8th order, Dirichlet) */
lira_status lira_laplacian(int dtype, int n, double h, int nbands, double scale, const void *x, void *out,
                           void *stream);
/* one damped-Jacobi sweep of the preconditioner (crates/qm-gpu/src/stencil.rs:62-73):
 * out_b = w_b + omega / max(diag_t + v - lam_b, 0.1) (r_b - (H - lam_b) w_b), lam device f64[nbands] */
lira_status lira_jacobi_sweep(int dtype, int n, double h, int nbands, const void *v, const void *w, const void *r,
                              const double *lam, void *out, void *stream);
/* xc_lsda (crates/qm-gpu/src/physics.rs:86-114): e = n eps_xc, vu, vd from (nu, nd), count points */
lira_status lira_xc_lsda(int dtype, long count, const void *nu, const void *nd, void *e, void *vu, void *vd,
                         void *stream);
/* nuclear_potential (crates/qm-gpu/src/physics.rs:48-76), nuclei on the host */
lira_status lira_nuclear_potential(int dtype, int n, double h, const lira_nucleus *nuclei, int n_nuclei, double sigma,
                                   void *out, void *stream);
/* tall-skinny Gram G = h^3 S^T U for band-major blocks S (m bands) and U (q bands); G is host f64 m x q. */
lira_status lira_gram(int dtype, int n, double h, int m, const void *s_block, int q, const void *u_block, double *g,
                      void *stream);

/* No counterpart in the reference: lira_gram for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). G = h^3 S^T U goes to g_dev, a device buffer of
 * m x q f64 (row-major), on `stream`, without waiting; the per-chunk partial sums live in a fixed device array of
 * LIRA_NOALLOC_SCRATCH_BYTES on each GPU the library runs on (held until the process ends), so the call allocates
 * nothing. The result is bitwise lira_gram's. A shape whose partials do not fit the array returns LIRA_INVALID before
 * any device work. Calls on one GPU share its array: calls that overlap in time on one GPU race (order them on one
 * stream, or with events). */
#define LIRA_NOALLOC_SCRATCH_BYTES (16u << 20)
lira_status lira_gram_noalloc(int dtype, int n, double h, int m, const void *s_block, int q, const void *u_block,
                              double *g_dev, void *stream);

/* No counterpart in the reference: the device's streaming ceiling, from hand-written kernels of 16-byte vector loads
 * and stores. kind 0 copies src to dst (2 x bytes moved), 1 reads src (bytes moved), 2 writes dst (bytes moved);
 * nontemporal != 0 uses nontemporal stores. bytes a multiple of 16, pointers 16-byte aligned, device memory. Runs two
 * untimed launches, then reps launches timed with HIP events on `stream`; *gbs = bytes moved / mean launch time. */
lira_status lira_stream_bandwidth(int kind, int nontemporal, const void *src, void *dst, unsigned long long bytes,
                                  int reps, void *stream, double *gbs);

/* stencil tile variants and the autotuner (the tile shape is picked per grid by timing) */
int lira_stencil_variant_count(void);
const char *lira_stencil_variant_name(int v);
/* times every variant on (dtype, n, nbands) and makes the fastest the default for that dtype; returns its index */
int lira_autotune_stencil(int dtype, int n, int nbands, double *ms_out);
void lira_set_stencil_variant(int dtype, int variant);

const char *lira_version(void);

#ifdef __cplusplus
}
#endif
#endif
