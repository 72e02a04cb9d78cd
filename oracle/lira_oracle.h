/*
 * lira_oracle.h — CPU f64 restatement of the reference's CPU Kohn–Sham solver (crates/qm).
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP product path: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it. The product never calls it.
 *
 * Every function cites the reference file:line whose algorithm it restates. Grids are cubic
 * n^3 with spacing h (Bohr), centred at the origin (origin = -(n-1)h/2), layout idx = i + n(j + n k).
 * Energies in Hartree, lengths in Bohr.
 */
#ifndef LIRA_ORACLE_H
#define LIRA_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define OR_L_MAX 6
#define OR_N_MOMENTS 49

/* status codes (qm/src/error.rs:1-33) */
#define OR_OK 0
#define OR_COINCIDENT 1
#define OR_NOT_CONVERGED 2
#define OR_STALLED 3
#define OR_NUMERICAL 4

/* grid and sums: qm/src/grid.rs:10-99 */
double or_neumaier(const double *x, long len);
double or_dot(int n, double h, const double *a, const double *b);
double or_integrate(int n, double h, const double *a);

/* stencil: qm/src/stencil.rs:21-53 ; H = -1/2 lap + V (qm/src/eigen.rs:124-129) */
void or_laplacian(int n, double h, const double *f, double *out);
void or_kinetic(int n, double h, const double *f, double *out);
void or_hamiltonian(int n, double h, const double *v, const double *x, double *out);
double or_jacobi_omega(void);
/* qm/src/eigen.rs:227-240: 4 damped-Jacobi sweeps on (H - lambda) w = r */
void or_jacobi_precond(int n, double h, const double *v, double lambda, const double *r, double *w);

/* Poisson: qm/src/poisson.rs:89-259 */
void or_solid_harmonics(const double r[3], double out[OR_N_MOMENTS]);
void or_multipole_moments(int n, double h, const double *rho, double q[OR_N_MOMENTS]);
/* the ghost value V_bc(r) of the multipole expansion (poisson.rs:175-190) */
double or_multipole_potential(const double q[OR_N_MOMENTS], const double r[3]);
/* b = 4 pi rho + L_bc (poisson.rs:210-214) */
void or_poisson_rhs(int n, double h, const double *rho, double *b);
/* returns iterations; *residual = ||b - A v|| / ||b|| at exit */
int or_poisson_solve(int n, double h, const double *rho, double *v, double tol, int max_iter, double *residual);
/* OpenMP threads of the grid loops (set / query) */
void or_set_threads(int threads);
int or_threads(void);

/* nuclei: qm/src/nuclear.rs:106-289 */
void or_deposit_charge(int n, double h, int nat, const double *pos, const int *z, double sigma, double *out);
void or_nuclear_potential(int n, double h, int nat, const double *pos, const int *z, double sigma, double *out);
/* returns OR_OK or OR_COINCIDENT (pair in *pi, *pj) */
int or_nuclear_repulsion(int nat, const double *pos, const int *z, double *e, int *pi, int *pj);

/* LSDA: qm/src/xc.rs:120-210 */
void or_exc_point(double nu, double nd, double *e, double *vu, double *vd);
void or_lsda(long len, const double *nu, const double *nd, double *e, double *vu, double *vd);
/* PW92 G(rs) and dG/drs for the PARA (0), FERRO (1) or MALPHA (2) fit (xc.rs:120-128) */
void or_pw92_g(double rs, int which, double *g, double *dg);

/* dense symmetric eigen (cyclic Jacobi; qm/src/eigen.rs:175-211). a is m x m row-major, destroyed.
 * evals ascending, evecs column k = eigenvector k (row-major m x m). */
void or_sym_eigen(int m, double *a, double *evals, double *evecs);

/* LOBPCG with full MGS (qm/src/eigen.rs:290-369). x: k vectors of n^3 (band-major), guess in, result out.
 * precond_lambda: NaN = no preconditioner. residuals (k) may be NULL. Returns iterations. */
int or_lobpcg(int n, double h, const double *v, int k, double *x, double *evals, double precond_lambda,
              double tol, int max_iter, double *max_residual, double *residuals);

/* Fermi–Dirac occupations of one channel (qm/src/occupations.rs:166-213). Returns OR_OK / OR_NUMERICAL. */
int or_fermi_dirac(int m, const double *evals, int n, double kt, double *f, double *mu, double *entropy);
double or_fd_entropy(double x);

/* the reference's deterministic orbital guess (qm/src/scf.rs:626-634) */
void or_initial_orbitals(int n, double h, int nat, const double *pos, int k, unsigned long long seed0, double *out);

/* SCF (qm/src/scf.rs:928-1098) */
typedef struct {
    double sigma, tol, mixing, eig_tol, kt; /* kt used when occupations == 1 */
    int max_iter, diis_history, occupations; /* 0 = Aufbau, 1 = Fermi–Dirac */
    int stall_window;
} or_settings;

typedef struct {
    int status, iterations;
    double delta_rho;
    double kinetic, external_, hartree, xc, nuclear, internal, ts, total;
    double boundary_density, max_eigen_residual, poisson_residual;
    int poisson_iterations;
    double entropy, mu[2], electrons[2];
    int k_up, k_dn; /* stored orbitals per spin */
    double seconds_poisson, seconds_eigen, seconds_total;
    int lobpcg_iterations, cg_iterations; /* totals over the solve (both spins; every Poisson solve of the loop) */
} or_scf_stats;

/* Runs the SCF. evals_up/evals_dn/occ_up/occ_dn have room for `cap` entries; rho_up_out/rho_dn_out (n^3, may be
 * NULL) receive the output densities; forces_out (3*nat, may be NULL) receives the Hellmann–Feynman forces of the
 * converged state (qm/src/forces.rs:189-236). orbitals_up_out/orbitals_dn_out (cap * n^3, may be NULL) receive the
 * orbitals. */
int or_scf(int n, double h, int nat, const double *pos, const int *z, int n_up, int n_dn, const or_settings *s,
           or_scf_stats *st, int cap, double *evals_up, double *evals_dn, double *occ_up, double *occ_dn,
           double *rho_up_out, double *rho_dn_out, double *forces_out, double *orbitals_up_out,
           double *orbitals_dn_out);

/* HF forces of a density (qm/src/forces.rs:189-236), electronic part + nuclear repulsion */
int or_forces(int n, double h, const double *rho, int nat, const double *pos, const int *z, double sigma,
              double *out);

/* ------------------------------------------------------------------------------------------------------------
 * Pseudised ions: the Kleinman–Bylander nonlocal operator, the ionic local potential and the partial core
 * (qm/src/nonlocal.rs, qm/src/ionic.rs, qm/src/nuclear.rs:290-382, qm/src/radial/filter.rs:302-363).
 * A species carries the uniform tables of qm::radial::filter::FilteredTable (spacing du, index 0 at d = 0);
 * a Z with no species entry is all-electron at sigma_ae (Ionic::AllElectron).
 * ------------------------------------------------------------------------------------------------------------ */
typedef struct {
    int l;            /* 0..2 */
    double e_kb;      /* Ha */
    double radius;    /* t_u is 0 beyond (FilteredProjector::radius) */
    const double *t_u; /* t_l = F_l / d^l on the uniform table */
    int n_t;
} or_projector;

typedef struct {
    int z;
    double z_v, sigma_lr, du;
    const double *v_sr_u;
    int n_v_sr;
    double r_sr;
    const double *rho_pc_u; /* may be NULL when pc_radius <= 0 */
    int n_pc;
    double pc_radius; /* <= 0: no partial core */
    const or_projector *proj;
    int n_proj;
} or_species;

typedef struct {
    const or_species *species;
    int n_species;
    double sigma_ae;
} or_ionic;

/* radial/filter.rs:302-318, 355-363: (value, d/dd) of the 4-point Lagrange interpolant, 0 beyond radius */
void or_table_read(const double *f, int len, double du, double radius, double d, double *val, double *der);
/* nuclear.rs:322-337: grid points with |r_p - c| <= radius, ascending; returns the count (out may be NULL) */
long or_points_within(int n, double h, const double c[3], double radius, long *out);
/* nuclear.rs:342-350 */
void or_ionic_local_potential(int n, double h, int nat, const double *pos, const int *z, const or_ionic *io,
                              double *out);
/* nuclear.rs:354-363; returns 1 if any centre carries a partial core (out is filled either way) */
int or_partial_core_field(int n, double h, int nat, const double *pos, const int *z, const or_ionic *io,
                          double *out);
/* nuclear.rs:371-382 */
int or_ionic_repulsion(int nat, const double *pos, const int *z, const or_ionic *io, double *e, int *pi, int *pj);
double or_valence_charge(const or_ionic *io, int z);
/* nonlocal.rs:63-73: gradients of R_lm for l <= 2 in lm_index order */
void or_solid_harmonic_gradients(const double r[3], double out[9][3]);

/* nonlocal.rs:81-174: spheres in (nucleus, projector) order; c has or_nl_rows() entries, (sphere, m) order */
typedef struct or_nl or_nl;
or_nl *or_nl_build(int n, double h, int nat, const double *pos, const int *z, const or_ionic *io, int with_grad);
void or_nl_free(or_nl *nl);
int or_nl_rows(const or_nl *nl);
int or_nl_spheres(const or_nl *nl);
long or_nl_sphere_points(const or_nl *nl, int s);
void or_nl_project(const or_nl *nl, const double *psi, double *c);
void or_nl_apply_add(const or_nl *nl, const double *psi, double *out);
/* sum_i f_i sum_lm E_KB c_lm[psi_i]^2, f == 0 orbitals skipped (f NULL: every f = 1) */
double or_nl_energy(const or_nl *nl, int k, const double *orbitals, const double *f);
/* H x = -1/2 lap x + v x (+ V_NL x when nl is not NULL) (nonlocal.rs:165-174) */
void or_pseudo_hamiltonian(int n, double h, const double *v, const or_nl *nl, const double *x, double *out);
int or_lobpcg_nl(int n, double h, const double *v, const or_nl *nl, int k, double *x, double *evals,
                 double precond_lambda, double tol, int max_iter, double *max_residual, double *residuals);

/* forces.rs:259-350: terms[t][I][a], t in [ion-ion, local LR, local SR, nonlocal, partial core]; orbitals are the
 * k spin orbitals (up then down) with occupations occ */
int or_pseudo_force_terms(int n, double h, const double *rho_up, const double *rho_dn, int k, const double *orbitals,
                          const double *occ, int nat, const double *pos, const int *z, const or_ionic *io,
                          double *terms);

/* or_scf under an ionic setup (scf.rs:115-135, 720-770, 1060-1075); io NULL is or_scf */
int or_scf_ionic(int n, double h, int nat, const double *pos, const int *z, int n_up, int n_dn, const or_settings *s,
                 const or_ionic *io, or_scf_stats *st, int cap, double *evals_up, double *evals_dn, double *occ_up,
                 double *occ_dn, double *rho_up_out, double *rho_dn_out, double *forces_out, double *orbitals_up_out,
                 double *orbitals_dn_out);

/* qm::Grid::trilinear (qm/src/grid.rs:54-86): f at p by trilinear interpolation, 0 outside the grid */
double or_trilinear(int n, double h, const double *f, const double p[3]);

/* multigrid f64 helpers (qm-gpu/src/multigrid.rs:207-258): the CPU model of the GPU V-cycle */
void or_mg_apply_a2(int n, double h, double d_low, double d_high, const double *x, double *out);
void or_mg_restrict(int n, const double *f, double *out);
void or_mg_prolong(int n, const double *c, double *out);

#ifdef __cplusplus
}
#endif
#endif
