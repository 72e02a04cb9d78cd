"""GPU parity of the Poisson, LOBPCG and SCF solvers against the f64 oracle and the reference's own pins.

Tolerances: FP64 against the oracle's converged answers to the solvers' tolerances; MIXED (f32 grid, f64 control)
to the reference's GPU-vs-CPU parity bounds (crates/validate/tests/gpu_vs_cpu.rs:152-160, 220-221:
|dE| <= 2e-6 Ha for H2, 3e-6 for LiH; |dF| <= 5e-6 Ha/Bohr for H2).
"""
import math

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lira_mod():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    import lira
    lira.lib()
    return lira


def _gaussian(n, h, q, sigma, c):
    x, y, z = O.grid_points(n, h)
    r2 = (x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2
    rho = q / (2 * math.pi * sigma * sigma) ** 1.5 * np.exp(-r2 / (2 * sigma * sigma))
    from scipy.special import erf
    r = np.sqrt(r2)
    with np.errstate(divide="ignore", invalid="ignore"):
        exact = np.where(r < 1e-12, q * math.sqrt(2 / math.pi) / sigma, q * erf(r / (sigma * math.sqrt(2))) / r)
    return rho, exact


# ------------------------------------------------------------------------------------------- Poisson ----

@pytest.mark.parametrize("precision", [0, 1])
def test_poisson_matches_oracle_and_closed_form(lira_mod, precision):
    lira = lira_mod
    n, h = 49, 0.25
    rho, exact = _gaussian(n, h, 1.0, 0.5, [1.3, -0.7, 0.4])
    s = lira.GpuSolver(n, h, precision)
    v, st = s.poisson(rho, tol=1e-10)
    print(f"poisson precision {precision}: {st.iterations} PCG iterations, {st.outer} corrections, "
          f"residual {st.residual:.2e}, {st.seconds * 1e3:.1f} ms")
    assert st.residual <= 1e-10
    vo, _, res = O.poisson_solve(n, h, rho, tol=1e-10, max_iter=3000)
    # both solve the same discrete system to 1e-10: the solutions agree to the system's conditioning
    assert np.abs(v - vo).max() < 5e-8
    assert np.abs(v - exact).max() < 1e-4  # qm/src/poisson.rs:304-307
    eh = 0.5 * np.sum(rho * v) * h ** 3
    # MIXED stores rho in f32 (relative rounding 6e-8 of the right-hand side)
    assert abs(eh - O.dot(n, h, rho, vo) * 0.5) < (1e-9 if precision == 0 else 1e-7)
    # a warm start from the solution needs no correction
    v2, st2 = s.poisson(rho, v0=v, tol=1e-10)
    assert st2.iterations == 0 and np.array_equal(v2, v)


def test_poisson_hartree_energy_of_gaussian(lira_mod):
    lira = lira_mod
    n, h = 49, 0.25
    rho, _ = _gaussian(n, h, 1.0, 0.5, [0, 0, 0])
    v, st = lira.GpuSolver(n, h, 1).poisson(rho, tol=1e-10)
    eh = 0.5 * np.sum(rho * v) * h ** 3
    assert abs(eh - 1.0 / (2 * 0.5 * math.sqrt(math.pi))) < 1e-4


def test_poisson_nan_density_is_numerical_error(lira_mod):
    lira = lira_mod
    n, h = 17, 0.3
    rho = np.zeros(n ** 3)
    rho[100] = float("nan")
    s = lira.GpuSolver(n, h, 0)
    with pytest.raises(lira.Numerical):
        s.poisson(rho)


# -------------------------------------------------------------------------------------------- LOBPCG ----

@pytest.mark.parametrize("precision,tol", [(0, 1e-6), (1, 2e-5)])
def test_lobpcg_oscillator(lira_mod, precision, tol):
    # 3D harmonic oscillator: 1.5, 2.5 (x3) (crates/qm-gpu/src/eigen.rs:1008-1033)
    lira = lira_mod
    n, h = 41, 0.25
    x, y, z = O.grid_points(n, h)
    v = 0.5 * (x * x + y * y + z * z)
    guess = O.initial_orbitals(n, h, [[0.0, 0.0, 0.0]], 4, 12345)
    s = lira.GpuSolver(n, h, precision)
    xs, ev, st = s.lobpcg(v, guess, tol, 500, precond_lambda=1.5)
    print(f"lobpcg precision {precision}: {st.iterations} iterations, residual {st.max_residual:.2e}, "
          f"fallbacks {st.fallbacks}, evals {ev}")
    assert st.max_residual < tol
    assert abs(ev[0] - 1.5) < 1e-4
    assert np.abs(ev[1:] - 2.5).max() < 1e-4
    g = xs @ xs.T * h ** 3
    assert np.abs(g - np.eye(4)).max() < (1e-8 if precision == 0 else 1e-5)
    # the same discrete problem in the oracle: eigenvalues agree to the residual^2 / gap scale
    _, evo, _, _, _ = O.lobpcg(n, h, v, guess, 1e-8, 500, precond_lambda=1.5)
    assert np.abs(ev - evo).max() < (1e-10 if precision == 0 else 2e-7)


def test_lobpcg_wide_block_anisotropic_oscillator(lira_mod):
    # k = 24 columns (S = [X, W, P] holds 72 > 64 columns): the lowest 24 levels of an anisotropic oscillator
    # E = sum_a w_a (n_a + 1/2), w = (1, 1.13, 1.31): distinct levels, each to the discretisation error
    lira = lira_mod
    n, h = 41, 0.3
    x, y, z = O.grid_points(n, h)
    w = (1.0, 1.13, 1.31)
    v = 0.5 * (w[0] ** 2 * x * x + w[1] ** 2 * y * y + w[2] ** 2 * z * z)
    k = 24
    levels = sorted(w[0] * (a + 0.5) + w[1] * (b + 0.5) + w[2] * (c + 0.5)
                    for a in range(8) for b in range(8) for c in range(8))[:k]
    guess = O.initial_orbitals(n, h, [[0.0, 0.0, 0.0]], k, 777)
    s = lira.GpuSolver(n, h, 0)
    xs, ev, st = s.lobpcg(v, guess, 1e-6, 800, precond_lambda=1.5)
    print(f"wide lobpcg: {st.iterations} iterations, residual {st.max_residual:.2e}")
    assert st.max_residual < 1e-6
    assert np.abs(np.sort(ev) - np.array(levels)).max() < 2e-3
    g = xs @ xs.T * h ** 3
    assert np.abs(g - np.eye(k)).max() < 1e-8


def test_lobpcg_block_beyond_64_columns(lira_mod):
    # k = 66 (S = [X, W, P] holds 198 > 192 columns, the size a 34-electron Fermi-Dirac solve reaches with its 32
    # witness columns): the lowest 66 levels of the anisotropic oscillator, in mixed precision
    lira = lira_mod
    n, h = 33, 0.35
    x, y, z = O.grid_points(n, h)
    w = (1.0, 1.13, 1.31)
    v = 0.5 * (w[0] ** 2 * x * x + w[1] ** 2 * y * y + w[2] ** 2 * z * z)
    k = 66
    levels = sorted(w[0] * (a + 0.5) + w[1] * (b + 0.5) + w[2] * (c + 0.5)
                    for a in range(12) for b in range(12) for c in range(12))[:k]
    guess = O.initial_orbitals(n, h, [[0.0, 0.0, 0.0]], k, 778)
    s = lira.GpuSolver(n, h, 1)
    xs, ev, st = s.lobpcg(v, guess, 1e-4, 600, precond_lambda=1.5)
    print(f"k = 66 lobpcg: {st.iterations} iterations, residual {st.max_residual:.2e}, "
          f"max level error {np.abs(np.sort(ev) - np.array(levels)).max():.2e}")
    assert st.max_residual < 1e-4
    assert np.abs(np.sort(ev) - np.array(levels)).max() < 3e-2
    g = xs @ xs.T * h ** 3
    assert np.abs(g - np.eye(k)).max() < 1e-5


# ----------------------------------------------------------------------------------------------- SCF ----

H2 = ([[0.7, 0.0, 0.0], [-0.7, 0.0, 0.0]], [1, 1], 1, 1)
LIH = ([[-0.8, 0.1, 0.0], [2.2, 0.0, 0.0]], [3, 1], 2, 2)


def _scf(lira, precision, system, occ=0, n=33, h=0.25):
    pos, z, nu, nd = system
    s = lira.GpuSolver(n, h, precision)
    st = s.solve(pos, z, nu, nd, lira.default_settings(h, occ))
    f = s.forces(pos, z, h)
    return s, st, f


# the pins are the reference CPU solver's stopping point (delta_rho < 1e-6, one preconditioner shift for every column);
# this solver shifts each column by its own Ritz value as the reference's GPU solver does, so it stops at another
# point within the same tolerance: energies agree to second order (1e-9), forces to first order (the reference's own
# Aufbau and FD H2 forces differ by 6.3e-7, qm/src/scf.rs:1277, 1291)
@pytest.mark.parametrize("precision,e_tol,f_tol", [(0, 1e-9, 1e-6), (1, 2e-6, 5e-6)])
def test_scf_h2_matches_reference_pins(lira_mod, precision, e_tol, f_tol):
    # qm/src/scf.rs:1276-1277: E = -0.952449176828, F_x = +0.128755968665 (33^3 x 0.25, Aufbau, 10 CPU steps)
    lira = lira_mod
    s, st, f = _scf(lira, precision, H2)
    e = st.energy.total
    print(f"H2 precision {precision}: E {e:.12f} F_x {f[0, 0]:+.12f} steps {st.iterations} lobpcg "
          f"{st.lobpcg_iterations} cg {st.cg_iterations} readbacks {st.readbacks} {st.seconds_total * 1e3:.1f} ms")
    assert abs(e - (-0.952449176828)) < e_tol
    assert abs(f[0, 0] - 0.128755968665) < f_tol
    assert abs(f[0, 0] + f[1, 0]) < 10 * f_tol
    assert abs(st.electrons[0] - 1) < 1e-6 and abs(st.electrons[1] - 1) < 1e-6
    assert st.restricted == 1


# forces and eigenvalues of a solve stopped at delta_rho < 1e-6 are first order in its residual (the reference's own
# Aufbau and FD H2 forces differ by 6.3e-7, qm/src/scf.rs:1277, 1291), energies second order: f_tol 2e-6 for FP64
@pytest.mark.parametrize("precision,e_tol,f_tol", [(0, 1e-9, 2e-6), (1, 3e-6, 2e-5)])
def test_scf_lih_matches_reference_pins(lira_mod, precision, e_tol, f_tol):
    # qm/src/scf.rs:1278-1279: E = -5.219911350989, F_x(Li) = -0.022680156352 (13 CPU steps)
    lira = lira_mod
    s, st, f = _scf(lira, precision, LIH)
    print(f"LiH precision {precision}: E {st.energy.total:.12f} F_x {f[0, 0]:+.12f} steps {st.iterations} "
          f"{st.seconds_total * 1e3:.1f} ms")
    assert abs(st.energy.total - (-5.219911350989)) < e_tol
    assert abs(f[0, 0] - (-0.022680156352)) < f_tol


@pytest.mark.parametrize("system", [H2, LIH], ids=["H2", "LiH"])
def test_scf_fp64_tight_state_matches_oracle(lira_mod, system):
    # both solvers driven to delta_rho < 1e-9 with eig_tol 1e-9: the same discrete fixed point, so energies,
    # eigenvalues, densities and forces agree to the tolerances, not to the first-order 1e-6 floor
    lira = lira_mod
    pos, z, nu, nd = system
    n, h = 33, 0.25
    s = lira.GpuSolver(n, h, 0)
    st = s.solve(pos, z, nu, nd, lira.default_settings(h, 0, tol=1e-9, eig_tol=1e-9))
    f = s.forces(pos, z, h)
    so, out = O.scf(n, h, pos, z, nu, nd, settings=O.default_settings(h, "aufbau", tol=1e-9, eig_tol=1e-9))
    assert so.status == O.OK
    state = s.state()
    print(f"tight {nu + nd} e-: dE {st.energy.total - so.total:.2e} dF {np.abs(f - out['forces']).max():.2e} "
          f"deps {np.abs(state.evals_up - np.sort(out['evals_up'])).max():.2e}")
    assert abs(st.energy.total - so.total) < 1e-10
    assert np.abs(state.evals_up - np.sort(out["evals_up"])).max() < 1e-8
    assert np.abs(f - out["forces"]).max() < 1e-8
    rho = state.rho_up + state.rho_dn
    rho_o = out["rho_up"] + out["rho_dn"]
    assert np.abs(rho - rho_o).max() < 1e-7 * rho_o.max()
    for name in ("kinetic", "external", "hartree", "xc", "nuclear"):
        assert abs(getattr(st.energy, name) - getattr(so, name)) < 1e-8, name


def test_scf_fermi_dirac_h2(lira_mod):
    # FD_H2_E = -0.952449176839 (qm/src/scf.rs:1290): gapped H2, TS ~ 0; per spin the occupied level, a tested witness
    # and (as the reference's GPU solver, crates/qm-gpu/src/scf.rs:874-900) one untested buffer level with f = 0
    lira = lira_mod
    s, st, f = _scf(lira, 0, H2, occ=1)
    assert abs(st.energy.total - (-0.952449176839)) < 1e-9
    assert st.n_stored[0] in (2, 3) and st.energy.ts < 1e-12
    ev, occ = s.levels(0)
    assert abs(occ.sum() - 1.0) < 1e-12 and occ[-1] < 1e-30


def test_scf_coincident_nuclei_keeps_state(lira_mod):
    lira = lira_mod
    s, st, f = _scf(lira, 1, H2)
    with pytest.raises(lira.CoincidentNuclei):
        s.solve([[0, 0, 0], [0, 0, 0]], [1, 1], 1, 1)
    # the converged state survives (the reference keeps its warm state on CoincidentNuclei)
    ev, _ = s.levels(0)
    assert ev.size == 1


def test_scf_warm_start_is_cheaper_and_lands_on_the_cold_energy(lira_mod):
    lira = lira_mod
    n, h = 33, 0.25
    s = lira.GpuSolver(n, h, 1)
    st = lira.default_settings(h, 0)
    s.solve([[0.7, 0, 0], [-0.7, 0, 0]], [1, 1], 1, 1, st)
    warm = s.solve([[0.72, 0, 0], [-0.7, 0, 0]], [1, 1], 1, 1, st, warm=True)
    cold = lira.GpuSolver(n, h, 1).solve([[0.72, 0, 0], [-0.7, 0, 0]], [1, 1], 1, 1, st)
    # the warm start skips the cold density's first steps; the f32 tail below 1e-5 takes its steps either way (its
    # last steps keep the orbitals within 1.5 x the eigen tolerance while the mixing converges the density)
    assert warm.iterations <= cold.iterations
    assert warm.seconds_total < cold.seconds_total
    assert abs(warm.energy.total - cold.energy.total) < 1e-5


def test_scf_abort_returns_aborted(lira_mod):
    lira = lira_mod
    s = lira.GpuSolver(25, 0.3, 1)
    calls = []

    def abort():
        calls.append(1)
        return len(calls) > 2

    with pytest.raises(lira.Aborted):
        s.solve(H2[0], H2[1], 1, 1, lira.default_settings(0.3, 0), abort=abort)


# ----------------------------------------------------------------- provider-facing controls and products ----

def test_profile_density_sampling_and_witness_cap(lira_mod):
    lira = lira_mod
    n, h = 33, 0.25
    s = lira.GpuSolver(n, h, 0)
    st = s.solve(H2[0], H2[1], 1, 1, lira.default_settings(h, 1))
    p, dr, et = s.profile()
    # SolveProfile (crates/qm-gpu/src/scf.rs:439-514): one delta_rho and eigen tolerance per step
    assert p.steps == st.iterations == len(dr) == len(et)
    assert dr[-1] < 1e-6 and np.all(et > 0)
    assert p.eigen_iterations == st.lobpcg_iterations and p.readbacks == st.readbacks
    assert p.peak_block >= 2 and p.wall >= p.eigen_wall > 0 and p.extrapolated == 0
    # sample_density (crates/qm-gpu/src/sample.rs:84) = qm::Grid::trilinear of the state's density: random points,
    # nodes (read bitwise) and points outside the grid (0)
    rho = s.field(2)
    rng = np.random.default_rng(21)
    x, y, z = O.grid_points(n, h)
    nodes = np.stack([x[[0, 5000, 17000, n ** 3 - 1]], y[[0, 5000, 17000, n ** 3 - 1]],
                      z[[0, 5000, 17000, n ** 3 - 1]]], axis=1)
    pts = np.concatenate([rng.uniform(-4.0, 4.0, (300, 3)), nodes, [[4.1, 0, 0], [0, -4.01, 0]]])
    got = s.sample_density(pts)
    want = O.trilinear(n, h, rho, pts)
    assert np.abs(got - want).max() <= 1e-14 * rho.max()
    assert np.array_equal(got[300:304], rho[[0, 5000, 17000, n ** 3 - 1]]) and np.all(got[304:] == 0.0)
    # GpuSolver::set_witness_cap clamps to 1..32 (scf.rs:1126)
    s.set_witness_cap(0)
    assert s.witness_cap() == 1
    s.set_witness_cap(99)
    assert s.witness_cap() == 32


def test_density_extrapolation_along_a_trajectory(lira_mod):
    # GpuSolver::set_density_extrapolation (crates/qm-gpu/src/scf.rs:1146-1162): the third and later warm solves of a
    # trajectory start from 2 rho_t - rho_{t-1}; they land on the cold answer; off drops it
    lira = lira_mod
    n, h = 33, 0.25
    st = lira.default_settings(h, 0)
    s = lira.GpuSolver(n, h, 1)
    s.set_density_extrapolation(True)
    used, steps = [], []
    for i, xa in enumerate([0.70, 0.71, 0.72, 0.73]):
        r = s.solve([[xa, 0.0, 0.0], [-0.7, 0.0, 0.0]], [1, 1], 1, 1, st, warm=i > 0)
        used.append(s.profile()[0].extrapolated)
        steps.append(r.iterations)
    print(f"extrapolated trajectory: steps {steps}")
    assert used == [0, 0, 1, 1]
    cold = lira.GpuSolver(n, h, 1).solve([[0.73, 0.0, 0.0], [-0.7, 0.0, 0.0]], [1, 1], 1, 1, st)
    assert abs(r.energy.total - cold.energy.total) < 1e-5
    s.set_density_extrapolation(False)
    s.solve([[0.74, 0.0, 0.0], [-0.7, 0.0, 0.0]], [1, 1], 1, 1, st, warm=True)
    assert s.profile()[0].extrapolated == 0


def test_forces_are_minus_the_energy_gradient(lira_mod):
    # dE/dR of the converged FP64 SCF by a Richardson central difference (forces.rs:151-166) against the GPU's
    # Hellmann-Feynman force: H2 at 33^3 x 0.25, tight tolerances, nucleus 0 moved along x
    lira = lira_mod
    n, h = 33, 0.25
    st = lira.default_settings(h, 0, tol=1e-9, eig_tol=1e-9)

    def run(dx):
        s = lira.GpuSolver(n, h, 0)
        r = s.solve([[0.7 + dx, 0.0, 0.0], [-0.7, 0.0, 0.0]], [1, 1], 1, 1, st)
        return r.energy.total, s.forces([[0.7 + dx, 0.0, 0.0], [-0.7, 0.0, 0.0]], [1, 1], h)

    _, f0 = run(0.0)
    d = 0.01
    fd1 = -(run(d)[0] - run(-d)[0]) / (2 * d)
    fd2 = -(run(2 * d)[0] - run(-2 * d)[0]) / (4 * d)
    rich = (4 * fd1 - fd2) / 3
    print(f"HF {f0[0, 0]:+.8f} vs Richardson {rich:+.8f}")
    assert abs(f0[0, 0] - rich) < 1e-4


def test_config1_h2_48_fermi_dirac_fp64_matches_oracle(lira_mod):
    # the small plain-method case: H2 at z = +-0.37 A (0.6992 Bohr) on 48^3 x 0.25 Bohr, LSDA singlet, Fermi-Dirac
    # kT 5 mHa, FP64; both solvers driven to delta_rho < 1e-9: energy, occupied level and forces against the oracle
    lira = lira_mod
    n, h = 48, 0.25
    pos, z = [[0.0, 0.0, 0.6992], [0.0, 0.0, -0.6992]], [1, 1]
    s = lira.GpuSolver(n, h, 0)
    st = s.solve(pos, z, 1, 1, lira.default_settings(h, 1, tol=1e-9, eig_tol=1e-9))
    f = s.forces(pos, z, h)
    so, out = O.scf(n, h, pos, z, 1, 1, settings=O.default_settings(h, "fd", tol=1e-9, eig_tol=1e-9))
    assert so.status == O.OK
    ev, occ = s.levels(0)
    print(f"H2 48^3 FD: GPU E {st.energy.total:.12f} ({st.iterations} steps, {st.n_stored[0]} stored) oracle "
          f"{so.total:.12f} ({so.iterations} steps)")
    assert abs(st.energy.total - so.total) < 1e-10
    assert abs(ev[0] - np.sort(out["evals_up"])[0]) < 1e-8
    assert np.abs(f - out["forces"]).max() < 1e-8
    assert abs(occ.sum() - 1.0) < 1e-12


def test_state_summary_spin_integral_and_density_max(lira_mod):
    # GpuSolver::spin_abs_integral / density_max (crates/qm-gpu/src/scf.rs:1834-1845): a singlet's channels are the
    # same field (0 exactly); a lone H electron's spin integral is its electron count; the maximum is the field's
    lira = lira_mod
    n, h = 33, 0.25
    for prec in (0, 1):
        s = lira.GpuSolver(n, h, prec)
        s.solve(H2[0], H2[1], 1, 1, lira.default_settings(h, 1))
        a, m = s.state_summary()
        assert a == 0.0 and m == s.field(2).max()
        s = lira.GpuSolver(n, h, prec)
        st = s.solve([[0.1, 0.0, -0.2]], [1], 1, 0, lira.default_settings(h, 1))
        a, m = s.state_summary()
        rho = s.field(2)
        assert abs(a - st.electrons[0]) < 1e-6 * (1 if prec == 0 else 10) and m == rho.max()
        assert abs(a - np.abs(s.field(0) - s.field(1)).sum() * h ** 3) < 1e-9


def test_provider_evaluation_policies(lira_mod):
    # GpuKs::evaluate_once's drive of the solver (crates/providers/src/ks.rs:1573-1642): forces of the converged
    # state; R15 (a failed warm attempt retried once cold, its steps counted); the extrapolated-start flag
    lira = lira_mod
    n, h = 33, 0.25
    s = lira.GpuSolver(n, h, 1)
    s.set_density_extrapolation(True)
    st = lira.default_settings(h, 1)
    traj = [0.70, 0.71, 0.72]
    for i, x in enumerate(traj):
        pos = [[x, 0.0, 0.0], [-0.7, 0.0, 0.0]]
        code, stats, f, flags, tol = s.evaluate(pos, [1, 1], 1, 1, st, warm=i > 0)
        assert code == lira.OK and tol == st.tol and not flags & lira.EVAL_TOL_RAISED  # all-electron: untouched
        assert np.array_equal(f, s.forces(pos, [1, 1], h))
        assert bool(flags & lira.EVAL_EXTRAPOLATED) == (i == 2)
    # a warm attempt that runs out of steps is retried cold; with the same cap the cold one fails too, and its
    # status is the result; both attempts' steps are counted
    one = lira.default_settings(h, 1, max_iter=1)
    code, stats, _, flags, _ = s.evaluate([[0.74, 0.0, 0.0], [-0.7, 0.0, 0.0]], [1, 1], 1, 1, one, warm=True)
    assert code == lira.NOT_CONVERGED and flags & lira.EVAL_RETRIED_COLD and not flags & lira.EVAL_WARM_STALLED
    assert stats.iterations == 2


def test_provider_evaluation_raises_the_tolerance_to_the_gpu_floor(lira_mod):
    # under an ionic setup tol is raised to max(tol, 2 * 5e-8 * N_valence) (crates/providers/src/pseudo.rs:437-440):
    # pseudised water has 8 valence electrons, so 1e-9 becomes 8e-7
    import synthetic_tables as T
    lira = lira_mod
    n, h = 25, 0.3
    s = lira.GpuSolver(n, h, 1)
    s.set_ionic([T.oxygen_like(h)])
    pos, z = [[0.05, -0.03, 0.02], [1.43, 1.11, 0.0], [-1.43, 1.11, 0.0]], [8, 1, 1]
    code, stats, f, flags, tol = s.evaluate(pos, z, 4, 4, lira.default_settings(h, 1, tol=1e-9))
    assert code == lira.OK and flags & lira.EVAL_TOL_RAISED and tol == pytest.approx(8e-7)
    assert stats.delta_rho < 8e-7 and np.array_equal(f, s.forces(pos, z, h))
