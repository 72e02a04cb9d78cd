"""GPU tests of the display products of a converged state (crates/qm-gpu/src/orbitals.rs, harris.rs): the unoccupied
levels of the state's own potential, and the Harris display solve in the potential of a given density.

Pins: the unoccupied levels against the oracle's f64 LOBPCG on the same output potential (the oracle's own H with the
GPU's converged density), the occupied levels of the display solve against the SCF's (the reference's runtime guard,
tau = max(100 tol, 1e-4) Ha), and the Harris map's fixed point: on the converged SCF density it reproduces the SCF's
levels and density (crates/qm-gpu/src/harris.rs, harris_on_the_converged_density_reproduces_the_scf).
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

H2 = ([[0.7, 0.0, 0.0], [-0.7, 0.0, 0.0]], [1, 1])
LIH = ([[-1.5, 0.0, 0.0], [1.5, 0.0, 0.0]], [3, 1])


@pytest.fixture(scope="module")
def lira_mod():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    import lira
    lira.lib()
    return lira


@pytest.mark.parametrize("occupations", [0, 1])
def test_unoccupied_levels_match_the_oracle(lira_mod, occupations):
    """H2 on 33^3 x 0.25 in FP64: the 4 lowest unoccupied levels (and 2 guards) of the converged state's own potential
    against the oracle's LOBPCG on the same potential (V_nuc + V_H[rho_out] + v_xc[rho_out], rebuilt on the host from
    the state's fields), to the eigen tolerance."""
    lira = lira_mod
    n, h = 33, 0.25
    st = lira.default_settings(h, occupations, tol=1e-8, eig_tol=1e-8)
    s = lira.GpuSolver(n, h, lira.FP64)
    s.solve(H2[0], H2[1], 1, 1, st)
    ev, res, occ_err, orbs = s.unoccupied(4, want_orbitals=True)
    print(f"unoccupied (occupations {occupations}): {np.round(ev[0], 6).tolist()}, residuals "
          f"{np.max(res[0]):.1e}, occupied error {occ_err}")
    assert len(ev[0]) == 6 and np.all(np.diff(ev[0]) >= 0) and np.array_equal(ev[0], ev[1])
    assert max(occ_err) < 1e-6
    # the state's output potential on the host: V_nuc + V_H + v_xc of the output densities (f64 oracle LSDA)
    ru, rd = s.field(0), s.field(1)
    _, vxu, _ = O.lsda(ru, rd)
    v = s.field(4) + s.field(3) + vxu
    # the oracle's lowest 1 + 6 levels from the GPU's occupied orbital and unoccupied vectors as the guess
    guess = np.concatenate([s.orbitals(0)[:1], orbs[0]])
    _, evo, it, mr, _ = O.lobpcg(n, h, v, guess, 1e-9, 400)
    evo = np.sort(evo)
    print(f"oracle: {np.round(evo, 6).tolist()} ({it} iterations)")
    # the displayed m levels (not the top guard, which converges slowest) agree to the eigen tolerance
    assert np.abs(ev[0][:4] - evo[1:5]).max() < 1e-6
    assert abs(evo[0] - s.levels(0)[0][0]) < 1e-7


def test_unoccupied_mixed_precision_and_warm_guess(lira_mod):
    """MIXED: the display's occupied levels match the SCF's within tau; a second call starts from the kept vectors
    and needs fewer iterations; the levels agree with the first call's to the f32 tolerance."""
    lira = lira_mod
    n, h = 33, 0.25
    s = lira.GpuSolver(n, h, lira.MIXED)
    s.solve(LIH[0], LIH[1], 2, 2, lira.default_settings(h, 1))
    ev1, res1, err1, _ = s.unoccupied(3)
    ev2, res2, err2, _ = s.unoccupied(3)
    assert max(err1) < 1e-4 and max(err2) < 1e-4
    assert np.abs(ev1[0][:3] - ev2[0][:3]).max() < 5e-5


def test_harris_on_the_converged_density_reproduces_the_scf(lira_mod):
    """The converged density is a fixed point of the Harris map: the display solve on rho_SCF gives the SCF's occupied
    levels and density (to the eigen tolerance), its output integrates to the electron counts, and it leaves the SCF
    state as it was."""
    lira = lira_mod
    n, h = 33, 0.25
    st = lira.default_settings(h, lira.AUFBAU, tol=1e-8, eig_tol=1e-8)
    s = lira.GpuSolver(n, h, lira.FP64)
    r0 = s.solve(LIH[0], LIH[1], 2, 2, st)
    ru, rd = s.field(0), s.field(1)
    ev_scf = s.levels(0)[0]
    f_before = s.forces(LIH[0], LIH[1], h)
    res, ou, od = s.harris(LIH[0], LIH[1], 2, 2, ru, rd, m=2, settings=st)
    print(f"Harris: levels {np.round(res.evals[0][:res.n_levels[0]], 6).tolist()}, n_occ {list(res.n_occ)}, "
          f"{list(res.iterations)} iterations, {res.seconds * 1e3:.1f} ms")
    assert list(res.n_occ) == [2, 2] and res.n_levels[0] == 2 + 2 + 2
    assert np.abs(np.array(res.evals[0][:2]) - np.sort(ev_scf)).max() < 1e-6
    assert abs(res.electrons[0] - 2.0) < 1e-8 and abs(res.electrons[1] - 2.0) < 1e-8
    assert np.abs(ou - ru).max() < 1e-5 * ru.max()
    # the SCF state is untouched: same fields and forces as before
    assert np.array_equal(s.field(0), ru) and np.array_equal(s.forces(LIH[0], LIH[1], h), f_before)
    assert s.state_summary()[1] == pytest.approx(float((ru + rd).max()), rel=1e-12)
    del r0


def test_harris_fermi_dirac_on_a_blob_density(lira_mod):
    """A Harris solve from a non-converged input (the SCF's sigma = 1 blob guess of H2, 1 + 1 electrons, Fermi-Dirac):
    occupations sum to the electron counts, the output density integrates to them, and a second call warm-starts."""
    lira = lira_mod
    n, h = 33, 0.3
    pos, z = H2
    x, y, zz = O.grid_points(n, h)
    blob = sum(np.exp(-((x - p[0]) ** 2 + (y - p[1]) ** 2 + (zz - p[2]) ** 2) / 2.0) / (2 * np.pi) ** 1.5
               for p in pos)
    s = lira.GpuSolver(n, h, lira.MIXED)
    st = lira.default_settings(h, lira.FERMI_DIRAC)
    r1, ou, od = s.harris(pos, z, 1, 1, 0.5 * blob, 0.5 * blob, m=1, settings=st)
    assert r1.cold == 1 and r1.n_occ[0] >= 1
    assert abs(sum(r1.occ[0][:r1.n_levels[0]]) - 1.0) < 1e-9
    assert abs(r1.electrons[0] - 1.0) < 1e-5 and abs(ou.sum() * h ** 3 - 1.0) < 1e-5
    r2, _, _ = s.harris(pos, z, 1, 1, 0.5 * blob, 0.5 * blob, m=1, settings=st)
    assert r2.cold == 0 and sum(r2.iterations) < sum(r1.iterations)
    assert abs(r2.evals[0][0] - r1.evals[0][0]) < 1e-4
