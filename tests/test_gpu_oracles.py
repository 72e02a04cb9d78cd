"""Physics oracles of the FP64 GPU solve (crates/validate/tests/oracles.rs): each quantity is computed on a ladder of
four nuclear smearing widths sigma = 1, 1.2, 1.4, 1.6 h, warm-started from one sigma to the next, and extrapolated to
the point nucleus with a + b s^2 + c s^3 + d s^4 (crates/qm/src/scf.rs:1122-1127). The answers are the NIST LSD
atomic energies and the LSDA H2 equilibrium (CCCBDB), with the reference's own bounds: they pin the whole solver
(grid, nuclear potential, Poisson, LSDA, eigensolver, SCF, forces) against physics rather than against the oracle.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lira_mod():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    import lira
    lira.lib()
    return lira


def _ladder(lira, n, h, pos, z, n_up, n_dn, tol, eig_tol, observe):
    s = lira.GpuSolver(n, h, lira.FP64)
    sig = lira.sigma_ladder(h)
    out = []
    for k, sg in enumerate(sig):
        st = lira.default_settings(h, lira.FERMI_DIRAC, sigma=sg, tol=tol, eig_tol=eig_tol)
        r = s.solve(pos, z, n_up, n_dn, st, warm=k > 0)
        out.append(observe(s, r, sg))
    return sig, out


def _atom(lira, z, n, h, n_up, n_dn):
    def observe(s, r, sg):
        assert r.energy.ts <= 1e-12, f"sigma {sg}: TS = {r.energy.ts:e} (a gapped atom carries no entropy)"
        return r.energy.total
    sig, e = _ladder(lira, n, h, [[0.0, 0.0, 0.0]], [z], n_up, n_dn, 1e-7, 1e-7, observe)
    a = lira.extrapolate_sigma(sig, e)
    print(f"Z = {z}: {n}^3, h = {h}: E(sigma) = {np.round(e, 6).tolist()} -> {a:+.6f} Ha")
    return a


def test_hydrogen_atom_lsd_energy(lira_mod):
    # NIST LSD H: -0.478671 Ha; 81^3 at h = 0.15 (crates/validate/tests/oracles.rs:48-56, bound 4e-3)
    a = _atom(lira_mod, 1, 81, 0.15, 1, 0)
    assert abs(a - (-0.478671)) < 4e-3


def test_helium_atom_lsd_energy(lira_mod):
    # NIST LSD He: -2.834836 Ha; 101^3 at h = 0.10 (oracles.rs:58-68, bound 1.5e-2: the smearing series is
    # pre-asymptotic at Z sigma_max = 0.32, the reference's own fit misses by 12.3 mHa)
    a = _atom(lira_mod, 2, 101, 0.10, 1, 1)
    assert abs(a - (-2.834836)) < 1.5e-2


def test_h2_force_vanishes_at_the_lsda_bond_length(lira_mod):
    # F_x on nucleus 0 at R = 1.446 Bohr (the LSDA R_e, NIST CCCBDB), tol 1e-9, eig_tol 1e-8, extrapolated to
    # sigma -> 0 (oracles.rs:173-190, bound 1e-2 Ha/Bohr)
    lira = lira_mod
    h, r = 0.15, 1.446
    pos = [[r / 2, 0.0, 0.0], [-r / 2, 0.0, 0.0]]

    def observe(s, res, sg):
        assert res.energy.ts <= 1e-12
        return float(s.forces(pos, [1, 1], sg)[0, 0])
    sig, fx = _ladder(lira, 81, h, pos, [1, 1], 1, 1, 1e-9, 1e-8, observe)
    f0 = lira.extrapolate_sigma(sig, fx)
    print(f"H2 at R = {r}: F_x0(sigma) = {np.round(fx, 6).tolist()} -> {f0:+.6f} Ha/Bohr")
    assert abs(f0) < 1e-2
