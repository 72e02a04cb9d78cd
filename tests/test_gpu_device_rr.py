"""The LOBPCG Rayleigh-Ritz step on the device (one f64 workgroup, Jacobi eigen-decompositions in LDS) against the
host step (crates/qm-gpu/src/eigen.rs:515-542; the host step is pinned by the reference's rr_i1 fixture and dense
generalised solves in test_abi.py), and the solver with the device step against the solver with the host step."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def lira_mod():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    import lira
    lira.lib()
    return lira


def _pencil(m, seed, cond=1e3):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((m, m)))
    b = q @ np.diag(np.logspace(0.0, -np.log10(cond), m)) @ q.T
    d = np.exp(rng.uniform(-2.0, 2.0, m))  # unequal column norms: the D-scaling matters
    b = d[:, None] * b * d[None, :]
    a = rng.standard_normal((m, m))
    return a + a.T, b


def _check_against_host(lira, a, b, nk, rtol_theta=1e-10):
    th_h, c_h = lira.rayleigh_ritz(a, b, nk)
    th_d, c_d, cp_d = lira.rayleigh_ritz_device(a, b, nk)
    scale = max(np.abs(th_h).max(), 1.0)
    assert np.abs(th_d - th_h).max() <= rtol_theta * scale
    # B-orthonormal Ritz vectors with Rayleigh quotients equal to the Ritz values
    assert np.abs(c_d.T @ b @ c_d - np.eye(nk)).max() < 1e-9
    for j in range(nk):
        cj = c_d[:, j]
        assert abs(cj @ a @ cj - th_d[j]) < 1e-9 * scale
    # the same vectors as the host step up to sign (the pencils here have no repeated Ritz value)
    sgn = np.sign(np.sum(c_d * c_h, axis=0))
    assert np.abs(c_d - c_h * sgn).max() < 1e-7 * np.abs(c_h).max()
    # P's coefficients: C minus its B-projection on the first nk columns
    b11, b1 = b[:nk, :nk], b[:nk, :]
    want = c_d.copy()
    want[:nk] -= np.linalg.solve(b11, b1 @ c_d)
    assert np.abs(cp_d - want).max() < 1e-8 * max(np.abs(want).max(), 1.0)


@pytest.mark.parametrize("m,nk", [(27, 9), (36, 12), (12, 4), (5, 2), (1, 1), (33, 11)])
def test_device_step_matches_host_step(lira_mod, m, nk):
    a, b = _pencil(m, 10 + m)
    _check_against_host(lira_mod, a, b, nk)


def test_device_step_on_the_reference_fixture(lira_mod):
    # crates/qm-gpu/src/eigen.rs:1182-1201: the measured 24x24 (A, B) of a failing step; theta0 = -0.8686699
    a = np.loadtxt(os.path.join(GOLD, "rr_i1_a.txt"))
    b = np.loadtxt(os.path.join(GOLD, "rr_i1_b.txt"))
    th, c, _ = lira_mod.rayleigh_ritz_device(a, b, 8)
    th_h, _ = lira_mod.rayleigh_ritz(a, b, 8)
    assert abs(th[0] + 0.8686699) < 1e-6
    assert np.all(np.diff(th) >= 0)
    assert np.abs(th - th_h).max() < 1e-9
    for k in range(8):
        ck = c[:, k]
        assert abs(ck @ a @ ck / (ck @ b @ ck) - th[k]) < 1e-9


def test_device_step_drops_dependent_directions(lira_mod):
    # the last column repeats the second: B~ has a zero direction, dropped at 1e-6 of the largest; with 2 of 3 needed
    # it still works, with 3 it is a rank loss (LIRA_NUMERICAL), as on the host
    a = np.diag([1.0, 2.0, 3.0])
    b = np.array([[1.0, 0.0, 0.0], [0.0, 1.0, 1.0], [0.0, 1.0, 1.0]])
    th, c, _ = lira_mod.rayleigh_ritz_device(a, b, 2)
    th_h, _ = lira_mod.rayleigh_ritz(a, b, 2)
    assert np.abs(th - th_h).max() < 1e-12
    with pytest.raises(Exception):
        lira_mod.rayleigh_ritz_device(a, b, 3)
    with pytest.raises(Exception):
        lira_mod.rayleigh_ritz_device(np.eye(37), np.eye(37), 4)  # beyond the device step's 36 columns


@pytest.mark.parametrize("precision", [0, 1])
def test_scf_with_the_device_step_equals_the_host_step(lira_mod, precision):
    """Water-shaped all-electron SCF (Fermi-Dirac, 41^3 x 0.3, delta_rho < 1e-6): the device step and the host step
    converge to the same state: FP64 energies within 1e-9 Ha (second order in the density), levels within 2e-6 Ha and
    forces within 1e-5 Ha/Bohr (first order, at the SCF tolerance); mixed within the f32 path's own noise."""
    import math
    lira = lira_mod
    n, h = 41, 0.3
    r = 0.96 * 1.8897261
    half = math.radians(104.5 / 2)
    pos = [[0.0, 0.0, 0.0], [r * math.sin(half), r * math.cos(half), 0.0], [-r * math.sin(half), r * math.cos(half), 0.0]]
    z = [8, 1, 1]
    settings = lira.default_settings(h, lira.FERMI_DIRAC)
    out = {}
    for dev in (True, False):
        s = lira.GpuSolver(n, h, precision)
        s.set_device_rr(dev)
        st = s.solve(pos, z, 5, 5, settings)
        out[dev] = (st.energy.total, s.levels(0), st.iterations, st.lobpcg_iterations, s.forces(pos, z, h))
        print(f"precision {precision} device {dev}: E {st.energy.total:.10f}, {st.iterations} steps, "
              f"{st.lobpcg_iterations} LOBPCG iterations, {st.readbacks} readbacks, {st.seconds_total:.3f} s")
    de = abs(out[True][0] - out[False][0])
    (ev_d, oc_d), (ev_h, oc_h) = out[True][1], out[False][1]
    held = [i for i in range(min(len(ev_d), len(ev_h))) if max(oc_d[i], oc_h[i]) >= 1e-6]
    deps = np.abs(ev_d[held] - ev_h[held]).max()
    df = np.abs(out[True][4] - out[False][4]).max()
    print(f"dE {de:.2e}, max d eps {deps:.2e}, max dF {df:.2e}")
    if precision == 0:
        assert de < 1e-9 and deps < 2e-6 and df < 1e-5
    else:
        assert de < 2e-6 and deps < 2e-5 and df < 1e-4
