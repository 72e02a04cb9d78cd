"""CPU checks of the C ABI: liblira loads, exports every symbol include/lira.h declares, and its host-only numerics
(the Rayleigh–Ritz step) are pinned by the reference's fixture. No GPU compute runs here."""
import os
import re

import numpy as np

import lira

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_library_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "lira.h")).read()
    declared = sorted(set(re.findall(r"\b(lira_[a-z0-9_]+)\s*\(", header)) - {"lira_abort_fn"})
    L = lira.lib()
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, missing
    assert set(declared) == set(lira.EXPORTS)
    assert L.lira_version().startswith(b"liblira")


def test_rayleigh_ritz_on_the_reference_fixture():
    # crates/qm-gpu/src/eigen.rs:1182-1201: the measured 24x24 (A, B) of a failing step; theta0 = -0.8686699 and every
    # returned column's Rayleigh quotient equals its Ritz value to 1e-9
    a = np.loadtxt(os.path.join(GOLD, "rr_i1_a.txt"))
    b = np.loadtxt(os.path.join(GOLD, "rr_i1_b.txt"))
    th, c = lira.rayleigh_ritz(a, b, 8)
    for k in range(8):
        ck = c[:, k]
        assert abs(ck @ a @ ck / (ck @ b @ ck) - th[k]) < 1e-9
    assert abs(th[0] + 0.8686699) < 1e-6
    assert np.all(np.diff(th) >= 0)


def test_rayleigh_ritz_matches_a_dense_generalised_solve():
    # an SPD B and symmetric A: the lowest Ritz values are scipy's generalised eigenvalues
    import scipy.linalg as sl
    rng = np.random.default_rng(5)
    m = 12
    q = rng.standard_normal((m, m))
    b = q @ q.T + m * np.eye(m)
    a = rng.standard_normal((m, m))
    a = a + a.T
    th, c = lira.rayleigh_ritz(a, b, 4)
    assert np.allclose(th, sl.eigh(a, b, eigvals_only=True)[:4], atol=1e-10)
    assert np.allclose(c.T @ b @ c, np.eye(4), atol=1e-10)  # B-orthonormal coefficients


def test_rank_loss_is_a_numerical_error():
    # two identical columns: B~ has a zero direction, so only 2 of 3 survive and 3 Ritz pairs cannot be formed
    a = np.diag([1.0, 2.0, 3.0])
    b = np.array([[1.0, 1.0, 0.0], [1.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
    try:
        lira.rayleigh_ritz(a, b, 3)
    except lira.Numerical:
        return
    raise AssertionError("expected lira.Numerical")


def test_bad_arguments_are_invalid_not_crashes():
    import ctypes as C
    h = C.c_void_p()
    assert lira.lib().lira_solver_create(1, 0.25, 0, C.byref(h)) == lira.INVALID
    assert lira.lib().lira_solver_create(33, -1.0, 0, C.byref(h)) == lira.INVALID
    assert lira.lib().lira_solver_create(33, 0.25, 7, C.byref(h)) == lira.INVALID
