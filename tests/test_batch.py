"""CPU checks of the batched SCF's host side: the molecule sampler's geometries and lira_solve_batch's argument
checks, which return before any device call."""
import ctypes as C
import math

import numpy as np
import pytest

import lira
import molecules as M


def _dist(p, i, j):
    return float(np.linalg.norm(p[i] - p[j]))


def _angle(p, c, i, j):
    a, b = p[i] - p[c], p[j] - p[c]
    return math.degrees(math.acos(float(a @ b) / (np.linalg.norm(a) * np.linalg.norm(b))))


def test_equilibrium_geometries_have_the_stated_bonds_and_angles():
    A = M.ANGSTROM
    p = M.equilibrium("H2")
    assert _dist(p, 0, 1) == pytest.approx(0.741 * A, abs=1e-12)
    assert _dist(M.equilibrium("LiH"), 0, 1) == pytest.approx(1.595 * A, abs=1e-12)
    p = M.equilibrium("H2O")
    assert _dist(p, 0, 1) == pytest.approx(0.9572 * A, abs=1e-12) and _dist(p, 0, 2) == pytest.approx(0.9572 * A)
    assert _angle(p, 0, 1, 2) == pytest.approx(104.52, abs=1e-9)
    p = M.equilibrium("NH3")
    for i in (1, 2, 3):
        assert _dist(p, 0, i) == pytest.approx(1.012 * A, abs=1e-12)
    for i, j in ((1, 2), (1, 3), (2, 3)):
        assert _angle(p, 0, i, j) == pytest.approx(106.67, abs=1e-9)
    p = M.equilibrium("CH4")
    for i in range(1, 5):
        assert _dist(p, 0, i) == pytest.approx(1.087 * A, abs=1e-12)
        for j in range(i + 1, 5):
            assert _angle(p, 0, i, j) == pytest.approx(math.degrees(math.acos(-1 / 3)), abs=1e-9)
    for name in M.SPECIES:
        assert np.abs(M.equilibrium(name).mean(axis=0)).max() < 1e-12, name


def test_sampler_is_seeded_bounded_and_covers_the_species():
    a, b = M.sample(64, seed=3), M.sample(64, seed=3)
    assert [m.name for m in a] == [m.name for m in b]
    assert all(np.array_equal(x.pos, y.pos) and x.n == y.n for x, y in zip(a, b))
    assert {m.name for m in a} == set(M.SPECIES)
    assert {m.n for m in a} <= set(range(48, 97, 8)) and len({m.n for m in a}) > 3
    for m in a:
        z, nu, nd = M.SPECIES[m.name]
        assert m.z == z and (m.n_up, m.n_dn) == (nu, nd) and m.h == 0.25
        d = m.pos - M.equilibrium(m.name)
        assert np.abs(d).max() <= 0.1 * M.ANGSTROM
        # every nucleus at least 3.5 Bohr inside its box
        assert np.abs(m.pos).max() < (m.n - 1) * m.h / 2 - 3.5


def test_batch_rejects_bad_arguments_before_any_device_call():
    L = lira.lib()
    nuc = lira.nuclei_array([[0.0, 0.0, 0.0]], [1])
    job = lira.BatchJob(33, 0.3, lira.System(C.cast(nuc, C.POINTER(lira.Nucleus)), 1, 1, 0))
    res = (lira.BatchResult * 1)()
    jobs = (lira.BatchJob * 1)(job)
    assert L.lira_solve_batch(jobs, 0, lira.MIXED, None, 2, res, None) == lira.OK  # nothing to do
    assert L.lira_solve_batch(jobs, -1, lira.MIXED, None, 2, res, None) == lira.INVALID
    assert L.lira_solve_batch(jobs, 1, lira.MIXED, None, 0, res, None) == lira.INVALID
    assert L.lira_solve_batch(jobs, 1, lira.MIXED, None, lira.BATCH_MAX_WORKERS + 1, res, None) == lira.INVALID
    assert L.lira_solve_batch(jobs, 1, 5, None, 1, res, None) == lira.INVALID
    assert L.lira_solve_batch(None, 1, lira.MIXED, None, 1, res, None) == lira.INVALID
    assert L.lira_solve_batch(jobs, 1, lira.MIXED, None, 1, None, None) == lira.INVALID
    for bad in (lira.BatchJob(2, 0.3, job.system), lira.BatchJob(lira.MAX_GRID + 1, 0.3, job.system),
                lira.BatchJob(33, 0.0, job.system), lira.BatchJob(33, 0.3, lira.System(None, 1, 1, 0)),
                lira.BatchJob(33, 0.3, lira.System(C.cast(nuc, C.POINTER(lira.Nucleus)), 0, 1, 0)),
                lira.BatchJob(33, 0.3, lira.System(C.cast(nuc, C.POINTER(lira.Nucleus)), 1, -1, 0))):
        assert L.lira_solve_batch((lira.BatchJob * 1)(bad), 1, lira.MIXED, None, 1, res, None) == lira.INVALID
