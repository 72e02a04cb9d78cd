"""HIP graph replay (lira_set_graphs) against one-by-one launches: the captured kernel sequences must give bitwise
the same results, on every replay, with other device work queued between the replays."""
import os
import sys

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

H2 = ([[0.7, 0.0, 0.0], [-0.7, 0.0, 0.0]], [1, 1])


@pytest.fixture(scope="module")
def lira_mod():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    import lira
    lira.lib()
    return lira


def _blob(n, h, centre, sigma, q):
    x, y, z = O.grid_points(n, h)
    r2 = (x - centre[0]) ** 2 + (y - centre[1]) ** 2 + (z - centre[2]) ** 2
    return q * np.exp(-r2 / (2 * sigma * sigma)) / (2 * np.pi * sigma * sigma) ** 1.5


@pytest.mark.parametrize("precision", [0, 1])
def test_poisson_graph_replays_are_bitwise_eager(lira_mod, precision):
    lira = lira_mod
    n, h = 41, 0.3
    rhos = [_blob(n, h, [0.3, -0.2, 0.1], 0.8, 1.0), _blob(n, h, [-1.0, 0.5, 0.7], 0.6, 2.0) -
            _blob(n, h, [1.2, 0.0, -0.4], 0.9, 1.0)]
    seq = [0, 1, 0, 1, 1]
    g = lira.GpuSolver(n, h, precision)
    g.set_graphs(True)
    e = lira.GpuSolver(n, h, precision)
    e.set_graphs(False)
    for i in seq:
        vg, sg = g.poisson(rhos[i], tol=1e-9)
        ve, se = e.poisson(rhos[i], tol=1e-9)
        assert sg.iterations == se.iterations and sg.iterations > 0
        assert np.array_equal(vg, ve), i


def test_scf_with_graphs_is_bitwise_the_eager_scf(lira_mod):
    lira = lira_mod
    n, h = 33, 0.25
    st = lira.default_settings(h, 1)
    g = lira.GpuSolver(n, h, 1)
    g.set_graphs(True)
    e = lira.GpuSolver(n, h, 1)
    e.set_graphs(False)
    for _ in range(2):
        a = g.solve(H2[0], H2[1], 1, 1, st)
        b = e.solve(H2[0], H2[1], 1, 1, st)
        assert a.energy.total == b.energy.total and a.iterations == b.iterations
        assert a.lobpcg_iterations == b.lobpcg_iterations and a.cg_iterations == b.cg_iterations
        assert np.array_equal(g.field(2), e.field(2))
        assert np.array_equal(g.forces(H2[0], H2[1], h), e.forces(H2[0], H2[1], h))


def test_pseudised_scf_with_graphs_is_bitwise_the_eager_scf(lira_mod):
    # the captured LOBPCG segment with the KB projections and scatter in it (a generated O table, water at 33^3)
    lira = lira_mod
    n, h = 33, 0.25
    r, half = 0.96 * 1.8897261, np.radians(104.5 / 2)
    pos = [[0.0, 0.0, 0.0], [r * np.sin(half), r * np.cos(half), 0.0], [-r * np.sin(half), r * np.cos(half), 0.0]]
    z = [8, 1, 1]
    oxygen = lira.generate_pseudo(8, h)
    st = lira.default_settings(h, 1)
    for precision in (1, 0):
        g = lira.GpuSolver(n, h, precision)
        g.set_graphs(True)
        g.set_ionic([oxygen])
        e = lira.GpuSolver(n, h, precision)
        e.set_graphs(False)
        e.set_ionic([oxygen])
        for _ in range(2):
            a = g.solve(pos, z, 4, 4, st)
            b = e.solve(pos, z, 4, 4, st)
            assert a.energy.total == b.energy.total and a.iterations == b.iterations, precision
            assert a.lobpcg_iterations == b.lobpcg_iterations and a.cg_iterations == b.cg_iterations
            assert np.array_equal(g.field(2), e.field(2))
            assert np.array_equal(g.forces(pos, z, h), e.forces(pos, z, h))
