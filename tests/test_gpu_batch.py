"""Batched SCF on the GPU (lira_solve_batch): molecules spread over host worker threads, each with its own solver and
stream, give exactly the results of lone solves on fresh solvers, and one molecule's failure stays its own."""
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


def _lone(lira, m, precision):
    s = lira.GpuSolver(m.n, m.h, precision)
    st = s.solve(m.pos, m.z, m.n_up, m.n_dn, lira.default_settings(m.h))
    return st, s.forces(m.pos, m.z, m.h)


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_equals_lone_solves_bit_for_bit(lira_mod, precision):
    lira = lira_mod
    import molecules as M
    mols = M.sample(6, seed=11, h=0.3, sizes=(33, 37, 41))
    print("batch:", [(m.name, m.n) for m in mols], flush=True)
    out = lira.solve_batch(mols, precision=precision, workers=3)
    assert len({o.worker for o in out}) >= 2
    for m, o in zip(mols, out):
        assert o.status == lira.OK, (m.name, o.message)
        assert o.end > o.start >= 0.0
        st, f = _lone(lira, m, precision)
        assert o.stats.iterations == st.iterations, m.name
        assert o.stats.energy.total == st.energy.total, (m.name, o.stats.energy.total - st.energy.total)
        assert o.stats.delta_rho == st.delta_rho
        assert np.array_equal(o.forces, f), m.name
        assert o.forces.shape == (len(m.z), 3)


def test_a_failing_molecule_does_not_stop_the_batch(lira_mod):
    lira = lira_mod
    import molecules as M
    mols = M.sample(3, seed=5, h=0.3, sizes=(33,))
    bad = lira.Molecule(n=33, h=0.3, pos=np.zeros((2, 3)), z=[1, 1], n_up=1, n_dn=1, name="coincident")
    out = lira.solve_batch(mols[:1] + [bad] + mols[1:], precision=1, workers=2)
    assert out[1].status == lira.COINCIDENT_NUCLEI and out[1].message
    assert [o.status for o in out[:1] + out[2:]] == [lira.OK] * 3
    st, f = _lone(lira, mols[1], 1)
    assert out[2].stats.energy.total == st.energy.total and np.array_equal(out[2].forces, f)


def test_batch_takes_per_molecule_settings(lira_mod):
    # explicit settings per molecule (Aufbau occupations, a tighter tolerance on one) are the ones each solve uses
    lira = lira_mod
    import molecules as M
    mols = M.sample(3, seed=8, h=0.3, sizes=(33,))
    sets = [lira.default_settings(m.h, lira.AUFBAU) for m in mols]
    sets[1].tol = 1e-8
    out = lira.solve_batch(mols, precision=0, workers=2, settings=sets)
    for m, st_, o in zip(mols, sets, out):
        assert o.status == lira.OK, o.message
        s = lira.GpuSolver(m.n, m.h, 0)
        lone = s.solve(m.pos, m.z, m.n_up, m.n_dn, st_)
        assert o.stats.energy.total == lone.energy.total and o.stats.iterations == lone.iterations
        assert o.stats.delta_rho < st_.tol and o.stats.energy.ts == 0.0


def test_batch_workload_grids_at_h_025(lira_mod):
    """The batched workload's own molecule set at its own spacing and grids (H2, LiH, H2O, NH3, CH4 with +-0.1 A
    seeded displacements, n from 48 to 96 at h = 0.25 Bohr, MIXED, 4 workers): every molecule converges, each result is
    its lone solve's bit for bit, each channel holds its electron count, and the net force (zero for the exact
    functional, broken only by the grid's egg-box effect) is small against the largest component."""
    lira = lira_mod
    import molecules as M
    mols = M.sample(10, seed=21, h=0.25)
    names = {m.name for m in mols}
    assert names == set(M.SPECIES) and min(m.n for m in mols) >= 48 and max(m.n for m in mols) <= 96
    print("batch:", [(m.name, m.n) for m in mols], flush=True)
    out = lira.solve_batch(mols, precision=lira.MIXED, workers=4)
    for m, o in zip(mols, out):
        assert o.status == lira.OK, (m.name, o.message)
        st, f = _lone(lira, m, lira.MIXED)
        assert o.stats.energy.total == st.energy.total and o.stats.iterations == st.iterations, m.name
        assert np.array_equal(o.forces, f), m.name
        assert abs(o.stats.electrons[0] - m.n_up) < 1e-5 and abs(o.stats.electrons[1] - m.n_dn) < 1e-5
        net = np.abs(o.forces.sum(axis=0)).max()
        print(f"{m.name} {m.n}^3: E {o.stats.energy.total:.8f}, {o.stats.iterations} steps, net force {net:.1e}, "
              f"max {np.abs(o.forces).max():.2e}")
        assert net < 5e-2 * np.abs(o.forces).max() + 1e-3
