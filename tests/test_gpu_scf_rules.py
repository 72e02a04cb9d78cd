"""GPU tests of the SCF's stopping and mixing rules and of the solve_polled contract at the boundary.

The rules restate the reference GPU solver's run loop (crates/qm-gpu/src/scf.rs:1237-1708): the partial-level stop of
the f32 solve (max(tol, 1e-5) on a step with a partly occupied level, :1519-1521, its test :2933-2967), the DIIS
restart when a frozen orbital set is re-solved and the freeze trim (:1538-1571), the per-channel charge
renormalisation (:1574-1582); the contract is solve_polled's (:1195-1235): the abort before each step, the inner abort
inside a Fermi-Dirac eigensolve, the observer, the parked iterate as the next warm start, and the readback-fault path
(LIRA_DEVICE with no state kept, :2429-2463).
"""
import threading

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


def test_open_shell_partial_level_stop(lira_mod):
    """All-electron C atom (nucleus smeared at sigma = h), 4 up + 2 down under Fermi-Dirac: the down channel holds one
    electron over three degenerate p levels (f = 1/3 each). The f32 solve tests those steps against max(tol, 1e-5)
    and reports it in tol_used; the f64 solve keeps tol (it has no occupation-noise floor). The energies differ by
    what stopping at 1e-5 costs: F is second order in the residual (the reference measured <= 2.9 uHa on all-electron
    O, scf.rs:276-283)."""
    lira = lira_mod
    n, h = 49, 0.25
    pos, z = [[0.0, 0.0, 0.0]], [6]
    st = lira.default_settings(h, lira.FERMI_DIRAC)
    out = {}
    for prec in (lira.MIXED, lira.FP64):
        s = lira.GpuSolver(n, h, prec)
        r = s.solve(pos, z, 4, 2, st)
        p, dr, _ = s.profile()
        out[prec] = r
        print(f"C atom precision {prec}: {r.iterations} steps, delta_rho {r.delta_rho:.2e}, tol_used {r.tol_used:.0e}, "
              f"E {r.energy.total:.9f}, partial {r.partial_levels()}, DIIS resets {p.diis_resets}, "
              f"frozen {p.frozen_steps}, trims {p.diis_freeze_trims}, locked calls {p.locked_calls}")
        thirds = [lv for lv in r.partial_levels() if abs(lv[3] - 1.0 / 3.0) < 1e-3]
        assert len(thirds) == 3 and all(sp == 1 for sp, _, _, _ in thirds)
    mixed, fp64 = out[lira.MIXED], out[lira.FP64]
    assert mixed.tol_used == 1e-5 and mixed.delta_rho < 1e-5 and mixed.iterations <= 60
    assert fp64.tol_used == 1e-6 and fp64.delta_rho < 1e-6
    assert abs(mixed.energy.total - fp64.energy.total) < 3e-6


def test_gapped_step_keeps_settings_tol(lira_mod):
    """A gapped Fermi-Dirac state (closed-shell H2) has no partial level: tol_used is settings.tol in both precisions."""
    lira = lira_mod
    for prec in (lira.MIXED, lira.FP64):
        s = lira.GpuSolver(33, 0.25, prec)
        r = s.solve(H2[0], H2[1], 1, 1, lira.default_settings(0.25, lira.FERMI_DIRAC))
        assert r.n_partial == 0 and r.tol_used == 1e-6 and r.delta_rho < 1e-6


def test_diis_history_15_matches_the_default_history(lira_mod):
    """The longest history the settings allow (15 entries: 34 reductions in the step's one readback, more than one
    reduction launch holds) converges to the default history's answer."""
    lira = lira_mod
    n, h = 33, 0.25
    base = lira.default_settings(h, lira.AUFBAU)
    long = lira.default_settings(h, lira.AUFBAU, diis_history=15)
    a = lira.GpuSolver(n, h, lira.FP64).solve(H2[0], H2[1], 1, 1, base)
    b = lira.GpuSolver(n, h, lira.FP64).solve(H2[0], H2[1], 1, 1, long)
    assert abs(a.energy.total - b.energy.total) < 1e-9
    with pytest.raises(lira.Invalid):
        lira.GpuSolver(n, h, lira.FP64).solve(H2[0], H2[1], 1, 1, lira.default_settings(h, 0, diis_history=16))


def test_observer_sees_every_step(lira_mod):
    """One progress record per step after its delta_rho readback: the sequence equals the profile's, the stall watch
    fields are the solve's, and the Harris-Foulkes energy of the last (converged) step's input density lands on the
    final energy. An observer changes no number of the solve."""
    lira = lira_mod
    n, h = 33, 0.25
    st = lira.default_settings(h, lira.FERMI_DIRAC)
    seen = []
    s = lira.GpuSolver(n, h, lira.FP64)
    status, r = s.solve_polled(H2[0], H2[1], 1, 1, st, observe=seen.append)
    assert status == lira.OK
    _, dr, _ = s.profile()
    assert [p["iteration"] for p in seen] == list(range(1, r.iterations + 1))
    assert np.array_equal(np.array([p["delta_rho"] for p in seen]), dr)
    best = np.minimum.accumulate(dr)
    assert np.array_equal(np.array([p["best_delta_rho"] for p in seen]), best)
    assert all(p["window"] == st.stall_window and p["tol"] == st.tol for p in seen)
    assert abs(seen[-1]["energy"] - r.energy.total) < 1e-5
    plain = lira.GpuSolver(n, h, lira.FP64).solve(H2[0], H2[1], 1, 1, st)
    assert plain.energy.total == r.energy.total and plain.iterations == r.iterations


def test_abort_parks_the_iterate_as_the_warm_start(lira_mod):
    """An abort before step 4 abandons the solve (no converged state) and parks its iterate: the next warm solve
    converges in fewer steps than a cold one and lands on the cold energy."""
    lira = lira_mod
    n, h = 33, 0.25
    st = lira.default_settings(h, lira.AUFBAU)
    s = lira.GpuSolver(n, h, lira.MIXED)
    calls = []

    def abort():
        calls.append(1)
        return len(calls) > 3

    status, r = s.solve_polled(H2[0], H2[1], 1, 1, st, abort=abort)
    p, _, _ = s.profile()
    assert status == lira.ABORTED and r.iterations == 3 and p.abandoned_before_step == 4
    with pytest.raises(lira.Invalid):
        s.levels(0)
    warm = s.solve(H2[0], H2[1], 1, 1, st, warm=True)
    cold = lira.GpuSolver(n, h, lira.MIXED).solve(H2[0], H2[1], 1, 1, st)
    print(f"after an abort at step 4: warm {warm.iterations} steps, cold {cold.iterations}")
    assert warm.iterations < cold.iterations
    assert abs(warm.energy.total - cold.energy.total) < 2e-6


def test_inner_abort_stops_a_fermi_dirac_eigensolve(lira_mod):
    """inner_abort is polled once per LOBPCG iteration of the Fermi-Dirac eigensolves: the solve is abandoned inside
    the step's eigensolve, the step does not count, and the parked iterate warm-starts the next solve."""
    lira = lira_mod
    n, h = 33, 0.25
    st = lira.default_settings(h, lira.FERMI_DIRAC)
    s = lira.GpuSolver(n, h, lira.MIXED)
    polls = []

    def inner():
        polls.append(1)
        return len(polls) > 12

    status, r = s.solve_polled(H2[0], H2[1], 1, 1, st, inner_abort=inner)
    p, _, _ = s.profile()
    assert status == lira.ABORTED and p.abandoned_in_step >= 1 and p.abandoned_before_step == 0
    assert r.iterations == p.abandoned_in_step - 1
    warm = s.solve(H2[0], H2[1], 1, 1, st, warm=True)
    cold = lira.GpuSolver(n, h, lira.MIXED).solve(H2[0], H2[1], 1, 1, st)
    assert abs(warm.energy.total - cold.energy.total) < 2e-6
    # Aufbau eigensolves are not polled
    polls.clear()
    status, _ = lira.GpuSolver(n, h, lira.MIXED).solve_polled(H2[0], H2[1], 1, 1, lira.default_settings(h, 0),
                                                              inner_abort=inner)
    assert status == lira.OK and not polls


@pytest.mark.parametrize("where", [0.0, 0.3, 1.0])
def test_readback_fault_is_a_device_error_and_drops_the_state(lira_mod, where):
    """A readback reported failed (host-side injection; every kernel runs normally) ends the solve with LIRA_DEVICE
    wherever it strikes (the first readback, one early in the solve, the last one: the energy's), and the solver holds
    no state: the next warm solve is a cold one, bit for bit."""
    lira = lira_mod
    n, h = 25, 0.3
    st = lira.default_settings(h, lira.AUFBAU)
    twin = lira.GpuSolver(n, h, lira.MIXED)  # the same solves without a fault count the warm solve's readbacks
    twin.solve(H2[0], H2[1], 1, 1, st)
    twin.solve(H2[0], H2[1], 1, 1, st, warm=True)
    count = twin.profile()[0].readbacks
    after = min(int(where * count), count - 1)
    s = lira.GpuSolver(n, h, lira.MIXED)
    s.solve(H2[0], H2[1], 1, 1, st)
    s.inject_readback_fault(after)
    with pytest.raises(lira.Device):
        s.solve(H2[0], H2[1], 1, 1, st, warm=True)
    with pytest.raises(lira.Invalid):
        s.levels(0)
    again = s.solve(H2[0], H2[1], 1, 1, st, warm=True)
    cold = lira.GpuSolver(n, h, lira.MIXED).solve(H2[0], H2[1], 1, 1, st)
    assert again.iterations == cold.iterations and again.energy.total == cold.energy.total


def test_core_overlaps_are_reported(lira_mod):
    """Two pseudised centres closer than the sum of their core radii: one core overlap, also a warning; the solve is
    still taken (crates/qm/src/ionic.rs:175-185, crates/qm/src/scf.rs:1083-1090)."""
    lira = lira_mod
    import synthetic_tables as T
    n, h = 33, 0.25
    sp = T.local_only(3, 1.0, h)
    sp.r_core = 1.2
    s = lira.GpuSolver(n, h, lira.FP64)
    s.set_ionic([sp])
    pos = [[1.0, 0.0, 0.0], [-1.0, 0.0, 0.0]]
    r = s.solve(pos, [3, 3], 1, 1, lira.default_settings(h, lira.AUFBAU))
    assert r.overlaps() == [(0, 1, pytest.approx(2.0, abs=1e-12))]
    assert any(w.startswith("cores overlap: 0-1") for w in r.warning_lines())
    sp.r_core = 0.9
    s.set_ionic([sp])
    r = s.solve(pos, [3, 3], 1, 1, lira.default_settings(h, lira.AUFBAU))
    assert r.n_core_overlaps == 0


def test_forces_on_many_nuclei(lira_mod):
    """Hellmann-Feynman forces of a converged density on 70 nuclei (more than the solver's fixed reduction scratch
    once held): against the oracle's forces on the same density."""
    lira = lira_mod
    n, h = 33, 0.3
    s = lira.GpuSolver(n, h, lira.FP64)
    st = lira.default_settings(h, lira.AUFBAU)
    s.solve(H2[0], H2[1], 1, 1, st)
    rng = np.random.default_rng(7)
    pos = rng.uniform(-4.0, 4.0, size=(70, 3))
    z = [1] * 70
    f = s.forces(pos, z, st.sigma)
    rho = s.field(2)
    _, fo = O.forces(n, h, rho, pos, z, st.sigma)
    scale = np.abs(fo).max()
    assert np.abs(f - fo).max() <= 1e-10 * max(scale, 1.0)


def test_solvers_on_two_threads_with_default_settings(lira_mod):
    """Two host threads, each solving on its own solver with the default settings, give the lone solves' results
    bit for bit."""
    lira = lira_mod
    n, h = 33, 0.25
    st = lira.default_settings(h, lira.FERMI_DIRAC)
    geoms = [H2, ([[0.75, 0.1, 0.0], [-0.7, 0.0, 0.0]], [1, 1])]
    lone = [lira.GpuSolver(n, h, lira.MIXED).solve(g[0], g[1], 1, 1, st).energy.total for g in geoms]
    got = [None, None]

    def work(i):
        got[i] = lira.GpuSolver(n, h, lira.MIXED).solve(geoms[i][0], geoms[i][1], 1, 1, st).energy.total

    th = [threading.Thread(target=work, args=(i,)) for i in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert got == lone


def test_evaluation_repeats_a_device_failure(lira_mod):
    """The provider's "GPU busy" policy: an evaluation whose readback fails is repeated after a pause and returns the
    clean evaluation's answer, flagged (crates/providers/src/ks.rs:317-350, 1688-1706)."""
    lira = lira_mod
    n, h = 25, 0.3
    st = lira.default_settings(h, lira.FERMI_DIRAC)
    clean = lira.GpuSolver(n, h, lira.MIXED).evaluate(H2[0], H2[1], 1, 1, st)
    s = lira.GpuSolver(n, h, lira.MIXED)
    s.inject_readback_fault(4)
    status, stats, f, flags, _ = s.evaluate(H2[0], H2[1], 1, 1, st)
    assert status == lira.OK and flags & lira.EVAL_DEVICE_RETRIED
    assert stats.energy.total == clean[1].energy.total and np.array_equal(f, clean[2])


def test_evaluation_takes_its_witness_cap_from_the_budget(lira_mod):
    """A device memory budget sets the witness cap to the largest that fits (never below 3): the solve's block stays
    within it and the evaluation says where its cap came from."""
    lira = lira_mod
    n, h = 33, 0.25
    budget = lira.solve_footprint_bytes(n, lira.MIXED, 1, 1, 5) + 1
    s = lira.GpuSolver(n, h, lira.MIXED)
    status, stats, _, flags, _ = s.evaluate(H2[0], H2[1], 1, 1, lira.default_settings(h, lira.FERMI_DIRAC),
                                            vram_budget=budget)
    assert status == lira.OK and flags & lira.EVAL_BUDGET_CAP
    assert s.witness_cap() == 5 and max(stats.n_stored) <= 1 + 5
