"""GPU solves on the generated pseudopotential tables (lira_pseudo_generate) against the reference's own numbers and
against the oracle.

- The reference's V4 pseudo-O cases (crates/validate/tests/gpu_vs_cpu.rs:798-836, recorded in
  crates/validate/references.md:341-345): the lone pseudo-O (partial core and projectors, tables filtered at h = 0.25)
  on the middle grid point of 49^3 x 0.25, Fermi-Dirac at kT = 5 mHa, sigma = h, tol 1e-6, at multiplicity 3 and 1.
  Bounds as the reference's: F within E_BOUND_PSEUDO = 6e-6 Ha of its CPU F (-16.114269705, -16.067788821), TS in the
  symmetric class (within 5 % of 3 kT s(1/3), 6 kT s(2/3)), forces within F_BOUND_PSEUDO = 6e-5 Ha/Bohr of the
  symmetric site's zero; MIXED stops at the partial-level tolerance 1e-5, FP64 at 1e-6.
- Tight FP64 parity with the oracle on the generated table: pseudised water (O pseudised, H all-electron) on
  33^3 x 0.25, energies 1e-10, levels and forces 1e-8, as the synthetic-table tests.
"""
import math
import os
import sys

import numpy as np
import pytest

import oracle as O

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "atomic-sandbox_amd"))

pytestmark = pytest.mark.gpu

KT = 5e-3
F_CPU = {3: -16.114269705, 1: -16.067788821}  # references.md:343-344, the CPU (crates/qm) column
E_BOUND_PSEUDO, F_BOUND_PSEUDO = 6e-6, 6e-5


def level_entropy(f):
    return -(f * math.log(f) + (1.0 - f) * math.log(1.0 - f))


TS_CLOSED = {3: 3.0 * KT * level_entropy(1.0 / 3.0), 1: 6.0 * KT * level_entropy(2.0 / 3.0)}


@pytest.fixture(scope="module")
def lira_mod():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    import lira
    lira.lib()
    return lira


@pytest.fixture(scope="module")
def oxygen(lira_mod):
    return lira_mod.generate_pseudo(8, 0.25)


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("mult", [3, 1])
def test_pseudo_oxygen_matches_the_reference_v4_case(lira_mod, oxygen, precision, mult):
    lira = lira_mod
    n, h = 49, 0.25
    nu, nd = (4, 2) if mult == 3 else (3, 3)
    s = lira.GpuSolver(n, h, precision)
    s.set_ionic([oxygen])
    st = s.solve([[0.0, 0.0, 0.0]], [8], nu, nd, lira.default_settings(h, lira.FERMI_DIRAC))
    f = s.forces([[0.0, 0.0, 0.0]], [8], h)
    print(f"pseudo-O M = {mult} precision {precision}: F {st.energy.total:+.9f} (reference CPU {F_CPU[mult]:+.9f}, "
          f"d {st.energy.total - F_CPU[mult]:+.2e}), TS {st.energy.ts:.6e}, {st.iterations} steps, tol_used "
          f"{st.tol_used:g}, |F_O| {np.abs(f).max():.1e}")
    assert abs(st.energy.total - F_CPU[mult]) < E_BOUND_PSEUDO
    assert abs(st.energy.ts - TS_CLOSED[mult]) < 0.05 * TS_CLOSED[mult]
    assert np.abs(f).max() < F_BOUND_PSEUDO
    assert st.tol_used == (1e-5 if precision == 1 else 1e-6)
    assert st.electrons[0] + st.electrons[1] == pytest.approx(6.0, abs=1e-6)


def test_generated_table_scf_matches_oracle(lira_mod, oxygen):
    # pseudised water (generated O at h = 0.25, the two H all-electron: 8 electrons, 4 + 4), both solvers at
    # delta_rho < 1e-9, eig_tol 1e-9: energies, levels and forces on the same discrete problem
    lira = lira_mod
    n, h = 33, 0.25
    r, half = 0.96 * 1.8897261, math.radians(104.5 / 2)
    pos = [[0.0, 0.0, 0.0], [r * math.sin(half), r * math.cos(half), 0.0], [-r * math.sin(half), r * math.cos(half), 0.0]]
    z = [8, 1, 1]
    s = lira.GpuSolver(n, h, 0)
    s.set_ionic([oxygen])
    st = s.solve(pos, z, 4, 4, lira.default_settings(h, 0, tol=1e-9, eig_tol=1e-9))
    so, out = O.scf(n, h, pos, z, 4, 4, settings=O.default_settings(h, "aufbau", tol=1e-9, eig_tol=1e-9),
                    species=[oxygen])
    assert so.status == O.OK
    print(f"generated-table water 33^3: GPU E {st.energy.total:.12f} ({st.iterations} steps) oracle "
          f"{so.total:.12f} ({so.iterations} steps)")
    assert abs(st.energy.total - so.total) < 1e-10
    assert np.abs(s.state().evals_up - np.sort(out["evals_up"])).max() < 1e-8
    assert np.abs(s.forces(pos, z, h) - out["forces"]).max() < 1e-8


def test_dimer_gate_keeps_the_reference_hydrogen_and_helium(lira_mod):
    # the molecular gate on f for Z <= 2 (crates/providers/src/pseudo.rs GridDimerGate; recorded in
    # crates/validate/references.md:3623-3631): H's f = 1.4 entries fail on R_e and f = 1.2 with the s local passes,
    # pseudo-H2 R_e 1.4340 Bohr, omega 0.018732 a.u. against the all-electron sigma -> 0 dimer's 1.4489, 0.019402 by the
    # same procedure; He has no all-electron force zero in [1, 4] r_peak and its pseudised wall agrees: not
    # applicable, f = 1.6. Bounds: 2e-3 Bohr on R_e (the reference's own GPU-vs-CPU spread, 2.2e-3) and 0.5 % on omega.
    lira = lira_mod
    h = lira.generate_pseudo(1, 0.25, gated=True).record
    print(f"gated H: f {h['f']} {h['local']} {h['dimer']}, PP {h['dimer_pp']}, AE {h['dimer_ae']}, {h['dimer_solves']} solves")
    assert (h["f"], h["local"], h["dimer"], h["provisional"]) == (1.2, "s-local", "pass", False)
    assert h["dimer_pp"][0] == pytest.approx(1.4340, abs=2e-3)
    assert h["dimer_pp"][1] == pytest.approx(0.018732, rel=5e-3)
    assert h["dimer_ae"][0] == pytest.approx(1.4489, abs=2e-3)
    assert h["dimer_ae"][1] == pytest.approx(0.019402, rel=5e-3)
    he = lira.generate_pseudo(2, 0.25, gated=True).record
    assert (he["f"], he["local"], he["dimer"], he["provisional"]) == (1.6, "s-local", "not applicable", False)
    # the ungated generation of the same element is the reference's provisional table
    assert lira.generate_pseudo(1, 0.25).record["provisional"] is True
