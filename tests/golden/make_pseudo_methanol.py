"""Writes the oracle's tight f64 SCF of pseudised methanol (the reduced pseudised organic of the adjoint-gradient
workload: C and O pseudised with the closed-form C-like and O-like tables, four all-electron H; 14 valence electrons,
7 + 7, Aufbau, 33^3 at h = 0.3 Bohr, delta_rho < 1e-9 and eig_tol 1e-9) to pseudo_methanol_33.json beside this script:
geometry, energy terms, eigenvalues, the five pseudised force terms and the forces. The GPU test compares its FP64
solve with these numbers, so the GPU box need not run the oracle's ~1-minute solve.

Run from the repository root: OMP_NUM_THREADS=<cores> python tests/golden/make_pseudo_methanol.py
"""
import json
import math
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "atomic-sandbox_amd"))

import oracle as O  # noqa: E402
import synthetic_tables as T  # noqa: E402

BOHR_PER_ANGSTROM = 1.8897261


def geometry():
    """Methanol (Angstrom: C-O 1.427, O-H 0.96 at 108.5 deg, C-H 1.093), converted to Bohr, centroid at the origin."""
    a = BOHR_PER_ANGSTROM
    t = math.radians(180.0 - 108.5)
    pos = np.array([[0.0, 0.0, 0.0], [1.427, 0.0, 0.0], [1.427 + 0.96 * math.cos(t), 0.96 * math.sin(t), 0.0],
                    [-0.364, 1.030, 0.0], [-0.364, -0.515, 0.892], [-0.364, -0.515, -0.892]]) * a
    return pos - pos.mean(axis=0), [6, 8, 1, 1, 1, 1]


def main():
    n, h = 33, 0.3
    pos, z = geometry()
    species = [T.carbon_like(h), T.oxygen_like(h)]
    st, out = O.scf(n, h, pos, z, 7, 7, settings=O.default_settings(h, "aufbau", tol=1e-9, eig_tol=1e-9),
                    species=species, want_orbitals=True)
    assert st.status == O.OK, st.status
    _, terms = O.pseudo_force_terms(n, h, out["rho_up"], out["rho_dn"], np.concatenate([out["orbitals_up"],
                                 out["orbitals_dn"]]), np.concatenate([out["occ_up"], out["occ_dn"]]), pos, z,
                                 species, h)
    res = {"n": n, "h": h, "pos": pos.tolist(), "z": z, "n_up": 7, "n_dn": 7, "steps": st.iterations,
           "energy": {k: getattr(st, k) for k in ("kinetic", "external", "hartree", "xc", "nuclear", "internal",
                                                 "ts", "total")},
           "evals_up": np.sort(out["evals_up"]).tolist(), "forces": out["forces"].tolist(),
           "force_terms": np.asarray(terms).tolist()}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pseudo_methanol_33.json")
    with open(path, "w") as fh:
        json.dump(res, fh, indent=1)
        fh.write("\n")
    print(f"methanol 33^3: E {st.total:.12f}, {st.iterations} steps")


if __name__ == "__main__":
    main()
