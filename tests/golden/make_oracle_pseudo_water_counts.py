"""Runs the oracle's own SCF of the water workload (H2O with the O pseudised by the generated pseudopotential filtered
for h and the H all-electron, 4 + 4 electrons, 95^3 at h = 0.15 Bohr, LSDA, Fermi-Dirac kT 5 mHa, delta_rho < 1e-6,
f64) to convergence once and writes its step, LOBPCG-iteration and CG-iteration counts, block sizes and energy to
oracle_pseudo_water95_counts.json beside this script. bench.py's CPU baseline times a bounded sample of the oracle on
the GPU box's host and scales it with these counts (the oracle's own, not the GPU's). The table comes from the
library's generator (host code: no GPU is used).

Run from the repository root: OMP_NUM_THREADS=<cores> python tests/golden/make_oracle_water_counts.py
"""
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "atomic-sandbox_amd"))

import lira  # noqa: E402
import oracle as O  # noqa: E402

BOHR_PER_ANGSTROM = 1.8897261


def main():
    r = 0.96 * BOHR_PER_ANGSTROM
    half = math.radians(104.5 / 2)
    pos = [[0.0, 0.0, 0.0], [r * math.sin(half), r * math.cos(half), 0.0], [-r * math.sin(half), r * math.cos(half), 0.0]]
    z = [8, 1, 1]
    n, h = 95, 0.15
    t0 = time.perf_counter()
    species = [lira.generate_pseudo(8, h)]
    st, out = O.scf(n, h, pos, z, 4, 4, settings=O.default_settings(h, "fd"), want_rho=False, species=species)
    wall = time.perf_counter() - t0
    res = {"workload": "water H2O, O pseudised by the generated table (f = 1.4, p local, partial core) and H "
                       "all-electron, 95^3 at h = 0.15 Bohr, LSDA, Fermi-Dirac kT 5 mHa, delta_rho < 1e-6",
           "status": st.status, "steps": st.iterations, "lobpcg_iterations": st.lobpcg_iterations,
           "cg_iterations": st.cg_iterations, "k_up": st.k_up, "k_dn": st.k_dn, "energy": st.total,
           "delta_rho": st.delta_rho, "seconds": wall, "seconds_poisson": st.seconds_poisson,
           "seconds_eigen": st.seconds_eigen, "threads": int(os.environ.get("OMP_NUM_THREADS", "0") or 0)}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "oracle_pseudo_water95_counts.json")
    with open(path, "w") as fh:
        json.dump(res, fh, indent=1)
        fh.write("\n")
    print(json.dumps(res))


if __name__ == "__main__":
    main()
