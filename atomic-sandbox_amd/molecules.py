"""Small all-electron molecules for batched SCF: equilibrium geometries (Bohr, centred on the nuclei's centroid) and a
seeded sampler of displaced copies on grids of uneven size, the kind of molecule set the reference's sampling passes
generate and evaluate one at a time (crates/providers/src/sampling/mod.rs:1635).

Equilibrium bond lengths and angles (NIST CCCBDB experimental): H2 0.741 A; LiH 1.595 A; H2O O-H 0.9572 A,
H-O-H 104.52 deg; NH3 N-H 1.012 A, H-N-H 106.67 deg; CH4 C-H 1.087 A (tetrahedral).
"""
import math

import numpy as np

from lira import Molecule

ANGSTROM = 1.8897261  # Bohr per Angstrom (scenes/aspirin.scene:3)

# name -> (Z list, spin-up electrons, spin-down electrons)
SPECIES = {
    "H2": ([1, 1], 1, 1),
    "LiH": ([3, 1], 2, 2),
    "H2O": ([8, 1, 1], 5, 5),
    "NH3": ([7, 1, 1, 1], 5, 5),
    "CH4": ([6, 1, 1, 1, 1], 5, 5),
}


def equilibrium(name: str) -> np.ndarray:
    """Nuclear positions (Bohr) of the equilibrium geometry, centroid at the origin."""
    a = ANGSTROM
    if name == "H2":
        p = [[0, 0, -0.3705 * a], [0, 0, 0.3705 * a]]
    elif name == "LiH":
        p = [[0, 0, 0], [0, 0, 1.595 * a]]
    elif name == "H2O":
        r, half = 0.9572 * a, math.radians(104.52) / 2
        p = [[0, 0, 0], [r * math.sin(half), r * math.cos(half), 0], [-r * math.sin(half), r * math.cos(half), 0]]
    elif name == "NH3":
        r, alpha = 1.012 * a, math.radians(106.67)
        # H at polar angle t from the -z axis: cos(alpha) = 1.5 cos^2 t - 0.5
        t = math.acos(math.sqrt((math.cos(alpha) + 0.5) / 1.5))
        p = [[0, 0, 0]] + [[r * math.sin(t) * math.cos(f), r * math.sin(t) * math.sin(f), -r * math.cos(t)]
                           for f in (0.0, 2 * math.pi / 3, 4 * math.pi / 3)]
    elif name == "CH4":
        d = 1.087 * a / math.sqrt(3)
        p = [[0, 0, 0], [d, d, d], [d, -d, -d], [-d, d, -d], [-d, -d, d]]
    else:
        raise KeyError(name)
    p = np.asarray(p, dtype=np.float64)
    return p - p.mean(axis=0)


def sample(count: int, seed: int = 0, h: float = 0.25, sizes=tuple(range(48, 97, 8)), displacement=0.1):
    """count molecules drawn uniformly from SPECIES, each at its equilibrium geometry plus a seeded uniform
    +-displacement Angstrom per nucleus and coordinate, on an n^3 grid with n drawn from sizes."""
    rng = np.random.default_rng(seed)
    names = sorted(SPECIES)
    out = []
    for _ in range(count):
        name = names[int(rng.integers(len(names)))]
        z, nu, nd = SPECIES[name]
        pos = equilibrium(name) + rng.uniform(-displacement, displacement, size=(len(z), 3)) * ANGSTROM
        n = int(sizes[int(rng.integers(len(sizes)))])
        out.append(Molecule(n=n, h=h, pos=pos, z=list(z), n_up=nu, n_dn=nd, name=name))
    return out
