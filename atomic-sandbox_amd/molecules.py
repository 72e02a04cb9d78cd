"""Molecules of the measured workloads. Small all-electron molecules for batched SCF: equilibrium geometries (Bohr,
centred on the nuclei's centroid) and a seeded sampler of displaced copies on grids of uneven size, the kind of
molecule set the reference's sampling passes generate and evaluate one at a time
(crates/providers/src/sampling/mod.rs:1635). And aspirin, the reference's largest shipped scene.

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


# Aspirin C9H8O4 (scenes/aspirin.scene:9-29, Bohr): PubChem CID 2244 conformer as the reference ships it; Z, x, y, z
ASPIRIN = [
    (8, -0.6738, -1.2212, 2.4999), (8, -3.9799, 4.4789, -1.5460), (8, -3.3422, 3.2919, 2.5142),
    (8, -3.2341, -3.6318, -0.1837), (6, 0.4778, -0.2199, 0.4251), (6, -0.4436, 1.9754, -0.7067),
    (6, 2.6070, -1.4187, -0.5641), (6, 0.7642, 2.9716, -2.8273), (6, 3.8149, -0.4224, -2.6846),
    (6, 2.8934, 1.7725, -3.8162), (6, -2.6620, 3.2579, 0.2971), (6, -2.5778, -2.9801, 1.9360),
    (6, -3.7237, -3.9678, 4.3332), (1, 3.3373, -3.1281, 0.3084), (1, 0.1091, 4.6924, -3.7371),
    (1, 5.4738, -1.3546, -3.4537), (1, 3.8379, 2.5512, -5.4637), (1, -2.2737, -4.9077, 5.4637),
    (1, -5.1973, -5.3422, 3.8736), (1, -4.5768, -2.4120, 5.3897), (1, -5.4738, 5.3422, -0.8680),
]


def aspirin(seed: int = 0, displacement=0.05):
    """Aspirin's nuclei (Bohr) with a seeded uniform +-displacement Angstrom per nucleus and coordinate, and their Z."""
    z = [a[0] for a in ASPIRIN]
    pos = np.array([a[1:] for a in ASPIRIN], dtype=np.float64)
    rng = np.random.default_rng(seed)
    return pos + rng.uniform(-displacement, displacement, size=pos.shape) * ANGSTROM, z
