"""Synthetic pseudo-ion tables: smooth, compactly supported stand-ins for qm::radial::filter::FilteredTable.

The reference builds its tables with a radial atom solver and a band-limiting filter (crates/qm/src/radial/*), which
sit outside the SCF hot path. These closed-form tables exercise the same consumers — the local V_SR correction, the
Kleinman–Bylander projectors (s and p) and the partial core — with the reference's table conventions
(crates/qm/src/radial/filter.rs:302-363): uniform spacing du = h / 256 from d = 0, an even function sampled at
d_i = i * du, exactly 0 beyond each table's radius, 4-point Lagrange reads with index -1 mirrored.

Every radial profile is g(d) = A exp(-d^2 / (2 w^2)) (1 - (d/R)^2)^4 for d < R and 0 beyond: even in d and C^3 at R.
"""
from __future__ import annotations

import math

import numpy as np

TABLE_POINTS_PER_H = 256  # crates/qm/src/radial/ledger.rs:163
SIGMA_LR_OVER_H = 2.0  # crates/qm/src/radial/ledger.rs:155


def bump_gaussian(d, amp, width, radius):
    d = np.asarray(d, dtype=np.float64)
    x = np.clip(d / radius, 0.0, 1.0)
    return np.where(d < radius, amp * np.exp(-d * d / (2.0 * width * width)) * (1.0 - x * x) ** 4, 0.0)


def uniform_table(du, radius, amp, width):
    """Samples g at i * du for i = 0 .. floor(radius / du) + 3 (the last stencil of a read at d <= radius)."""
    m = int(math.floor(radius / du)) + 4
    return bump_gaussian(np.arange(m) * du, amp, width, radius)


class Projector:
    """FilteredProjector (filter.rs:323): l, E_KB (Ha), radius (Bohr), t_u = F_l / d^l on the uniform table."""

    def __init__(self, l, e_kb, radius, t_u):
        self.l, self.e_kb, self.radius = int(l), float(e_kb), float(radius)
        self.t_u = np.ascontiguousarray(t_u, dtype=np.float64)


class Species:
    """The fields of FilteredTable (filter.rs:330-338) the grid reads: Z, Z_v, sigma_LR, du, the V_SR table and its
    radius, the partial-core table and its radius (<= 0: no core), the projectors, and the derived core radius r_core
    (two cores closer than the sum of theirs warn)."""

    def __init__(self, z, z_v, sigma_lr, du, v_sr_u, r_sr, rho_pc_u=None, pc_radius=0.0, projectors=(), r_core=0.0):
        self.z, self.z_v, self.sigma_lr, self.du = int(z), float(z_v), float(sigma_lr), float(du)
        self.v_sr_u = np.ascontiguousarray(v_sr_u, dtype=np.float64)
        self.r_sr = float(r_sr)
        self.rho_pc_u = None if rho_pc_u is None else np.ascontiguousarray(rho_pc_u, dtype=np.float64)
        self.pc_radius = float(pc_radius) if rho_pc_u is not None else 0.0
        self.projectors = list(projectors)
        self.r_core = float(r_core)


# profile parameters (amplitude, width, radius) of the O-like ion: Z = 8, Z_v = 6
O_V_SR = (4.0, 0.45, 1.6)
O_PC = (2.0, 0.35, 1.2)
O_S = (1.0, 0.45, 1.4, 1.5)  # (amplitude, width, radius, E_KB)
O_P = (1.0, 0.50, 1.5, 0.6)


def oxygen_like(h, partial_core=True, projectors=True):
    """An O-like pseudo-ion (Z = 8, Z_v = 6) filtered for spacing h: a repulsive V_SR, an s and a p projector with
    E_KB > 0, and a partial core of ~1.4 electrons."""
    du = h / TABLE_POINTS_PER_H
    projs = []
    if projectors:
        for l, (a, w, r, e) in ((0, O_S), (1, O_P)):
            projs.append(Projector(l, e, r, uniform_table(du, r, a, w)))
    pc = uniform_table(du, O_PC[2], O_PC[0], O_PC[1]) if partial_core else None
    return Species(8, 6.0, SIGMA_LR_OVER_H * h, du, uniform_table(du, O_V_SR[2], O_V_SR[0], O_V_SR[1]), O_V_SR[2],
                   pc, O_PC[2] if partial_core else 0.0, projs)


# the C-like ion: Z = 6, Z_v = 4 (same profile family, softer and wider)
C_V_SR = (3.0, 0.50, 1.7)
C_PC = (1.0, 0.40, 1.3)
C_S = (1.0, 0.50, 1.5, 1.2)
C_P = (1.0, 0.55, 1.6, 0.5)


def carbon_like(h, partial_core=True, projectors=True):
    """A C-like pseudo-ion (Z = 6, Z_v = 4) filtered for spacing h: a repulsive V_SR, an s and a p projector with
    E_KB > 0, and a partial core."""
    du = h / TABLE_POINTS_PER_H
    projs = []
    if projectors:
        for l, (a, w, r, e) in ((0, C_S), (1, C_P)):
            projs.append(Projector(l, e, r, uniform_table(du, r, a, w)))
    pc = uniform_table(du, C_PC[2], C_PC[0], C_PC[1]) if partial_core else None
    return Species(6, 4.0, SIGMA_LR_OVER_H * h, du, uniform_table(du, C_V_SR[2], C_V_SR[0], C_V_SR[1]), C_V_SR[2],
                   pc, C_PC[2] if partial_core else 0.0, projs)


def local_only(z, z_v, h, sigma_lr=None):
    """A pseudo-ion with V_SR = 0, no projectors and no core: the smeared charge Z_v at sigma_LR alone."""
    du = h / TABLE_POINTS_PER_H
    return Species(z, z_v, SIGMA_LR_OVER_H * h if sigma_lr is None else sigma_lr, du, np.zeros(8), 4 * du)
