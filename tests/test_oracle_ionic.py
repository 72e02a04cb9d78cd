"""Pins the oracle's pseudised-ion restatement (qm/src/nonlocal.rs, ionic.rs, nuclear.rs:290-382,
radial/filter.rs:302-363, forces.rs:259-350) before it checks the HIP path.

The tables are the closed-form synthetic ones of synthetic_tables (the reference's own tables come from its radial
atom solver, which is outside the path). Every check is a closed form (the profiles are known functions), an identity
(symmetry, rank, the all-electron limit) or a derivative: each force term must be minus the central difference of
its own frozen energy, the reference's test strategy (forces.rs:494-560).
"""
import math
import os
import sys

import numpy as np
import pytest

import oracle as O

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "atomic-sandbox_amd"))
import synthetic_tables as T  # noqa: E402

H = 0.25
N = 25
WATER = ([[0.05, -0.03, 0.02], [1.43, 1.11, 0.0], [-1.43, 1.11, 0.0]], [8, 1, 1])


def _closed(d, prof):
    return T.bump_gaussian(d, prof[0], prof[1], prof[2])


# ------------------------------------------------------------------------- tables (filter.rs:302-363) ----

def test_table_read_is_exact_for_even_quadratics_and_zero_beyond_radius():
    # a 4-point Lagrange interpolant reproduces cubics; an even quadratic sampled from d = 0 with index -1 mirrored
    # is reproduced everywhere, value and derivative, including the first interval
    du, radius = 0.01, 0.5
    f = 1.5 - 3.0 * (np.arange(60) * du) ** 2
    for d in (0.0, 0.3 * du, du, 1.7 * du, 0.1234, 0.4999, radius):
        v, dv = O.table_read(f, du, radius, d)
        assert abs(v - (1.5 - 3.0 * d * d)) < 1e-13
        assert abs(dv - (-6.0 * d)) < 1e-11
    assert O.table_read(f, du, radius, radius + 1e-9) == (0.0, 0.0)
    # a cubic term in d is odd: the mirrored first interval does not reproduce it (the table holds even functions)
    g = (np.arange(60) * du) ** 3
    assert abs(O.table_read(g, du, radius, 0.5 * du)[0] - (0.5 * du) ** 3) > 1e-9


def test_table_read_of_the_synthetic_profiles_matches_the_closed_form():
    sp = T.oxygen_like(H)
    rng = np.random.default_rng(1)
    for d in rng.uniform(0, 1.7, 200):
        v, _ = O.table_read(sp.v_sr_u, sp.du, sp.r_sr, d)
        assert abs(v - float(_closed(d, T.O_V_SR))) < 1e-9 * T.O_V_SR[0]


def test_points_within_is_the_brute_force_ball():
    c = np.array([0.31, -0.47, 1.9])
    x, y, z = O.grid_points(N, H)
    want = np.nonzero((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2 <= 1.3 ** 2)[0]
    got = O.points_within(N, H, c, 1.3)
    assert np.array_equal(got, want)
    assert O.points_within(N, H, [50.0, 0.0, 0.0], 1.0).size == 0  # a sphere that misses the box


# ------------------------------------------------------------- local potential and core (nuclear.rs) ----

def test_local_only_pseudo_ion_is_the_all_electron_nucleus():
    # V_SR = 0, Z_v = Z, sigma_LR = sigma_AE: the smeared-nucleus potential of the all-electron path
    pos, z = WATER
    sp = [T.local_only(8, 8.0, H, sigma_lr=H)]
    v = O.ionic_local_potential(N, H, pos, z, sp, H)
    assert np.abs(v - O.nuclear_potential(N, H, pos, z, H)).max() < 1e-13 * np.abs(v).max()


def test_local_potential_is_lr_erf_plus_the_sr_table():
    pos, z = WATER
    sp = T.oxygen_like(H)
    v = O.ionic_local_potential(N, H, pos, z, [sp], H)
    x, y, zz = O.grid_points(N, H)
    want = np.zeros(N ** 3)
    from scipy.special import erf
    for p, q, s in ((pos[0], 6.0, sp.sigma_lr), (pos[1], 1.0, H), (pos[2], 1.0, H)):
        d = np.sqrt((x - p[0]) ** 2 + (y - p[1]) ** 2 + (zz - p[2]) ** 2)
        want += -q * erf(d / (s * math.sqrt(2))) / d
    d0 = np.sqrt((x - pos[0][0]) ** 2 + (y - pos[0][1]) ** 2 + (zz - pos[0][2]) ** 2)
    want += _closed(d0, T.O_V_SR)
    assert np.abs(v - want).max() < 1e-8


def test_partial_core_is_the_profile_and_absent_without_a_core():
    pos, z = WATER
    pc, has = O.partial_core_field(N, H, pos, z, [T.oxygen_like(H)])
    x, y, zz = O.grid_points(N, H)
    d0 = np.sqrt((x - pos[0][0]) ** 2 + (y - pos[0][1]) ** 2 + (zz - pos[0][2]) ** 2)
    assert has and np.abs(pc - _closed(d0, T.O_PC)).max() < 1e-9
    pc2, has2 = O.partial_core_field(N, H, pos, z, [T.oxygen_like(H, partial_core=False)])
    assert not has2 and not pc2.any()


def test_ionic_repulsion_uses_valence_charges():
    pos, z = WATER
    e, st, _ = O.ionic_repulsion(pos, z, [T.oxygen_like(H)])
    p = np.asarray(pos)
    want = (6 * 1 / np.linalg.norm(p[0] - p[1]) + 6 * 1 / np.linalg.norm(p[0] - p[2])
            + 1 / np.linalg.norm(p[1] - p[2]))
    assert st == O.OK and abs(e - want) < 1e-13
    assert O.ionic_repulsion([[0, 0, 0], [0, 0, 5e-9]], [8, 1], [T.oxygen_like(H)])[1] == O.COINCIDENT


# ------------------------------------------------------------------------ nonlocal (nonlocal.rs:81-174) ----

def _beta_closed(n, h, c, l, prof):
    """beta_lm = t(d) sqrt((2l+1)/4pi) R_lm(d), R_lm Racah-normalised (R00 = 1; R1-1, R10, R11 = y, z, x)."""
    x, y, z = O.grid_points(n, h)
    d = np.stack([x - c[0], y - c[1], z - c[2]])
    r = np.sqrt((d * d).sum(0))
    t = _closed(r, prof) * math.sqrt((2 * l + 1) / (4 * math.pi))
    if l == 0:
        return [t]
    return [t * d[1], t * d[2], t * d[0]]


def test_projections_match_closed_form_projectors():
    pos, z = WATER
    nl = O.Nonlocal(N, H, pos, z, [T.oxygen_like(H)])
    assert nl.spheres == 2 and nl.rows == 4
    psi = np.random.default_rng(3).standard_normal(N ** 3)
    c = nl.project(psi)
    betas = _beta_closed(N, H, pos[0], 0, T.O_S) + _beta_closed(N, H, pos[0], 1, T.O_P)
    want = np.array([H ** 3 * np.dot(b, psi) for b in betas])
    assert np.abs(c - want).max() < 1e-9 * np.abs(want).max()


def test_nonlocal_is_symmetric_low_rank_and_positive():
    pos, z = WATER
    nl = O.Nonlocal(N, H, pos, z, [T.oxygen_like(H)])
    rng = np.random.default_rng(4)
    a, b = rng.standard_normal(N ** 3), rng.standard_normal(N ** 3)
    va = nl.apply_add(a, np.zeros(N ** 3))
    vb = nl.apply_add(b, np.zeros(N ** 3))
    assert abs(np.dot(a, vb) - np.dot(va, b)) < 1e-12 * np.abs(va).max() * np.abs(b).sum()
    assert np.dot(a, va) > 0  # both E_KB > 0
    # V_NL maps into the span of the 4 projectors: V a is a combination of the closed-form beta_lm
    betas = np.stack(_beta_closed(N, H, pos[0], 0, T.O_S) + _beta_closed(N, H, pos[0], 1, T.O_P))
    coef, *_ = np.linalg.lstsq(betas.T, va, rcond=None)
    assert np.abs(betas.T @ coef - va).max() < 1e-8 * np.abs(va).max()
    # and the coefficients are E_KB c_lm
    c = nl.project(a)
    e = np.array([T.O_S[3], T.O_P[3], T.O_P[3], T.O_P[3]])
    assert np.abs(coef - e * c).max() < 1e-7 * np.abs(e * c).max()


def test_nonlocal_energy_is_weighted_sum_of_squares():
    pos, z = WATER
    nl = O.Nonlocal(N, H, pos, z, [T.oxygen_like(H)])
    orbs = np.random.default_rng(5).standard_normal((3, N ** 3))
    f = np.array([1.0, 0.0, 0.25])
    e = np.array([T.O_S[3], T.O_P[3], T.O_P[3], T.O_P[3]])
    want = sum(fi * np.dot(e, nl.project(o) ** 2) for fi, o in zip(f, orbs))
    assert abs(nl.energy(orbs, f) - want) < 1e-12 * abs(want)
    assert abs(nl.energy(orbs) - sum(np.dot(e, nl.project(o) ** 2) for o in orbs)) < 1e-11


def test_pseudo_hamiltonian_is_local_plus_nonlocal():
    pos, z = WATER
    nl = O.Nonlocal(N, H, pos, z, [T.oxygen_like(H)])
    rng = np.random.default_rng(6)
    v, x = rng.uniform(-3, 0, N ** 3), rng.standard_normal(N ** 3)
    want = nl.apply_add(x, O.hamiltonian(N, H, v, x))
    assert np.abs(nl.hamiltonian(v, x) - want).max() < 1e-12 * np.abs(want).max()


def test_solid_harmonic_gradients_are_derivatives_of_solid_harmonics():
    r = np.array([0.37, -0.81, 0.52])
    g = O.solid_harmonic_gradients(r)
    eps = 1e-6
    for a in range(3):
        dp, dm = r.copy(), r.copy()
        dp[a] += eps
        dm[a] -= eps
        fd = (O.solid_harmonics(dp)[:9] - O.solid_harmonics(dm)[:9]) / (2 * eps)
        assert np.abs(fd - g[:, a]).max() < 1e-8


# ------------------------------------------------------------------- force terms (forces.rs:259-350) ----

def _shifted(pos, i, a, d):
    p = [list(q) for q in pos]
    p[i][a] += d
    return p


def _frozen_state(n, h):
    rng = np.random.default_rng(11)
    x, y, z = O.grid_points(n, h)
    r2 = x * x + y * y + z * z
    up = np.exp(-r2 / 2.0) * (1 + 0.1 * rng.standard_normal(n ** 3) ** 2)
    dn = 0.7 * np.exp(-r2 / 3.0)
    orbs = np.stack([np.exp(-r2 / 2.5) * (1 + 0.2 * x), np.exp(-r2 / 2.0) * y, np.exp(-r2 / 3.0) * (x - z)])
    return up, dn, orbs, np.array([1.0, 0.6, 0.3])


@pytest.mark.parametrize("term", ["local", "nonlocal", "core"])
def test_force_terms_are_derivatives_of_their_frozen_energies(term):
    # forces.rs:494-560: each term is -dE/dR of its own energy with rho, rho_sigma, psi and f held fixed
    pos, z = WATER
    sp = [T.oxygen_like(H)]
    up, dn, orbs, occ = _frozen_state(N, H)
    st, terms = O.pseudo_force_terms(N, H, up, dn, orbs, occ, pos, z, sp, H)
    assert st == O.OK

    def energy(p):
        if term == "local":
            return H ** 3 * np.dot(up + dn, O.ionic_local_potential(N, H, p, z, sp, H))
        if term == "nonlocal":
            return O.Nonlocal(N, H, p, z, sp).energy(orbs, occ)
        pc, _ = O.partial_core_field(N, H, p, z, sp)
        e, _, _ = O.lsda(up + 0.5 * pc, dn + 0.5 * pc)
        return H ** 3 * O.neumaier(e)

    got = {"local": terms[1] + terms[2], "nonlocal": terms[3], "core": terms[4]}[term]
    delta = 1e-4
    for i, a in ((0, 0), (0, 1), (1, 0), (2, 2)):
        fd = -(energy(_shifted(pos, i, a, delta)) - energy(_shifted(pos, i, a, -delta))) / (2 * delta)
        assert abs(got[i, a] - fd) < 2e-7 + 1e-5 * abs(fd), (term, i, a, got[i, a], fd)
    if term == "local":
        assert np.abs(terms[2][1:]).max() == 0.0  # all-electron H: no V_SR
    if term in ("nonlocal", "core"):
        assert np.abs(got[1:]).max() == 0.0  # only the O carries projectors and a core


def test_ion_ion_force_term_is_coulomb():
    pos, z = WATER
    up, dn, orbs, occ = _frozen_state(N, H)
    _, terms = O.pseudo_force_terms(N, H, up, dn, orbs, occ, pos, z, [T.oxygen_like(H)], H)
    p = np.asarray(pos)
    q = np.array([6.0, 1.0, 1.0])
    for i in range(3):
        want = sum(q[i] * q[j] * (p[i] - p[j]) / np.linalg.norm(p[i] - p[j]) ** 3 for j in range(3) if j != i)
        assert np.abs(terms[0][i] - want).max() < 1e-13


# -------------------------------------------------------------------------------------- SCF (scf.rs) ----

def test_scf_local_only_species_reproduce_the_all_electron_scf():
    # H2 33^3 x 0.25 with both H as local-only pseudo-ions of Z_v = 1 at sigma_LR = sigma_AE: the all-electron bits
    pos, z = [[0.7, 0.0, 0.0], [-0.7, 0.0, 0.0]], [1, 1]
    st, out = O.scf(33, 0.25, pos, z, 1, 1, species=[T.local_only(1, 1.0, 0.25, sigma_lr=0.25)])
    assert st.status == O.OK
    assert abs(st.total - (-0.952449176828)) < 1e-10  # qm/src/scf.rs:1276
    assert abs(out["forces"][0, 0] - 0.128755968665) < 1e-9


@pytest.mark.slow
def test_scf_pseudo_water_force_matches_energy_difference():
    # HF force of the pseudised SCF vs the central difference of its energy (forces.rs:151-166: |HF - FD| well under
    # 1e-4 Ha/Bohr at delta = 0.01 with Richardson; here a symmetric stretch of the O along y at tol 1e-8)
    n, h = 25, 0.3
    sp = [T.oxygen_like(h)]
    s = O.default_settings(h, "aufbau", tol=1e-8, eig_tol=1e-8)
    base = [[0.0, 0.0, 0.0], [1.43, 1.11, 0.0], [-1.43, 1.11, 0.0]]

    def run(dy):
        p = _shifted(base, 0, 1, dy)
        st, out = O.scf(n, h, p, [8, 1, 1], 4, 4, settings=s, species=sp, want_rho=False)
        assert st.status == O.OK
        return st.total, out["forces"]

    e0, f0 = run(0.0)
    d = 0.01
    ep, _ = run(d)
    em, _ = run(-d)
    ep2, _ = run(2 * d)
    em2, _ = run(-2 * d)
    fd1 = -(ep - em) / (2 * d)
    fd2 = -(ep2 - em2) / (4 * d)
    rich = (4 * fd1 - fd2) / 3
    assert abs(f0[0, 1] - rich) < 1e-4, (f0[0, 1], rich)
