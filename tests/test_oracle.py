"""Pins the CPU f64 oracle (oracle/) before it is trusted as the checker of the HIP path.

Every check is either a value the reference's own tests hold (file:line cited), a closed form, or an identity that
a plausible mistake (a dropped term, a wrong sign or index) breaks. None retypes the oracle's formula.
"""
import math
import os

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


# ---------------------------------------------------------------- stencil (qm/src/stencil.rs:56-104) ----

def _gauss_lap_err(n, h):
    x, y, z = O.grid_points(n, h)
    r2 = x * x + y * y + z * z
    f = np.exp(-r2 / 2)
    exact = (r2 - 3.0) * f
    out = O.laplacian(n, h, f).reshape(n, n, n)[4:-4, 4:-4, 4:-4]
    return np.abs(out - exact.reshape(n, n, n)[4:-4, 4:-4, 4:-4]).max()


def test_laplacian_of_gaussian_closed_form():
    # qm/src/stencil.rs:88-90: lap exp(-r^2/2) = (r^2 - 3) exp(-r^2/2) to 1e-4 at 49^3 x 0.25
    assert _gauss_lap_err(49, 0.25) < 1e-4


def test_laplacian_converges_eighth_order():
    # qm/src/stencil.rs:92-97: h -> h/2 reduces the error by > 64 (8th order predicts 256)
    ratio = _gauss_lap_err(25, 0.5) / _gauss_lap_err(49, 0.25)
    assert ratio > 64, ratio


def test_laplacian_annihilates_constants_away_from_walls_and_is_symmetric():
    n, h = 13, 0.3
    lap = O.laplacian(n, h, np.ones(n ** 3)).reshape(n, n, n)
    assert np.abs(lap[4:-4, 4:-4, 4:-4]).max() < 1e-12
    rng = np.random.default_rng(3)
    a, b = rng.standard_normal(n ** 3), rng.standard_normal(n ** 3)
    # L_int is symmetric (qm/src/poisson.rs:19-22): <a, L b> = <L a, b>
    assert abs(a @ O.laplacian(n, h, b) - O.laplacian(n, h, a) @ b) < 1e-9 * np.abs(a @ O.laplacian(n, h, b))


def test_hamiltonian_is_kinetic_plus_potential():
    n, h = 11, 0.4
    rng = np.random.default_rng(0)
    x, v = rng.standard_normal(n ** 3), rng.uniform(-5, 0, n ** 3)
    np.testing.assert_allclose(O.hamiltonian(n, h, v, x), -0.5 * O.laplacian(n, h, x) + v * x, rtol=0, atol=1e-12)


def test_jacobi_omega_value():
    # qm/src/eigen.rs:85-87: omega = 4 / (3 rho), rho = 2.284 -> 0.584
    assert abs(O.lib().or_jacobi_omega() - 0.5837) < 1e-3


# ------------------------------------------------------------- sums (qm/src/grid.rs:118-122) ----

def test_neumaier_is_compensated():
    assert O.neumaier([1e16, 1.0, -1e16]) == 1.0


def test_integrate_gaussian():
    # qm/src/grid.rs:110-116: int exp(-r^2) = pi^{3/2} to 1e-9 at 49^3 x 0.25
    n, h = 49, 0.25
    x, y, z = O.grid_points(n, h)
    assert abs(O.integrate(n, h, np.exp(-(x * x + y * y + z * z))) - math.pi ** 1.5) < 1e-9


# --------------------------------------------------------- Poisson (qm/src/poisson.rs:262-440) ----

def test_solid_harmonics_closed_forms():
    x, y, z = 0.7, -0.4, 1.1
    r2 = x * x + y * y + z * z
    h = O.solid_harmonics([x, y, z])
    s3 = math.sqrt(3)
    idx = lambda l, m: l * l + m + l
    expect = {(0, 0): 1.0, (1, 0): z, (1, 1): x, (1, -1): y, (2, 0): (3 * z * z - r2) / 2, (2, 1): s3 * x * z,
              (2, -1): s3 * y * z, (2, 2): s3 * (x * x - y * y) / 2, (2, -2): s3 * x * y}
    for (l, m), v in expect.items():
        assert abs(h[idx(l, m)] - v) < 1e-14


def test_solid_harmonics_addition_theorem():
    # qm/src/poisson.rs:344-367: partial sums converge to 1/|a-b| inside the Legendre tail bound
    a, b = np.array([0.3, -0.2, 0.5]), np.array([2.0, 1.5, -1.0])
    ha, hb = O.solid_harmonics(a), O.solid_harmonics(b)
    na, nb = np.linalg.norm(a), np.linalg.norm(b)
    exact = 1.0 / np.linalg.norm(a - b)
    partial, prev = 0.0, np.inf
    for l in range(7):
        for m in range(-l, l + 1):
            i = l * l + m + l
            partial += ha[i] * hb[i] / nb ** (2 * l + 1)
        err = abs(partial - exact)
        assert err <= (na / nb) ** (l + 1) / (nb - na)
        assert err < prev
        prev = err


def _gaussian(n, h, q, sigma, c):
    x, y, z = O.grid_points(n, h)
    r2 = (x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2
    rho = q / (2 * math.pi * sigma * sigma) ** 1.5 * np.exp(-r2 / (2 * sigma * sigma))
    r = np.sqrt(r2)
    from scipy.special import erf
    with np.errstate(divide="ignore", invalid="ignore"):
        exact = np.where(r < 1e-12, q * math.sqrt(2 / math.pi) / sigma, q * erf(r / (sigma * math.sqrt(2))) / r)
    return rho, exact


def test_moments_of_spherical_charge_are_harmonics_at_centre():
    # qm/src/poisson.rs:397-407 (mean-value property), to 1e-9
    d = [1.3, -0.7, 0.4]
    rho, _ = _gaussian(49, 0.25, 1.0, 0.5, d)
    q = O.multipole_moments(49, 0.25, rho)
    np.testing.assert_allclose(q, O.solid_harmonics(d), rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("centre", [[0.0, 0.0, 0.0], [1.3, -0.7, 0.4]])
def test_poisson_gaussian_potential(centre):
    # qm/src/poisson.rs:299-307: |V - q erf(r/(sigma sqrt2))/r| < 1e-4 at 49^3 x 0.25, centred and off-centre
    rho, exact = _gaussian(49, 0.25, 1.0, 0.5, centre)
    v, it, res = O.poisson_solve(49, 0.25, rho, tol=1e-10, max_iter=2000)
    assert res < 1e-10
    assert np.abs(v - exact).max() < 1e-4


def test_hartree_energy_of_gaussian():
    # qm/src/poisson.rs:316-323: E_H = q^2 / (2 sigma sqrt(pi)) = 0.5641896 at sigma = 0.5
    rho, _ = _gaussian(49, 0.25, 1.0, 0.5, [0, 0, 0])
    v, _, _ = O.poisson_solve(49, 0.25, rho, tol=1e-10, max_iter=2000)
    eh = 0.5 * O.dot(49, 0.25, rho, v)
    assert abs(eh - 1.0 / (2 * 0.5 * math.sqrt(math.pi))) < 1e-4


# ------------------------------------------------------------- nuclei (qm/src/nuclear.rs:381-) ----

def test_nuclear_potential_on_site_and_far_field():
    n, h, sigma = 9, 0.5, 0.5
    v = O.nuclear_potential(n, h, [[0, 0, 0]], [2], sigma)
    centre = v.reshape(n, n, n)[4, 4, 4]
    assert abs(centre + 2 * math.sqrt(2 / math.pi) / sigma) < 1e-14
    corner = v.reshape(n, n, n)[0, 0, 0]
    d = math.sqrt(3) * 2.0
    assert abs(corner + 2 * math.erf(d / (sigma * math.sqrt(2))) / d) < 1e-14


def test_nuclear_potential_equals_minus_poisson_of_deposit():
    # qm/src/nuclear.rs:36-41 and the sigma table at :59-65: sigma = 0.8, h = 0.25, Z = 2 -> 1.49e-6
    n, h, sigma = 49, 0.25, 0.8
    rho = O.deposit_charge(n, h, [[0, 0, 0]], [2], sigma)
    phi, _, res = O.poisson_solve(n, h, rho, tol=1e-10, max_iter=3000)
    v = O.nuclear_potential(n, h, [[0, 0, 0]], [2], sigma)
    assert np.abs(v + phi).max() < 3e-6


def test_nuclear_repulsion_and_coincidence():
    e, st, _ = O.nuclear_repulsion([[0, 0, 0], [0, 0, 2.0]], [1, 3])
    assert st == O.OK and abs(e - 1.5) < 1e-15
    _, st, pair = O.nuclear_repulsion([[0, 0, 0], [1, 0, 0], [1, 0, 1e-9]], [1, 1, 1])
    assert st == O.COINCIDENT and pair == (1, 2)


# ------------------------------------------------------------------- XC (qm/src/xc.rs:214-300) ----

def test_pw92_reference_values():
    for rs, eps in [(2.0, -0.044760), (10.0, -0.018572)]:
        n = 3 / (4 * math.pi * rs ** 3)
        e, _, _ = O.exc_point(n / 2, n / 2)
        ex = -(3 / 4) * (3 / math.pi) ** (1 / 3) * n ** (4 / 3)
        assert abs((e - ex) / n - eps) < 2e-5


def test_dirac_exchange_identity():
    n = 0.1
    e, vu, vd = O.exc_point(n / 2, n / 2)
    g0, _ = O.pw92_g((3 / (4 * math.pi * n)) ** (1 / 3), 0)
    ex = -(3 / 4) * (3 / math.pi) ** (1 / 3) * n ** (4 / 3)
    assert abs((e - n * g0) - ex) < 1e-12
    assert vu == vd


def test_fully_polarised_uses_ferro_fit():
    n = 0.02
    e, _, _ = O.exc_point(n, 0.0)
    rs = (3 / (4 * math.pi * n)) ** (1 / 3)
    g1, _ = O.pw92_g(rs, 1)
    ex = -(3 / 4) * (3 / math.pi) ** (1 / 3) * 2 ** (1 / 3) * n ** (4 / 3)
    assert abs((e - ex) / n - g1) < 1e-12


@pytest.mark.parametrize("nu,nd", [(0.3, 0.3), (0.05, 0.2), (0.001, 0.0004), (0.1, 0.0)])
def test_xc_potentials_are_energy_derivatives(nu, nd):
    # qm/src/xc.rs:281-292: v_sigma = d(n eps)/dn_sigma by central differences
    _, vu, vd = O.exc_point(nu, nd)
    d = 1e-6 * (nu + nd)
    fu = (O.exc_point(nu + d, nd)[0] - O.exc_point(nu - d, nd)[0]) / (2 * d)
    assert abs(vu - fu) <= 1e-7 + 1e-6 * abs(fu)
    if nd > 0:
        fd = (O.exc_point(nu, nd + d)[0] - O.exc_point(nu, nd - d)[0]) / (2 * d)
        assert abs(vd - fd) <= 1e-7 + 1e-6 * abs(fd)


def test_xc_vacuum_is_zero_not_nan():
    e, vu, vd = O.lsda(np.array([0.0, 1e-30, 0.1, -1e-3]), np.array([0.0, 0.0, 0.1, 0.0]))
    assert e[0] == 0 and vu[0] == 0 and vd[0] == 0
    assert np.all(np.isfinite(e)) and np.all(np.isfinite(vu)) and np.all(np.isfinite(vd))


# ----------------------------------------------------- dense eigen and LOBPCG (qm/src/eigen.rs) ----

def test_sym_eigen_clustered_and_spread():
    # qm/src/eigen.rs:456-475: a near-degenerate cluster and a wide spread, decomposed to rounding
    n = 30
    rng = np.random.default_rng(7)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.array([-0.8686 + 1e-12 * i if i < 3 else 70.0 * i / n for i in range(n)])
    a = q @ np.diag(lam) @ q.T
    a = 0.5 * (a + a.T)
    ev, vec = O.sym_eigen(a)
    assert np.abs(a @ vec - vec * ev).max() < 1e-10
    assert np.abs(vec.T @ vec - np.eye(n)).max() < 1e-10
    assert np.abs(ev - np.sort(lam)).max() < 1e-10


def test_lobpcg_oscillator_spectrum():
    # qm/src/eigen.rs:394-404: H = -1/2 lap + r^2/2 -> 1.5, 2.5 (x3)
    n, h = 41, 0.25
    x, y, z = O.grid_points(n, h)
    v = 0.5 * (x * x + y * y + z * z)
    guess = O.initial_orbitals(n, h, [[0.0, 0.0, 0.0]], 4, 12345)
    xs, ev, it, mr, _ = O.lobpcg(n, h, v, guess, 1e-6, 500, precond_lambda=1.5)
    assert mr < 1e-6
    assert abs(ev[0] - 1.5) < 1e-4
    assert np.abs(ev[1:] - 2.5).max() < 1e-4
    g = xs @ xs.T * h ** 3
    assert np.abs(g - np.eye(4)).max() < 1e-8


def test_preconditioner_reduces_iterations():
    n, h = 25, 0.3
    x, y, z = O.grid_points(n, h)
    v = 0.5 * (x * x + y * y + z * z)
    guess = O.initial_orbitals(n, h, [[0.0, 0.0, 0.0]], 2, 99)
    _, _, plain, _, _ = O.lobpcg(n, h, v, guess, 1e-6, 500)
    _, _, pre, _, _ = O.lobpcg(n, h, v, guess, 1e-6, 500, precond_lambda=1.5)
    assert pre < plain


# -------------------------------------------------- occupations (qm/src/occupations.rs; references.md:342) ----

def s_of(f):
    return -(f * math.log(f) + (1 - f) * math.log(1 - f))


@pytest.mark.parametrize("m,n", [(2, 1), (4, 2), (3, 1), (3, 2), (6, 4)])
def test_fd_entropy_closed_forms_of_degenerate_shells(m, n):
    # an exactly degenerate shell of m levels holding n electrons: f = n/m each, S = m s(n/m) (4kT ln2 for m=4, n=2)
    ev = np.array([-1.0] * m + [5.0])
    st, f, mu, s = O.fermi_dirac(ev, n, 5e-3)
    assert st == O.OK
    np.testing.assert_allclose(f[:m], n / m, rtol=0, atol=1e-12)
    assert abs(s - m * s_of(n / m)) < 1e-10
    assert abs(f.sum() - n) < 1e-12


def test_fd_gapped_and_refusals():
    # a gapped channel: sum f == 2 in f64 on the plateau where e^{-(0.4-mu)/kT} < 2^-52 and e^{-(mu+0.3)/kT} < 2^-53,
    # i.e. mu in about [-0.116, 0.220]; mu is that plateau's midpoint (qm/src/occupations.rs:200-204)
    st, f, mu, s = O.fermi_dirac(np.array([-0.5, -0.3, 0.4]), 2, 5e-3)
    assert st == O.OK and abs(f[0] - 1) < 1e-15 and abs(f[1] - 1) < 1e-15
    lo, hi = -0.3 + 5e-3 * 53 * math.log(2), 0.4 - 5e-3 * 52 * math.log(2)
    assert abs(mu - 0.5 * (lo + hi)) < 5e-3
    assert abs(f[2] - math.exp(-(0.4 - mu) / 5e-3)) < 1e-12 * f[2] and s < 1e-14
    assert O.fermi_dirac(np.array([-0.5, -0.3]), 2, 5e-3)[0] == O.NUMERICAL  # needs n + 1 levels
    assert O.fermi_dirac(np.array([-0.5, float("nan"), 1.0]), 1, 5e-3)[0] == O.NUMERICAL


def test_fd_entropy_formula_limits():
    assert O.lib().or_fd_entropy(float("inf")) == 0.0
    assert abs(O.lib().or_fd_entropy(0.0) - math.log(2)) < 1e-15


# -------------------------------------------------------- multigrid model (qm-gpu/src/multigrid.rs:366-371) ----

@pytest.mark.parametrize("n", [9, 10])
def test_restriction_is_prolongation_transpose_over_eight(n):
    rng = np.random.default_rng(n)
    f = rng.standard_normal(n ** 3)
    c = rng.standard_normal((n // 2) ** 3)
    assert abs(O.mg_restrict(n, f) @ c - f @ O.mg_prolong(n, c) / 8) < 1e-12


def test_a2_is_symmetric_positive():
    n = 8
    rng = np.random.default_rng(1)
    a, b = rng.standard_normal(n ** 3), rng.standard_normal(n ** 3)
    for dl, dh in [(1.0, 1.0), (1.0, 0.5), (1.5, 0.75)]:
        ab = a @ O.mg_apply_a2(n, 0.5, dl, dh, b)
        ba = b @ O.mg_apply_a2(n, 0.5, dl, dh, a)
        assert abs(ab - ba) < 1e-10 * abs(ab)
        assert a @ O.mg_apply_a2(n, 0.5, dl, dh, a) > 0


# ---------------------------------------------------------- SCF golden pins (qm/src/scf.rs:1258-1294) ----

H2 = ([[0.7, 0.0, 0.0], [-0.7, 0.0, 0.0]], [1, 1], 1, 1)
LIH = ([[-0.8, 0.1, 0.0], [2.2, 0.0, 0.0]], [3, 1], 2, 2)


def test_scf_h2_aufbau_matches_reference_bits():
    # BASE_H2_E = -0.952449176828, BASE_H2_FX = +0.128755968665, 10 SCF steps (qm/src/scf.rs:1276-1277)
    pos, z, nu, nd = H2
    st, out = O.scf(33, 0.25, pos, z, nu, nd)
    assert st.status == O.OK and st.iterations == 10
    assert abs(st.total - (-0.952449176828)) < 2e-12
    assert abs(out["forces"][0, 0] - 0.128755968665) < 1e-11
    # H2 is mirror-symmetric about x = 0: the forces are equal and opposite
    assert abs(out["forces"][0, 0] + out["forces"][1, 0]) < 1e-9
    assert abs(st.electrons[0] - 1) < 1e-9 and abs(st.electrons[1] - 1) < 1e-9


def test_scf_lih_aufbau_matches_reference_bits():
    # BASE_LIH_E = -5.219911350989, BASE_LIH_FX = -0.022680156352, 13 SCF steps (qm/src/scf.rs:1278-1279)
    pos, z, nu, nd = LIH
    st, out = O.scf(33, 0.25, pos, z, nu, nd)
    assert st.status == O.OK and st.iterations == 13
    assert abs(st.total - (-5.219911350989)) < 5e-12
    assert abs(out["forces"][0, 0] - (-0.022680156352)) < 1e-9


def test_scf_h2_fermi_dirac_matches_reference():
    # FD_H2_E = -0.952449176839 (qm/src/scf.rs:1290); the reference's FD cold guess is the free-atom density, ours
    # the sigma = 1 blobs (the same start as its plan-Task-7 pin -0.952449176832): agreement to 1e-10
    pos, z, nu, nd = H2
    st, out = O.scf(33, 0.25, pos, z, nu, nd, settings=O.default_settings(0.25, "fd"))
    assert st.status == O.OK
    assert abs(st.total - (-0.952449176839)) < 1e-10
    assert st.k_up == 2 and st.k_dn == 2 and st.ts < 1e-15


def test_scf_coincident_nuclei_is_an_error():
    st, _ = O.scf(17, 0.3, [[0, 0, 0], [0, 0, 0]], [1, 1], 1, 1)
    assert st.status == O.COINCIDENT


# ------------------------------------------------------------------ trilinear sampling (qm/src/grid.rs:54-86) ----

def test_trilinear_reproduces_linear_fields_reads_nodes_and_is_zero_outside():
    # qm/src/grid.rs:143-170: a linear field is reproduced exactly inside [origin, origin + (n-1)h]^3 (off the upper
    # faces, whose +1 neighbours read as 0), a node returns its stored value bitwise, and points outside read 0
    n, h = 9, 0.5
    x, y, z = O.grid_points(n, h)
    f = 1.0 + 2.0 * x - y + 3.0 * z
    rng = np.random.default_rng(5)
    pts = rng.uniform(-2.0, 1.99, (50, 3))
    got = O.trilinear(n, h, f, pts)
    want = 1.0 + 2.0 * pts[:, 0] - pts[:, 1] + 3.0 * pts[:, 2]
    assert np.abs(got - want).max() < 1e-12
    g = rng.standard_normal(n ** 3)
    idx = [0, 17, 300, n ** 3 - 1]
    nodes = np.stack([x[idx], y[idx], z[idx]], axis=1)
    assert np.array_equal(O.trilinear(n, h, g, nodes), g[idx])
    assert np.all(O.trilinear(n, h, f, [[2.01, 0, 0], [0, -2.0001, 0], [0, 0, 9.0]]) == 0.0)
