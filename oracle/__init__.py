"""ctypes binding of the CPU f64 oracle (liblira_oracle.so).

TEST INFRASTRUCTURE ONLY: the oracle restates the reference's CPU solver (crates/qm) and is the checker for the
HIP product path. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this module;
the product (atomic-sandbox_amd/) never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liblira_oracle.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")

OK, COINCIDENT, NOT_CONVERGED, STALLED, NUMERICAL = 0, 1, 2, 3, 4


class Settings(C.Structure):
    _fields_ = [("sigma", C.c_double), ("tol", C.c_double), ("mixing", C.c_double), ("eig_tol", C.c_double),
                ("kt", C.c_double), ("max_iter", C.c_int), ("diis_history", C.c_int), ("occupations", C.c_int),
                ("stall_window", C.c_int)]


class ScfStats(C.Structure):
    _fields_ = [("status", C.c_int), ("iterations", C.c_int), ("delta_rho", C.c_double),
                ("kinetic", C.c_double), ("external", C.c_double), ("hartree", C.c_double), ("xc", C.c_double),
                ("nuclear", C.c_double), ("internal", C.c_double), ("ts", C.c_double), ("total", C.c_double),
                ("boundary_density", C.c_double), ("max_eigen_residual", C.c_double),
                ("poisson_residual", C.c_double), ("poisson_iterations", C.c_int), ("entropy", C.c_double),
                ("mu", C.c_double * 2), ("electrons", C.c_double * 2), ("k_up", C.c_int), ("k_dn", C.c_int),
                ("seconds_poisson", C.c_double), ("seconds_eigen", C.c_double), ("seconds_total", C.c_double)]


def build() -> str:
    """Compiles the oracle with its Makefile (gcc, no fast-math)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.or_neumaier.restype = C.c_double
        L.or_neumaier.argtypes = [_dp, C.c_long]
        L.or_dot.restype = C.c_double
        L.or_dot.argtypes = [C.c_int, C.c_double, _dp, _dp]
        L.or_integrate.restype = C.c_double
        L.or_integrate.argtypes = [C.c_int, C.c_double, _dp]
        for f in (L.or_laplacian, L.or_kinetic):
            f.restype = None
            f.argtypes = [C.c_int, C.c_double, _dp, _dp]
        L.or_hamiltonian.restype = None
        L.or_hamiltonian.argtypes = [C.c_int, C.c_double, _dp, _dp, _dp]
        L.or_jacobi_omega.restype = C.c_double
        L.or_jacobi_precond.restype = None
        L.or_jacobi_precond.argtypes = [C.c_int, C.c_double, _dp, C.c_double, _dp, _dp]
        L.or_solid_harmonics.restype = None
        L.or_solid_harmonics.argtypes = [_dp, _dp]
        L.or_multipole_moments.restype = None
        L.or_multipole_moments.argtypes = [C.c_int, C.c_double, _dp, _dp]
        L.or_multipole_potential.restype = C.c_double
        L.or_multipole_potential.argtypes = [_dp, _dp]
        L.or_poisson_rhs.restype = None
        L.or_poisson_rhs.argtypes = [C.c_int, C.c_double, _dp, _dp]
        L.or_poisson_solve.restype = C.c_int
        L.or_poisson_solve.argtypes = [C.c_int, C.c_double, _dp, _dp, C.c_double, C.c_int, C.POINTER(C.c_double)]
        for f in (L.or_deposit_charge, L.or_nuclear_potential):
            f.restype = None
            f.argtypes = [C.c_int, C.c_double, C.c_int, _dp, _ip, C.c_double, _dp]
        L.or_nuclear_repulsion.restype = C.c_int
        L.or_nuclear_repulsion.argtypes = [C.c_int, _dp, _ip, C.POINTER(C.c_double), C.POINTER(C.c_int),
                                           C.POINTER(C.c_int)]
        L.or_exc_point.restype = None
        L.or_exc_point.argtypes = [C.c_double, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                   C.POINTER(C.c_double)]
        L.or_lsda.restype = None
        L.or_lsda.argtypes = [C.c_long, _dp, _dp, _dp, _dp, _dp]
        L.or_pw92_g.restype = None
        L.or_pw92_g.argtypes = [C.c_double, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.or_sym_eigen.restype = None
        L.or_sym_eigen.argtypes = [C.c_int, _dp, _dp, _dp]
        L.or_lobpcg.restype = C.c_int
        L.or_lobpcg.argtypes = [C.c_int, C.c_double, _dp, C.c_int, _dp, _dp, C.c_double, C.c_double, C.c_int,
                                C.POINTER(C.c_double), _dp]
        L.or_fermi_dirac.restype = C.c_int
        L.or_fermi_dirac.argtypes = [C.c_int, _dp, C.c_int, C.c_double, _dp, C.POINTER(C.c_double),
                                     C.POINTER(C.c_double)]
        L.or_fd_entropy.restype = C.c_double
        L.or_fd_entropy.argtypes = [C.c_double]
        L.or_initial_orbitals.restype = None
        L.or_initial_orbitals.argtypes = [C.c_int, C.c_double, C.c_int, _dp, C.c_int, C.c_ulonglong, _dp]
        L.or_scf.restype = C.c_int
        L.or_scf.argtypes = [C.c_int, C.c_double, C.c_int, _dp, _ip, C.c_int, C.c_int, C.POINTER(Settings),
                             C.POINTER(ScfStats), C.c_int, _dp, _dp, _dp, _dp, C.c_void_p, C.c_void_p, C.c_void_p,
                             C.c_void_p, C.c_void_p]
        L.or_forces.restype = C.c_int
        L.or_forces.argtypes = [C.c_int, C.c_double, _dp, C.c_int, _dp, _ip, C.c_double, _dp]
        L.or_mg_apply_a2.restype = None
        L.or_mg_apply_a2.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, _dp, _dp]
        for f in (L.or_mg_restrict, L.or_mg_prolong):
            f.restype = None
            f.argtypes = [C.c_int, _dp, _dp]
        _lib = L
    return _lib


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def grid_points(n: int, h: float):
    """(x, y, z) coordinate arrays in the qm layout idx = i + n(j + n k) (qm/src/grid.rs:11-26)."""
    o = -(n - 1) * h / 2.0
    ax = o + h * np.arange(n)
    z, y, x = np.meshgrid(ax, ax, ax, indexing="ij")
    return x.ravel(), y.ravel(), z.ravel()


def laplacian(n, h, f):
    out = np.empty(n ** 3)
    lib().or_laplacian(n, h, _f64(f), out)
    return out


def hamiltonian(n, h, v, x):
    out = np.empty(n ** 3)
    lib().or_hamiltonian(n, h, _f64(v), _f64(x), out)
    return out


def jacobi_precond(n, h, v, lam, r):
    out = np.empty(n ** 3)
    lib().or_jacobi_precond(n, h, _f64(v), lam, _f64(r), out)
    return out


def dot(n, h, a, b):
    return lib().or_dot(n, h, _f64(a), _f64(b))


def integrate(n, h, a):
    return lib().or_integrate(n, h, _f64(a))


def neumaier(x):
    x = _f64(x)
    return lib().or_neumaier(x, x.size)


def solid_harmonics(r):
    out = np.empty(49)
    lib().or_solid_harmonics(_f64(r), out)
    return out


def multipole_moments(n, h, rho):
    out = np.empty(49)
    lib().or_multipole_moments(n, h, _f64(rho), out)
    return out


def multipole_potential(q, r):
    return lib().or_multipole_potential(_f64(q), _f64(r))


def poisson_rhs(n, h, rho):
    out = np.empty(n ** 3)
    lib().or_poisson_rhs(n, h, _f64(rho), out)
    return out


def poisson_solve(n, h, rho, v0=None, tol=1e-10, max_iter=5000):
    v = np.zeros(n ** 3) if v0 is None else _f64(v0).copy()
    res = C.c_double()
    it = lib().or_poisson_solve(n, h, _f64(rho), v, tol, max_iter, C.byref(res))
    return v, it, res.value


def _nuc(pos, z):
    return _f64(np.asarray(pos).reshape(-1)), np.ascontiguousarray(z, dtype=np.int32)


def nuclear_potential(n, h, pos, z, sigma):
    p, zz = _nuc(pos, z)
    out = np.empty(n ** 3)
    lib().or_nuclear_potential(n, h, len(zz), p, zz, sigma, out)
    return out


def deposit_charge(n, h, pos, z, sigma):
    p, zz = _nuc(pos, z)
    out = np.empty(n ** 3)
    lib().or_deposit_charge(n, h, len(zz), p, zz, sigma, out)
    return out


def nuclear_repulsion(pos, z):
    p, zz = _nuc(pos, z)
    e, i, j = C.c_double(), C.c_int(), C.c_int()
    st = lib().or_nuclear_repulsion(len(zz), p, zz, C.byref(e), C.byref(i), C.byref(j))
    return (e.value if st == OK else None), st, (i.value, j.value)


def exc_point(nu, nd):
    e, vu, vd = C.c_double(), C.c_double(), C.c_double()
    lib().or_exc_point(nu, nd, C.byref(e), C.byref(vu), C.byref(vd))
    return e.value, vu.value, vd.value


def lsda(nu, nd):
    nu, nd = _f64(nu), _f64(nd)
    e, vu, vd = np.empty_like(nu), np.empty_like(nu), np.empty_like(nu)
    lib().or_lsda(nu.size, nu, nd, e, vu, vd)
    return e, vu, vd


def pw92_g(rs, which):
    g, dg = C.c_double(), C.c_double()
    lib().or_pw92_g(rs, which, C.byref(g), C.byref(dg))
    return g.value, dg.value


def sym_eigen(a):
    a = _f64(a).copy()
    m = a.shape[0]
    ev, vec = np.empty(m), np.empty((m, m))
    lib().or_sym_eigen(m, a, ev, vec)
    return ev, vec


def lobpcg(n, h, v, x0, tol, max_iter=500, precond_lambda=None):
    x = _f64(x0).copy()
    k = x.shape[0]
    ev = np.empty(k)
    res = np.empty(k)
    mr = C.c_double()
    lam = float("nan") if precond_lambda is None else float(precond_lambda)
    it = lib().or_lobpcg(n, h, _f64(v), k, x, ev, lam, tol, max_iter, C.byref(mr), res)
    return x, ev, it, mr.value, res


def fermi_dirac(evals, n, kt):
    ev = _f64(evals)
    f = np.empty_like(ev)
    mu, s = C.c_double(), C.c_double()
    st = lib().or_fermi_dirac(ev.size, ev, n, kt, f, C.byref(mu), C.byref(s))
    return st, f, mu.value, s.value


def initial_orbitals(n, h, pos, k, seed):
    p, _ = _nuc(pos, [0] * (len(np.asarray(pos).reshape(-1)) // 3))
    out = np.empty((k, n ** 3))
    lib().or_initial_orbitals(n, h, len(p) // 3, p, k, seed, out)
    return out


def default_settings(h, occupations="aufbau", tol=1e-6, eig_tol=1e-6, max_iter=200):
    """ScfSettings::for_grid (qm/src/scf.rs:375-385): sigma = h, alpha = 0.3, DIIS 6, kT = 5 mHa, stall W = 50."""
    return Settings(sigma=h, tol=tol, mixing=0.3, eig_tol=eig_tol, kt=5e-3, max_iter=max_iter, diis_history=6,
                    occupations=0 if occupations == "aufbau" else 1, stall_window=50)


def scf(n, h, pos, z, n_up, n_dn, settings=None, cap=40, want_rho=True, want_orbitals=False):
    """Runs the oracle SCF; returns (stats, dict of arrays)."""
    s = settings if settings is not None else default_settings(h)
    p, zz = _nuc(pos, z)
    st = ScfStats()
    ev_u, ev_d, oc_u, oc_d = np.zeros(cap), np.zeros(cap), np.zeros(cap), np.zeros(cap)
    N = n ** 3
    ru = np.zeros(N) if want_rho else None
    rd = np.zeros(N) if want_rho else None
    forces = np.zeros(3 * len(zz))
    ou = np.zeros(cap * N) if want_orbitals else None
    od = np.zeros(cap * N) if want_orbitals else None

    def ptr(a):
        return None if a is None else a.ctypes.data_as(C.c_void_p)

    lib().or_scf(n, h, len(zz), p, zz, n_up, n_dn, C.byref(s), C.byref(st), cap, ev_u, ev_d, oc_u, oc_d,
                 ptr(ru), ptr(rd), ptr(forces), ptr(ou), ptr(od))
    out = {"evals_up": ev_u[:st.k_up].copy(), "evals_dn": ev_d[:st.k_dn].copy(),
           "occ_up": oc_u[:st.k_up].copy(), "occ_dn": oc_d[:st.k_dn].copy(),
           "forces": forces.reshape(-1, 3), "rho_up": ru, "rho_dn": rd}
    if want_orbitals:
        out["orbitals_up"] = ou[:st.k_up * N].reshape(st.k_up, N)
        out["orbitals_dn"] = od[:st.k_dn * N].reshape(st.k_dn, N)
    return st, out


def forces(n, h, rho, pos, z, sigma):
    p, zz = _nuc(pos, z)
    out = np.empty(3 * len(zz))
    st = lib().or_forces(n, h, _f64(rho), len(zz), p, zz, sigma, out)
    return st, out.reshape(-1, 3)


def mg_apply_a2(n, h, d_low, d_high, x):
    out = np.empty(n ** 3)
    lib().or_mg_apply_a2(n, h, d_low, d_high, _f64(x), out)
    return out


def mg_restrict(n, f):
    out = np.empty((n // 2) ** 3)
    lib().or_mg_restrict(n, _f64(f), out)
    return out


def mg_prolong(n, c):
    out = np.empty(n ** 3)
    lib().or_mg_prolong(n, _f64(c), out)
    return out
