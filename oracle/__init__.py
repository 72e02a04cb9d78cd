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
                ("seconds_poisson", C.c_double), ("seconds_eigen", C.c_double), ("seconds_total", C.c_double),
                ("lobpcg_iterations", C.c_int), ("cg_iterations", C.c_int)]


class _Projector(C.Structure):
    _fields_ = [("l", C.c_int), ("e_kb", C.c_double), ("radius", C.c_double), ("t_u", C.c_void_p), ("n_t", C.c_int)]


class _Species(C.Structure):
    _fields_ = [("z", C.c_int), ("z_v", C.c_double), ("sigma_lr", C.c_double), ("du", C.c_double),
                ("v_sr_u", C.c_void_p), ("n_v_sr", C.c_int), ("r_sr", C.c_double), ("rho_pc_u", C.c_void_p),
                ("n_pc", C.c_int), ("pc_radius", C.c_double), ("proj", C.c_void_p), ("n_proj", C.c_int)]


class _Ionic(C.Structure):
    _fields_ = [("species", C.c_void_p), ("n_species", C.c_int), ("sigma_ae", C.c_double)]


class Ionic:
    """An or_ionic built from objects with the Species / Projector attributes of synthetic_tables (keeps the
    arrays alive while the oracle reads them)."""

    def __init__(self, species, sigma_ae):
        self._keep = []
        arr = (_Species * max(1, len(species)))()
        for i, sp in enumerate(species):
            projs = (_Projector * max(1, len(sp.projectors)))()
            for j, pr in enumerate(sp.projectors):
                t = _f64(pr.t_u)
                self._keep.append(t)
                projs[j] = _Projector(pr.l, pr.e_kb, pr.radius, t.ctypes.data, t.size)
            self._keep.append(projs)
            v = _f64(sp.v_sr_u)
            self._keep.append(v)
            pc = None if sp.rho_pc_u is None else _f64(sp.rho_pc_u)
            if pc is not None:
                self._keep.append(pc)
            arr[i] = _Species(sp.z, sp.z_v, sp.sigma_lr, sp.du, v.ctypes.data, v.size, sp.r_sr,
                              None if pc is None else pc.ctypes.data, 0 if pc is None else pc.size,
                              sp.pc_radius if pc is not None else 0.0, C.cast(projs, C.c_void_p).value,
                              len(sp.projectors))
        self._keep.append(arr)
        self.c = _Ionic(C.cast(arr, C.c_void_p).value, len(species), sigma_ae)

    @property
    def ref(self):
        return C.byref(self.c)


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
        L.or_set_threads.restype = None
        L.or_set_threads.argtypes = [C.c_int]
        L.or_threads.restype = C.c_int
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
        L.or_table_read.restype = None
        L.or_table_read.argtypes = [_dp, C.c_int, C.c_double, C.c_double, C.c_double, C.POINTER(C.c_double),
                                    C.POINTER(C.c_double)]
        L.or_points_within.restype = C.c_long
        L.or_points_within.argtypes = [C.c_int, C.c_double, _dp, C.c_double, C.c_void_p]
        L.or_ionic_local_potential.restype = None
        L.or_ionic_local_potential.argtypes = [C.c_int, C.c_double, C.c_int, _dp, _ip, C.c_void_p, _dp]
        L.or_partial_core_field.restype = C.c_int
        L.or_partial_core_field.argtypes = [C.c_int, C.c_double, C.c_int, _dp, _ip, C.c_void_p, _dp]
        L.or_ionic_repulsion.restype = C.c_int
        L.or_ionic_repulsion.argtypes = [C.c_int, _dp, _ip, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int),
                                         C.POINTER(C.c_int)]
        L.or_solid_harmonic_gradients.restype = None
        L.or_solid_harmonic_gradients.argtypes = [_dp, _dp]
        L.or_nl_build.restype = C.c_void_p
        L.or_nl_build.argtypes = [C.c_int, C.c_double, C.c_int, _dp, _ip, C.c_void_p, C.c_int]
        L.or_nl_free.restype = None
        L.or_nl_free.argtypes = [C.c_void_p]
        for f in (L.or_nl_rows, L.or_nl_spheres):
            f.restype = C.c_int
            f.argtypes = [C.c_void_p]
        L.or_nl_sphere_points.restype = C.c_long
        L.or_nl_sphere_points.argtypes = [C.c_void_p, C.c_int]
        for f in (L.or_nl_project, L.or_nl_apply_add):
            f.restype = None
            f.argtypes = [C.c_void_p, _dp, _dp]
        L.or_nl_energy.restype = C.c_double
        L.or_nl_energy.argtypes = [C.c_void_p, C.c_int, _dp, C.c_void_p]
        L.or_pseudo_hamiltonian.restype = None
        L.or_pseudo_hamiltonian.argtypes = [C.c_int, C.c_double, _dp, C.c_void_p, _dp, _dp]
        L.or_lobpcg_nl.restype = C.c_int
        L.or_lobpcg_nl.argtypes = [C.c_int, C.c_double, _dp, C.c_void_p, C.c_int, _dp, _dp, C.c_double, C.c_double,
                                   C.c_int, C.POINTER(C.c_double), _dp]
        L.or_pseudo_force_terms.restype = C.c_int
        L.or_pseudo_force_terms.argtypes = [C.c_int, C.c_double, _dp, _dp, C.c_int, _dp, _dp, C.c_int, _dp, _ip,
                                            C.c_void_p, _dp]
        L.or_scf_ionic.restype = C.c_int
        L.or_scf_ionic.argtypes = [C.c_int, C.c_double, C.c_int, _dp, _ip, C.c_int, C.c_int, C.POINTER(Settings),
                                   C.c_void_p, C.POINTER(ScfStats), C.c_int, _dp, _dp, _dp, _dp, C.c_void_p,
                                   C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.or_trilinear.restype = C.c_double
        L.or_trilinear.argtypes = [C.c_int, C.c_double, _dp, _dp]
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


def set_threads(threads: int):
    """OpenMP threads of the oracle's grid loops (1 = the reference's single-threaded solver)."""
    lib().or_set_threads(int(threads))


def threads() -> int:
    return lib().or_threads()


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


def scf(n, h, pos, z, n_up, n_dn, settings=None, cap=40, want_rho=True, want_orbitals=False, species=None):
    """Runs the oracle SCF; returns (stats, dict of arrays). species: pseudised ions (the rest all-electron)."""
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

    io = None if species is None else Ionic(species, s.sigma)
    lib().or_scf_ionic(n, h, len(zz), p, zz, n_up, n_dn, C.byref(s), None if io is None else io.ref, C.byref(st),
                       cap, ev_u, ev_d, oc_u, oc_d, ptr(ru), ptr(rd), ptr(forces), ptr(ou), ptr(od))
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


# ------------------------------------------------------------------------------------------ pseudised ions ----

def table_read(f, du, radius, d):
    f = _f64(f)
    v, dv = C.c_double(), C.c_double()
    lib().or_table_read(f, f.size, du, radius, d, C.byref(v), C.byref(dv))
    return v.value, dv.value


def points_within(n, h, c, radius):
    cnt = lib().or_points_within(n, h, _f64(c), radius, None)
    out = np.zeros(max(cnt, 1), dtype=np.int64)
    lib().or_points_within(n, h, _f64(c), radius, out.ctypes.data_as(C.c_void_p))
    return out[:cnt]


def ionic_local_potential(n, h, pos, z, species, sigma_ae):
    p, zz = _nuc(pos, z)
    io = Ionic(species, sigma_ae)
    out = np.empty(n ** 3)
    lib().or_ionic_local_potential(n, h, len(zz), p, zz, io.ref, out)
    return out


def partial_core_field(n, h, pos, z, species):
    p, zz = _nuc(pos, z)
    io = Ionic(species, 1.0)
    out = np.empty(n ** 3)
    any_core = lib().or_partial_core_field(n, h, len(zz), p, zz, io.ref, out)
    return out, bool(any_core)


def ionic_repulsion(pos, z, species):
    p, zz = _nuc(pos, z)
    io = Ionic(species, 1.0)
    e, i, j = C.c_double(), C.c_int(), C.c_int()
    st = lib().or_ionic_repulsion(len(zz), p, zz, io.ref, C.byref(e), C.byref(i), C.byref(j))
    return (e.value if st == OK else None), st, (i.value, j.value)


def solid_harmonic_gradients(r):
    out = np.empty(27)
    lib().or_solid_harmonic_gradients(_f64(r), out)
    return out.reshape(9, 3)


class Nonlocal:
    """The oracle's NonlocalOp (qm/src/nonlocal.rs:81-174) at one geometry."""

    def __init__(self, n, h, pos, z, species, with_grad=False):
        p, zz = _nuc(pos, z)
        self.n, self.h = n, h
        self._io = Ionic(species, 1.0)
        self._p = C.c_void_p(lib().or_nl_build(n, h, len(zz), p, zz, self._io.ref, int(with_grad)))

    def __del__(self):
        if getattr(self, "_p", None):
            lib().or_nl_free(self._p)
            self._p = None

    @property
    def rows(self):
        return lib().or_nl_rows(self._p)

    @property
    def spheres(self):
        return lib().or_nl_spheres(self._p)

    def sphere_points(self, s):
        return lib().or_nl_sphere_points(self._p, s)

    def project(self, psi):
        c = np.empty(max(self.rows, 1))
        lib().or_nl_project(self._p, _f64(psi), c)
        return c[:self.rows]

    def apply_add(self, psi, out):
        lib().or_nl_apply_add(self._p, _f64(psi), out)
        return out

    def energy(self, orbitals, f=None):
        orbs = _f64(orbitals).reshape(-1)
        k = orbs.size // self.n ** 3
        fa = None if f is None else _f64(f)
        return lib().or_nl_energy(self._p, k, orbs, None if fa is None else fa.ctypes.data_as(C.c_void_p))

    def hamiltonian(self, v, x):
        out = np.empty(self.n ** 3)
        lib().or_pseudo_hamiltonian(self.n, self.h, _f64(v), self._p, _f64(x), out)
        return out

    def lobpcg(self, v, x0, tol, max_iter=500, precond_lambda=None):
        x = _f64(x0).copy()
        k = x.shape[0]
        ev, res, mr = np.empty(k), np.empty(k), C.c_double()
        lam = float("nan") if precond_lambda is None else float(precond_lambda)
        it = lib().or_lobpcg_nl(self.n, self.h, _f64(v), self._p, k, x, ev, lam, tol, max_iter, C.byref(mr), res)
        return x, ev, it, mr.value, res


def pseudo_force_terms(n, h, rho_up, rho_dn, orbitals, occ, pos, z, species, sigma_ae):
    """[ion-ion, local LR, local SR, nonlocal, partial core] x nucleus x 3 (qm/src/forces.rs:259-350)."""
    p, zz = _nuc(pos, z)
    io = Ionic(species, sigma_ae)
    orbs = _f64(orbitals).reshape(-1)
    k = orbs.size // n ** 3
    terms = np.zeros(15 * len(zz))
    st = lib().or_pseudo_force_terms(n, h, _f64(rho_up), _f64(rho_dn), k, orbs, _f64(occ), len(zz), p, zz, io.ref,
                                     terms)
    return st, terms.reshape(5, len(zz), 3)


def trilinear(n, h, f, points):
    """qm::Grid::trilinear of f at each point (count x 3)."""
    f = _f64(f)
    pts = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    return np.array([lib().or_trilinear(n, h, f, _f64(p)) for p in pts])
