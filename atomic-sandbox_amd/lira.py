"""Python binding of liblira (ctypes), mirroring the reference's `qm_gpu::GpuSolver` interface.

`GpuSolver(n, h, precision)` wraps `lira_solver_create` (crates/qm-gpu/src/scf.rs:1083 `GpuSolver::new`);
`solve` is `solve_polled` (scf.rs:1195-1235), `forces` is `GpuSolver::forces` (forces.rs:239), `state()` returns the
converged densities, potentials and levels (scf.rs:1172, 1763-1851). Errors raise `QmError` subclasses named after
the reference's variants (crates/qm/src/error.rs:1-33). All compute runs in liblira's HIP kernels on the GPU; this
module only marshals arguments. It refuses to run without the built library.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblira.so")
_lib = None

FP64, MIXED = 0, 1
AUFBAU, FERMI_DIRAC = 0, 1
OK, COINCIDENT_NUCLEI, NOT_CONVERGED, STALLED, NUMERICAL, DEVICE, INVALID, ABORTED = range(8)

#: every symbol include/lira.h declares
EXPORTS = [
    "lira_default_settings", "lira_solver_create", "lira_solver_destroy", "lira_last_error", "lira_solve",
    "lira_forces", "lira_get_field", "lira_get_levels", "lira_get_orbitals", "lira_effective_eig_tol", "lira_reset",
    "lira_poisson", "lira_lobpcg", "lira_rayleigh_ritz", "lira_hamiltonian", "lira_laplacian", "lira_jacobi_sweep",
    "lira_xc_lsda", "lira_nuclear_potential", "lira_gram", "lira_stencil_variant_count", "lira_stencil_variant_name",
    "lira_autotune_stencil", "lira_set_stencil_variant", "lira_version", "lira_set_ionic", "lira_force_terms",
    "lira_ionic_fields", "lira_nonlocal_apply", "lira_free_atom_levels", "lira_set_witness_cap", "lira_witness_cap",
    "lira_set_density_extrapolation", "lira_get_profile", "lira_sample_density", "lira_set_graphs",
    "lira_state_summary", "lira_effective_gpu_tol", "lira_evaluate", "lira_solve_batch", "lira_solve_polled",
    "lira_inject_readback_fault", "lira_gram_noalloc", "lira_mg_vcycle", "lira_mg_level_op",
    "lira_solve_footprint_bytes", "lira_witness_cap_for_budget", "lira_unoccupied", "lira_harris",
    "lira_stream_bandwidth", "lira_rayleigh_ritz_device", "lira_set_device_rr", "lira_pseudo_generate",
    "lira_pseudo_error", "lira_pseudo_record_of", "lira_pseudo_species", "lira_pseudo_free", "lira_pseudo_generate_gated",
    "lira_pseudo_check",
]


class Profile(C.Structure):
    """lira_profile (SolveProfile, crates/qm-gpu/src/scf.rs:439-514)."""
    _fields_ = [("steps", C.c_int), ("readbacks", C.c_int), ("poisson_iterations", C.c_int),
                ("eigen_iterations", C.c_int), ("eigen_fallbacks", C.c_int), ("witness_growths", C.c_int),
                ("witness_trims", C.c_int), ("peak_block", C.c_int), ("extrapolated", C.c_int),
                ("wall", C.c_double), ("start_wall", C.c_double), ("poisson_wall", C.c_double),
                ("eigen_wall", C.c_double), ("eig_tol_used", C.c_double), ("diis_resets", C.c_int),
                ("diis_freeze_trims", C.c_int), ("frozen_steps", C.c_int), ("locked_calls", C.c_int),
                ("abandoned_in_step", C.c_int), ("abandoned_before_step", C.c_int)]


class Projector(C.Structure):
    """lira_projector (FilteredProjector, crates/qm/src/radial/filter.rs:323)."""
    _fields_ = [("l", C.c_int), ("e_kb", C.c_double), ("radius", C.c_double), ("t_u", C.c_void_p), ("n_t", C.c_int)]


class Species(C.Structure):
    """lira_species (FilteredTable, crates/qm/src/radial/filter.rs:330-338)."""
    _fields_ = [("z", C.c_int), ("z_v", C.c_double), ("sigma_lr", C.c_double), ("du", C.c_double),
                ("v_sr_u", C.c_void_p), ("n_v_sr", C.c_int), ("r_sr", C.c_double), ("rho_pc_u", C.c_void_p),
                ("n_rho_pc", C.c_int), ("pc_radius", C.c_double), ("projectors", C.c_void_p),
                ("n_projectors", C.c_int), ("r_core", C.c_double)]


class PseudoRecord(C.Structure):
    """lira_pseudo_record (qm::radial::record::GenerationRecord, crates/qm/src/radial/record.rs)."""
    _fields_ = [("z", C.c_int), ("f", C.c_double), ("local", C.c_int), ("local_l", C.c_int), ("z_v", C.c_double),
                ("n_core", C.c_int), ("n_channels", C.c_int), ("channel_l", C.c_int * 4), ("r_c", C.c_double * 4),
                ("eps", C.c_double * 4), ("t5", C.c_double * 4), ("n_projectors", C.c_int), ("proj_l", C.c_int * 3),
                ("e_kb", C.c_double * 3), ("gss", C.c_double * 3), ("r_beta", C.c_double * 3), ("r_nlc", C.c_double),
                ("r_core", C.c_double), ("r_cmp", C.c_double), ("r_pc_support", C.c_double), ("h", C.c_double),
                ("provisional", C.c_int), ("dimer", C.c_int), ("dimer_pp", C.c_double * 2), ("dimer_ae", C.c_double * 2),
                ("dimer_solves", C.c_int), ("h_req", C.c_double), ("h_eb", C.c_double)]


class QmError(RuntimeError):
    status = -1


class CoincidentNuclei(QmError):
    status = COINCIDENT_NUCLEI


class ScfNotConverged(QmError):
    status = NOT_CONVERGED


class ScfStalled(QmError):
    status = STALLED


class Numerical(QmError):
    status = NUMERICAL


class Device(QmError):
    status = DEVICE


class Invalid(QmError):
    status = INVALID


class Aborted(QmError):
    status = ABORTED


_ERRORS = {c.status: c for c in (CoincidentNuclei, ScfNotConverged, ScfStalled, Numerical, Device, Invalid, Aborted)}


class Nucleus(C.Structure):
    _fields_ = [("pos", C.c_double * 3), ("z", C.c_int)]


class System(C.Structure):
    _fields_ = [("nuclei", C.POINTER(Nucleus)), ("n_nuclei", C.c_int), ("n_up", C.c_int), ("n_dn", C.c_int)]


class Settings(C.Structure):
    _fields_ = [("sigma", C.c_double), ("tol", C.c_double), ("max_iter", C.c_int), ("mixing", C.c_double),
                ("eig_tol", C.c_double), ("diis_history", C.c_int), ("occupations", C.c_int), ("kt", C.c_double),
                ("stall_window", C.c_int)]


class Energy(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("kinetic", "external", "hartree", "xc", "nuclear", "internal", "ts",
                                           "total")]


MAX_PARTIAL, MAX_OVERLAPS, WARNINGS_BYTES, MAX_DIIS_HISTORY = 16, 16, 512, 15


class PartialLevel(C.Structure):
    """lira_partial_level (qm::PartialLevel, crates/qm/src/occupations.rs:425-437)."""
    _fields_ = [("spin", C.c_int), ("index", C.c_int), ("eps", C.c_double), ("f", C.c_double)]


class CoreOverlap(C.Structure):
    _fields_ = [("i", C.c_int), ("j", C.c_int), ("r", C.c_double)]


class ScfStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("delta_rho", C.c_double), ("energy", Energy),
                ("boundary_density", C.c_double), ("max_eigen_residual", C.c_double),
                ("poisson_residual", C.c_double), ("poisson_iterations", C.c_int), ("entropy", C.c_double),
                ("mu", C.c_double * 2), ("electrons", C.c_double * 2), ("tol_used", C.c_double),
                ("n_stored", C.c_int * 2), ("restricted", C.c_int), ("lobpcg_iterations", C.c_int),
                ("cg_iterations", C.c_int), ("readbacks", C.c_int), ("seconds_total", C.c_double),
                ("seconds_poisson", C.c_double), ("seconds_eigen", C.c_double), ("seconds_other", C.c_double),
                ("max_guard_residual", C.c_double), ("n_partial", C.c_int), ("partial", PartialLevel * MAX_PARTIAL),
                ("n_core_overlaps", C.c_int), ("core_overlaps", CoreOverlap * MAX_OVERLAPS),
                ("n_warnings", C.c_int), ("warnings", C.c_char * WARNINGS_BYTES)]

    def partial_levels(self):
        """[(spin, index, eps, f)] of the final step's partly occupied levels."""
        return [(p.spin, p.index, p.eps, p.f) for p in self.partial[:self.n_partial]]

    def overlaps(self):
        return [(o.i, o.j, o.r) for o in self.core_overlaps[:self.n_core_overlaps]]

    def warning_lines(self):
        return [w for w in self.warnings.decode(errors="replace").split("\n") if w]


UNOCC_MAX, HARRIS_MAX_LEVELS = 34, 96


class UnoccupiedResult(C.Structure):
    """lira_unoccupied_result (OrbitalView::unoccupied's record, crates/qm-gpu/src/orbitals.rs:997)."""
    _fields_ = [("count", C.c_int * 2), ("evals", (C.c_double * UNOCC_MAX) * 2),
                ("residuals", (C.c_double * UNOCC_MAX) * 2), ("occupied_error", C.c_double * 2),
                ("iterations", C.c_int * 2)]


class HarrisResult(C.Structure):
    """lira_harris_result (HarrisSolve, crates/qm-gpu/src/harris.rs:210-238)."""
    _fields_ = [("n_levels", C.c_int * 2), ("n_occ", C.c_int * 2), ("evals", (C.c_double * HARRIS_MAX_LEVELS) * 2),
                ("occ", (C.c_double * HARRIS_MAX_LEVELS) * 2), ("electrons", C.c_double * 2),
                ("iterations", C.c_int * 2), ("window_short", C.c_int * 2), ("cold", C.c_int),
                ("seconds", C.c_double)]


class Progress(C.Structure):
    """lira_progress (qm::ScfProgress, crates/qm/src/scf.rs:482-498)."""
    _fields_ = [("iteration", C.c_int), ("max_iter", C.c_int), ("delta_rho", C.c_double), ("tol", C.c_double),
                ("energy", C.c_double), ("best_delta_rho", C.c_double), ("since_gain", C.c_int), ("window", C.c_int),
                ("stall_level", C.c_double)]


class BatchJob(C.Structure):
    _fields_ = [("n", C.c_int), ("h", C.c_double), ("system", System)]


class BatchResult(C.Structure):
    _fields_ = [("status", C.c_int), ("stats", ScfStats), ("worker", C.c_int), ("start", C.c_double),
                ("end", C.c_double), ("message", C.c_char * 128)]


MAX_GRID, BATCH_MAX_WORKERS = 644, 64


class PoissonStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("outer", C.c_int), ("residual", C.c_double), ("seconds", C.c_double)]


class EigenStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("max_residual", C.c_double), ("fallbacks", C.c_int)]


ABORT_FN = C.CFUNCTYPE(C.c_int, C.c_void_p)
OBSERVE_FN = C.CFUNCTYPE(None, C.POINTER(Progress), C.c_void_p)
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def lib():
    """Loads liblira.so; raises if it has not been built (there is no fallback path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"liblira.so is not built at {LIB_PATH}: run __graft_entry__.build() (make -C "
                               f"atomic-sandbox_amd); there is no CPU fallback")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.lira_default_settings.argtypes = [C.c_double, C.POINTER(Settings)]
        L.lira_default_settings.restype = None
        L.lira_solver_create.argtypes = [C.c_int, C.c_double, C.c_int, C.POINTER(vp)]
        L.lira_solver_destroy.argtypes = [vp]
        L.lira_solver_destroy.restype = None
        L.lira_last_error.argtypes = [vp]
        L.lira_last_error.restype = C.c_char_p
        L.lira_solve.argtypes = [vp, C.POINTER(System), C.POINTER(Settings), C.c_int, ABORT_FN, vp,
                                 C.POINTER(ScfStats)]
        L.lira_solve_polled.argtypes = [vp, C.POINTER(System), C.POINTER(Settings), C.c_int, ABORT_FN, ABORT_FN,
                                        OBSERVE_FN, vp, C.POINTER(ScfStats)]
        L.lira_inject_readback_fault.argtypes = [vp, C.c_int]
        L.lira_unoccupied.argtypes = [vp, C.c_int, C.POINTER(UnoccupiedResult), vp]
        L.lira_harris.argtypes = [vp, C.POINTER(System), C.POINTER(Settings), _dp, _dp, C.c_int,
                                  C.POINTER(HarrisResult), vp, vp]
        L.lira_forces.argtypes = [vp, C.POINTER(Nucleus), C.c_int, C.c_double, _dp]
        L.lira_get_field.argtypes = [vp, C.c_int, _dp]
        L.lira_get_levels.argtypes = [vp, C.c_int, _dp, _dp, C.c_int, C.POINTER(C.c_int)]
        L.lira_get_orbitals.argtypes = [vp, C.c_int, _dp, C.c_int, C.POINTER(C.c_int)]
        L.lira_effective_eig_tol.argtypes = [vp, C.c_double]
        L.lira_effective_eig_tol.restype = C.c_double
        L.lira_reset.argtypes = [vp]
        L.lira_reset.restype = None
        L.lira_poisson.argtypes = [vp, _dp, _dp, C.c_double, C.c_int, C.POINTER(PoissonStats)]
        L.lira_lobpcg.argtypes = [vp, _dp, C.c_int, _dp, _dp, C.c_double, C.c_double, C.c_int, C.POINTER(EigenStats)]
        L.lira_rayleigh_ritz.argtypes = [C.c_int, _dp, _dp, C.c_int, _dp, _dp]
        L.lira_rayleigh_ritz_device.argtypes = [C.c_int, _dp, _dp, C.c_int, _dp, _dp, _dp]
        L.lira_set_device_rr.argtypes = [vp, C.c_int]
        L.lira_hamiltonian.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, vp, vp, vp, vp]
        L.lira_laplacian.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, C.c_double, vp, vp, vp]
        L.lira_jacobi_sweep.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, vp, vp, vp, vp, vp, vp]
        L.lira_xc_lsda.argtypes = [C.c_int, C.c_long, vp, vp, vp, vp, vp, vp]
        L.lira_nuclear_potential.argtypes = [C.c_int, C.c_int, C.c_double, C.POINTER(Nucleus), C.c_int, C.c_double,
                                             vp, vp]
        L.lira_gram.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, vp, C.c_int, vp, _dp, vp]
        L.lira_gram_noalloc.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, vp, C.c_int, vp, vp, vp]
        L.lira_mg_vcycle.argtypes = [vp, _dp, _dp]
        L.lira_stream_bandwidth.argtypes = [C.c_int, C.c_int, vp, vp, C.c_ulonglong, C.c_int, vp,
                                            C.POINTER(C.c_double)]
        L.lira_mg_level_op.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, vp, vp, vp, vp]
        L.lira_stencil_variant_count.restype = C.c_int
        L.lira_stencil_variant_name.argtypes = [C.c_int]
        L.lira_stencil_variant_name.restype = C.c_char_p
        L.lira_autotune_stencil.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double)]
        L.lira_autotune_stencil.restype = C.c_int
        L.lira_set_stencil_variant.argtypes = [C.c_int, C.c_int]
        L.lira_set_ionic.argtypes = [vp, vp, C.c_int]
        L.lira_free_atom_levels.argtypes = [C.c_int, _dp, C.POINTER(C.c_double)]
        L.lira_pseudo_generate.argtypes = [C.c_int, C.c_double, C.c_int, C.c_int, C.POINTER(vp)]
        L.lira_pseudo_generate_gated.argtypes = [C.c_int, C.c_double, C.POINTER(vp)]
        L.lira_pseudo_check.argtypes = [vp]
        L.lira_pseudo_error.argtypes = []
        L.lira_pseudo_error.restype = C.c_char_p
        L.lira_pseudo_record_of.argtypes = [vp, C.POINTER(PseudoRecord)]
        L.lira_pseudo_species.argtypes = [vp, C.POINTER(Species), vp, C.c_int]
        L.lira_pseudo_free.argtypes = [vp]
        L.lira_pseudo_free.restype = None
        L.lira_set_witness_cap.argtypes = [vp, C.c_int]
        L.lira_witness_cap.restype = C.c_int
        L.lira_witness_cap.argtypes = [vp]
        L.lira_set_density_extrapolation.argtypes = [vp, C.c_int]
        L.lira_get_profile.argtypes = [vp, C.POINTER(Profile), _dp, _dp, C.c_int, C.POINTER(C.c_int)]
        L.lira_sample_density.argtypes = [vp, _dp, C.c_int, _dp]
        L.lira_set_graphs.argtypes = [vp, C.c_int]
        L.lira_state_summary.argtypes = [vp, _dp, _dp]
        L.lira_effective_gpu_tol.argtypes = [C.c_double, C.c_int]
        L.lira_effective_gpu_tol.restype = C.c_double
        L.lira_evaluate.argtypes = [vp, C.POINTER(System), C.POINTER(Settings), C.c_int, ABORT_FN, ABORT_FN,
                                    OBSERVE_FN, vp, C.c_ulonglong, C.POINTER(ScfStats), _dp, C.POINTER(C.c_int),
                                    C.POINTER(C.c_double)]
        L.lira_solve_footprint_bytes.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.lira_solve_footprint_bytes.restype = C.c_ulonglong
        L.lira_witness_cap_for_budget.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_ulonglong]
        L.lira_witness_cap_for_budget.restype = C.c_int
        L.lira_force_terms.argtypes = [vp, C.POINTER(Nucleus), C.c_int, C.c_double, _dp]
        L.lira_ionic_fields.argtypes = [vp, C.POINTER(Nucleus), C.c_int, C.c_double, vp, vp]
        L.lira_nonlocal_apply.argtypes = [vp, _dp, C.c_int, _dp, vp, C.c_int, C.POINTER(C.c_int)]
        L.lira_solve_batch.argtypes = [C.POINTER(BatchJob), C.c_int, C.c_int, C.POINTER(Settings), C.c_int,
                                       C.POINTER(BatchResult), vp]
        L.lira_set_stencil_variant.restype = None
        L.lira_version.restype = C.c_char_p
        _lib = L
    return _lib


def _raise(status: int, handle=None, what: str = ""):
    if status == OK:
        return
    if handle is not None:
        msg = lib().lira_last_error(handle).decode()
    elif what == "lira_solver_create":  # a failed creation leaves its reason in the library's creation error
        msg = lib().lira_last_error(None).decode()
    else:
        msg = ""
    raise _ERRORS.get(status, QmError)(f"{what}: {msg}" if what else msg)


def default_settings(h: float, occupations: int = FERMI_DIRAC, **kw) -> Settings:
    """ScfSettings::for_grid (crates/qm/src/scf.rs:380-385); keyword arguments override fields."""
    s = Settings()
    lib().lira_default_settings(h, C.byref(s))
    s.occupations = occupations
    for k, v in kw.items():
        setattr(s, k, v)
    return s


EVAL_TOL_RAISED, EVAL_RETRIED_COLD, EVAL_WARM_STALLED, EVAL_EXTRAPOLATED = 1, 2, 4, 8
EVAL_DEVICE_RETRIED, EVAL_BUDGET_CAP = 16, 32


def solve_footprint_bytes(n, precision, n_up, n_dn, witness, diis_history=6):
    """Device bytes one solve holds (lira_solve_footprint_bytes)."""
    return lib().lira_solve_footprint_bytes(n, precision, n_up, n_dn, witness, diis_history)


def witness_cap_for_budget(n, precision, n_up, n_dn, budget, diis_history=6):
    """The witness cap a device memory budget leaves a solve (lira_witness_cap_for_budget)."""
    return lib().lira_witness_cap_for_budget(n, precision, n_up, n_dn, diis_history, int(budget))


def effective_gpu_tol(tol: float, n_valence: int) -> float:
    """max(tol, 2 * 5e-8 * n_valence), the SCF tolerance the GPU reaches (lira_effective_gpu_tol)."""
    return lib().lira_effective_gpu_tol(tol, n_valence)


def sigma_ladder(h: float):
    """The four nuclear smearing widths of a sigma -> 0 series: 1, 1.2, 1.4, 1.6 h (crates/validate/src/lib.rs:87)."""
    return [f * h for f in (1.0, 1.2, 1.4, 1.6)]


def extrapolate_sigma(sigmas, values) -> float:
    """Point-nucleus limit of a smeared-nucleus series (crates/qm/src/scf.rs:1122-1127): solves
    v_i = a + b s_i^2 + c s_i^3 + d s_i^4 for the four (s_i, v_i) and returns a. Host arithmetic only."""
    s = np.asarray(sigmas, dtype=np.float64)
    v = np.asarray(values, dtype=np.float64)
    if s.shape != (4,) or v.shape != (4,) or len(set(s.tolist())) != 4:
        raise Invalid("extrapolate_sigma: four distinct sigmas and four values")
    a = np.stack([np.ones(4), s ** 2, s ** 3, s ** 4], axis=1)
    return float(np.linalg.solve(a, v)[0])


def nuclei_array(pos, z):
    pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
    arr = (Nucleus * len(pos))()
    for i, (p, zz) in enumerate(zip(pos, z)):
        arr[i].pos[:] = [float(p[0]), float(p[1]), float(p[2])]
        arr[i].z = int(zz)
    return arr


@dataclass
class Molecule:
    """One all-electron molecule of a batch on its own n^3 grid at spacing h (Bohr)."""
    n: int
    h: float
    pos: np.ndarray
    z: list
    n_up: int
    n_dn: int
    name: str = ""


@dataclass
class BatchOutcome:
    status: int
    stats: ScfStats
    worker: int
    start: float
    end: float
    message: str
    forces: np.ndarray


def solve_batch(molecules, precision: int = MIXED, workers: int = 4, settings=None, forces: bool = True):
    """lira_solve_batch: a cold SCF and the forces of every molecule on the current HIP device, over `workers` host
    threads that each own a solver. settings: None (for_grid(h) per molecule) or one Settings per molecule. Returns one
    BatchOutcome per molecule in input order; a failed molecule is a status, not an exception."""
    mols = list(molecules)
    nj = len(mols)
    keep = []
    jobs = (BatchJob * max(nj, 1))()
    for i, m in enumerate(mols):
        nuc = nuclei_array(m.pos, m.z)
        keep.append(nuc)
        jobs[i].n, jobs[i].h = int(m.n), float(m.h)
        jobs[i].system = System(C.cast(nuc, C.POINTER(Nucleus)), len(nuc), int(m.n_up), int(m.n_dn))
    sets = None
    if settings is not None:
        if len(settings) != nj:
            raise Invalid("solve_batch: one Settings per molecule")
        sets = (Settings * nj)(*settings)
    res = (BatchResult * max(nj, 1))()
    off = np.concatenate([[0], np.cumsum([3 * len(np.asarray(m.z)) for m in mols])]).astype(int)
    fout = np.zeros(max(int(off[-1]), 1))
    st = lib().lira_solve_batch(jobs, nj, int(precision), C.cast(sets, C.POINTER(Settings)) if sets else None,
                                int(workers), res, fout.ctypes.data if forces else None)
    _raise(st, None, "lira_solve_batch")
    out = []
    for i in range(nj):
        r = res[i]
        out.append(BatchOutcome(r.status, r.stats, r.worker, r.start, r.end, r.message.decode(errors="replace"),
                                fout[off[i]:off[i + 1]].reshape(-1, 3).copy() if forces else None))
    return out


@dataclass
class State:
    """A converged Kohn–Sham state copied to the host (f64, qm layout)."""
    rho_up: np.ndarray
    rho_dn: np.ndarray
    v_h: np.ndarray
    evals_up: np.ndarray
    evals_dn: np.ndarray
    occ_up: np.ndarray
    occ_dn: np.ndarray
    extra: dict = field(default_factory=dict)


class GpuSolver:
    """One n^3 grid at spacing h on the current HIP device; precision FP64 or MIXED."""

    def __init__(self, n: int, h: float, precision: int = MIXED):
        handle = C.c_void_p()
        st = lib().lira_solver_create(n, h, precision, C.byref(handle))
        if st != OK:
            _raise(st, None, "lira_solver_create")
        self._h = handle
        self.n, self.h, self.precision = n, h, precision
        self.N = n ** 3

    def close(self):
        if getattr(self, "_h", None):
            lib().lira_solver_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve(self, pos, z, n_up, n_dn, settings: Settings | None = None, warm=False, abort=None) -> ScfStats:
        nuc = nuclei_array(pos, z)
        sysd = System(C.cast(nuc, C.POINTER(Nucleus)), len(nuc), n_up, n_dn)
        s = settings if settings is not None else default_settings(self.h)
        stats = ScfStats()
        cb = ABORT_FN(lambda _u: int(bool(abort()))) if abort else ABORT_FN(0)
        st = lib().lira_solve(self._h, C.byref(sysd), C.byref(s), int(warm), cb, None, C.byref(stats))
        _raise(st, self._h, "lira_solve")
        return stats

    def solve_polled(self, pos, z, n_up, n_dn, settings: Settings | None = None, warm=False, abort=None,
                     inner_abort=None, observe=None):
        """lira_solve_polled (GpuSolver::solve_polled): abort() before each step, inner_abort() once per Fermi-Dirac
        LOBPCG iteration, observe(dict) once per step. Returns (status, stats); an abandoned solve is ABORTED, not an
        exception, and parks its iterate as the next warm start."""
        nuc = nuclei_array(pos, z)
        sysd = System(C.cast(nuc, C.POINTER(Nucleus)), len(nuc), n_up, n_dn)
        s = settings if settings is not None else default_settings(self.h)
        stats = ScfStats()
        cb = ABORT_FN(lambda _u: int(bool(abort()))) if abort else ABORT_FN(0)
        icb = ABORT_FN(lambda _u: int(bool(inner_abort()))) if inner_abort else ABORT_FN(0)

        def _obs(p, _u):
            q = p.contents
            observe({k: getattr(q, k) for k, _ in Progress._fields_})
        ocb = OBSERVE_FN(_obs) if observe else OBSERVE_FN(0)
        st = lib().lira_solve_polled(self._h, C.byref(sysd), C.byref(s), int(warm), cb, icb, ocb, None,
                                     C.byref(stats))
        if st not in (OK, ABORTED, NOT_CONVERGED, STALLED):
            _raise(st, self._h, "lira_solve_polled")
        return st, stats

    def unoccupied(self, m: int, want_orbitals=False):
        """lira_unoccupied: per spin the m + 2 lowest unoccupied levels of the converged state, their residuals and
        (optionally) orbitals ((2, m + 2, n^3)). Returns (evals [2][m + 2], residuals, occupied_error, orbitals)."""
        r = UnoccupiedResult()
        ku = m + 2
        orbs = np.zeros(2 * ku * self.N) if want_orbitals else None
        _raise(lib().lira_unoccupied(self._h, int(m), C.byref(r), None if orbs is None else orbs.ctypes.data),
               self._h, "lira_unoccupied")
        ev = [np.array(r.evals[s][:r.count[s]]) for s in range(2)]
        res = [np.array(r.residuals[s][:r.count[s]]) for s in range(2)]
        return ev, res, list(r.occupied_error), None if orbs is None else orbs.reshape(2, ku, self.N)

    def harris(self, pos, z, n_up, n_dn, rho_up, rho_dn, m=0, settings: Settings | None = None, want_rho=True):
        """lira_harris: the non-self-consistent display solve in the potential of (rho_up, rho_dn). Returns
        (HarrisResult, rho_out_up, rho_out_dn)."""
        nuc = nuclei_array(pos, z)
        sysd = System(C.cast(nuc, C.POINTER(Nucleus)), len(nuc), n_up, n_dn)
        s = settings if settings is not None else default_settings(self.h)
        r = HarrisResult()
        ou = np.zeros(self.N) if want_rho else None
        od = np.zeros(self.N) if want_rho else None
        _raise(lib().lira_harris(self._h, C.byref(sysd), C.byref(s), np.ascontiguousarray(rho_up, dtype=np.float64),
                                 np.ascontiguousarray(rho_dn, dtype=np.float64), int(m), C.byref(r),
                                 None if ou is None else ou.ctypes.data, None if od is None else od.ctypes.data),
               self._h, "lira_harris")
        return r, ou, od

    def inject_readback_fault(self, after: int):
        """lira_inject_readback_fault: the readback `after` readbacks from now reports a device failure."""
        _raise(lib().lira_inject_readback_fault(self._h, int(after)), self._h, "lira_inject_readback_fault")

    def evaluate(self, pos, z, n_up, n_dn, settings: Settings | None = None, warm=False, abort=None,
                 inner_abort=None, observe=None, vram_budget=0):
        """lira_evaluate, one provider evaluation (the budget's witness cap, the floor-aware tolerance, the cold retry of
        a failed warm solve, the retry of a device failure, forces): returns (status, stats, forces (nuclei x 3),
        flags, tol_used); a failed solve is a status, not an exception, so its flags stay readable."""
        nuc = nuclei_array(pos, z)
        sysd = System(C.cast(nuc, C.POINTER(Nucleus)), len(nuc), n_up, n_dn)
        s = settings if settings is not None else default_settings(self.h)
        stats = ScfStats()
        f = np.zeros(3 * len(nuc))
        flags, tol = C.c_int(), C.c_double()
        cb = ABORT_FN(lambda _u: int(bool(abort()))) if abort else ABORT_FN(0)
        icb = ABORT_FN(lambda _u: int(bool(inner_abort()))) if inner_abort else ABORT_FN(0)

        def _obs(p, _u):
            q = p.contents
            observe({k: getattr(q, k) for k, _ in Progress._fields_})
        ocb = OBSERVE_FN(_obs) if observe else OBSERVE_FN(0)
        st = lib().lira_evaluate(self._h, C.byref(sysd), C.byref(s), int(warm), cb, icb, ocb, None, int(vram_budget),
                                 C.byref(stats), f, C.byref(flags), C.byref(tol))
        return st, stats, f.reshape(-1, 3), flags.value, tol.value

    def set_ionic(self, species):
        """lira_set_ionic: species are objects with the attributes of synthetic_tables.Species (z, z_v, sigma_lr,
        du, v_sr_u, r_sr, rho_pc_u, pc_radius, projectors[l, e_kb, radius, t_u]); an empty list clears the setup."""
        keep = []
        arr = (Species * max(1, len(species)))()
        for i, sp in enumerate(species):
            projs = (Projector * max(1, len(sp.projectors)))()
            for j, pr in enumerate(sp.projectors):
                t = np.ascontiguousarray(pr.t_u, dtype=np.float64)
                keep.append(t)
                projs[j] = Projector(pr.l, pr.e_kb, pr.radius, t.ctypes.data, t.size)
            keep.append(projs)
            v = np.ascontiguousarray(sp.v_sr_u, dtype=np.float64)
            keep.append(v)
            pc = None if sp.rho_pc_u is None else np.ascontiguousarray(sp.rho_pc_u, dtype=np.float64)
            if pc is not None:
                keep.append(pc)
            arr[i] = Species(sp.z, sp.z_v, sp.sigma_lr, sp.du, v.ctypes.data, v.size, sp.r_sr,
                             None if pc is None else pc.ctypes.data, 0 if pc is None else pc.size,
                             sp.pc_radius if pc is not None else 0.0, C.cast(projs, C.c_void_p).value,
                             len(sp.projectors), float(getattr(sp, "r_core", 0.0)))
        _raise(lib().lira_set_ionic(self._h, C.cast(arr, C.c_void_p), len(species)), self._h, "lira_set_ionic")

    def force_terms(self, pos, z, sigma) -> np.ndarray:
        """[ion-ion, local LR, local SR, nonlocal, partial core] x nucleus x 3 (lira_force_terms)."""
        nuc = nuclei_array(pos, z)
        out = np.zeros(15 * len(nuc))
        _raise(lib().lira_force_terms(self._h, nuc, len(nuc), sigma, out), self._h, "lira_force_terms")
        return out.reshape(5, len(nuc), 3)

    def ionic_fields(self, pos, z, sigma):
        """(V_loc, rho_pc) of a geometry under the current ionic setup (lira_ionic_fields)."""
        nuc = nuclei_array(pos, z)
        v, pc = np.zeros(self.N), np.zeros(self.N)
        _raise(lib().lira_ionic_fields(self._h, nuc, len(nuc), sigma, v.ctypes.data_as(C.c_void_p),
                                       pc.ctypes.data_as(C.c_void_p)), self._h, "lira_ionic_fields")
        return v, pc

    def nonlocal_apply(self, x):
        """(V_NL x, projections c[rows][band]) for host vectors x (nb x n^3) (lira_nonlocal_apply)."""
        x = np.ascontiguousarray(np.atleast_2d(x), dtype=np.float64)
        nb = x.shape[0]
        out = np.zeros_like(x)
        c = np.zeros(64 * nb)
        rows = C.c_int()
        _raise(lib().lira_nonlocal_apply(self._h, x.reshape(-1), nb, out.reshape(-1), c.ctypes.data_as(C.c_void_p),
                                         c.size, C.byref(rows)), self._h, "lira_nonlocal_apply")
        return out, c[:rows.value * nb].reshape(rows.value, nb)

    def set_witness_cap(self, cap: int):
        _raise(lib().lira_set_witness_cap(self._h, cap), self._h, "lira_set_witness_cap")

    def witness_cap(self) -> int:
        return lib().lira_witness_cap(self._h)

    def set_density_extrapolation(self, on: bool):
        _raise(lib().lira_set_density_extrapolation(self._h, int(bool(on))), self._h, "lira_set_density_extrapolation")

    def profile(self):
        """(Profile, delta_rho per step, eigen tolerance per step) of the last solve (lira_get_profile)."""
        p = Profile()
        dr, et = np.zeros(1024), np.zeros(1024)
        cnt = C.c_int()
        _raise(lib().lira_get_profile(self._h, C.byref(p), dr, et, 1024, C.byref(cnt)), self._h, "lira_get_profile")
        k = min(cnt.value, 1024)
        return p, dr[:k].copy(), et[:k].copy()

    def sample_density(self, points) -> np.ndarray:
        """rho_up + rho_dn of the converged state at points (count x 3, Bohr) (lira_sample_density)."""
        pts = np.ascontiguousarray(np.asarray(points, dtype=np.float64).reshape(-1, 3))
        out = np.zeros(len(pts))
        _raise(lib().lira_sample_density(self._h, pts.reshape(-1), len(pts), out), self._h, "lira_sample_density")
        return out

    def state_summary(self):
        """(integral of |rho_up - rho_dn|, max of rho) of the converged state (lira_state_summary)."""
        a, m = np.zeros(1), np.zeros(1)
        _raise(lib().lira_state_summary(self._h, a, m), self._h, "lira_state_summary")
        return float(a[0]), float(m[0])

    def set_graphs(self, on):
        """Replay the fixed kernel sequences as HIP graphs (on) or launch kernels one by one (off, the default)
        (lira_set_graphs)."""
        _raise(lib().lira_set_graphs(self._h, int(bool(on))), self._h, "lira_set_graphs")

    def set_device_rr(self, on):
        """LOBPCG's Rayleigh-Ritz step on the device from each eigensolve's second iteration on (on, the default) or
        on the host (off) (lira_set_device_rr)."""
        _raise(lib().lira_set_device_rr(self._h, int(bool(on))), self._h, "lira_set_device_rr")

    def forces(self, pos, z, sigma) -> np.ndarray:
        nuc = nuclei_array(pos, z)
        out = np.zeros(3 * len(nuc))
        _raise(lib().lira_forces(self._h, nuc, len(nuc), sigma, out), self._h, "lira_forces")
        return out.reshape(-1, 3)

    def field(self, which: int) -> np.ndarray:
        out = np.zeros(self.N)
        _raise(lib().lira_get_field(self._h, which, out), self._h, "lira_get_field")
        return out

    def levels(self, spin: int):
        ev, oc = np.zeros(256), np.zeros(256)
        cnt = C.c_int()
        _raise(lib().lira_get_levels(self._h, spin, ev, oc, 256, C.byref(cnt)), self._h, "lira_get_levels")
        return ev[:cnt.value].copy(), oc[:cnt.value].copy()

    def orbitals(self, spin: int, cap: int = 64) -> np.ndarray:
        out = np.zeros(cap * self.N)
        cnt = C.c_int()
        _raise(lib().lira_get_orbitals(self._h, spin, out, cap, C.byref(cnt)), self._h, "lira_get_orbitals")
        return out[:min(cap, cnt.value) * self.N].reshape(-1, self.N)

    def state(self) -> State:
        eu, ou = self.levels(0)
        ed, od = self.levels(1)
        return State(self.field(0), self.field(1), self.field(3), eu, ed, ou, od)

    def effective_eig_tol(self, requested: float) -> float:
        return lib().lira_effective_eig_tol(self._h, requested)

    def reset(self):
        lib().lira_reset(self._h)

    def poisson(self, rho, v0=None, tol=1e-10, max_iter=5000):
        v = np.zeros(self.N) if v0 is None else np.ascontiguousarray(v0, dtype=np.float64).copy()
        st = PoissonStats()
        _raise(lib().lira_poisson(self._h, np.ascontiguousarray(rho, dtype=np.float64), v, tol, max_iter,
                                  C.byref(st)), self._h, "lira_poisson")
        return v, st

    def mg_vcycle(self, r):
        """z = M r, one V-cycle of the Poisson preconditioner (lira_mg_vcycle)."""
        z = np.zeros(self.N)
        _raise(lib().lira_mg_vcycle(self._h, np.ascontiguousarray(r, dtype=np.float64), z), self._h, "lira_mg_vcycle")
        return z

    def lobpcg(self, v, x0, tol, max_iter=500, precond_lambda=-1.0):
        x = np.ascontiguousarray(x0, dtype=np.float64).copy()
        k = x.shape[0]
        ev = np.zeros(k)
        st = EigenStats()
        _raise(lib().lira_lobpcg(self._h, np.ascontiguousarray(v, dtype=np.float64), k, x, ev, precond_lambda, tol,
                                 max_iter, C.byref(st)), self._h, "lira_lobpcg")
        return x, ev, st


def free_atom_levels(z):
    """(1s, 2s, 3s, 2p, 3p eigenvalues, electron count) of the free LSDA atom behind the Fermi-Dirac cold guess."""
    eps = np.zeros(5)
    el = C.c_double()
    _raise(lib().lira_free_atom_levels(z, eps, C.byref(el)), None, "lira_free_atom_levels")
    return eps, el.value


PSEUDO_LOCAL_HIGHEST, PSEUDO_LOCAL_POLYNOMIAL = 0, 1
F_LADDER = (1.6, 1.4, 1.2, 1.0)


class PseudoProjector:
    """One filtered KB projector (FilteredProjector, crates/qm/src/radial/filter.rs:323): l, E_KB, radius, t_u."""

    def __init__(self, l, e_kb, radius, t_u):
        self.l, self.e_kb, self.radius, self.t_u = l, e_kb, radius, t_u


class PseudoSpecies:
    """A generated ion filtered for spacing h (qm::radial::filter::FilteredTable, filter.rs:330-338), in the form
    GpuSolver.set_ionic takes, with its generation record (`record`, a dict of lira_pseudo_record's fields) and the
    verdict of qm::ionic::check at h (`usable`, and `refusal`: the reason when it is not)."""

    def __init__(self, handle):
        st = lib().lira_pseudo_check(handle)  # ionic::check at the table's own spacing
        self.usable, self.refusal = st == 0, ("" if st == 0 else lib().lira_pseudo_error().decode())
        rec = PseudoRecord()
        _raise(lib().lira_pseudo_record_of(handle, C.byref(rec)), None, "lira_pseudo_record_of")
        sp = Species()
        projs = (Projector * 3)()
        _raise(lib().lira_pseudo_species(handle, C.byref(sp), C.cast(projs, C.c_void_p), 3), None, "lira_pseudo_species")
        copy = lambda ptr, n: np.ctypeslib.as_array(C.cast(ptr, C.POINTER(C.c_double)), shape=(n,)).copy()
        self.z, self.z_v, self.sigma_lr, self.du = sp.z, sp.z_v, sp.sigma_lr, sp.du
        self.v_sr_u, self.r_sr = copy(sp.v_sr_u, sp.n_v_sr), sp.r_sr
        self.rho_pc_u = copy(sp.rho_pc_u, sp.n_rho_pc) if sp.rho_pc_u else None
        self.pc_radius = sp.pc_radius
        self.projectors = [PseudoProjector(p.l, p.e_kb, p.radius, copy(p.t_u, p.n_t)) for p in projs[:sp.n_projectors]]
        self.r_core = sp.r_core
        nc, npj = rec.n_channels, rec.n_projectors
        self.record = {
            "z": rec.z, "f": rec.f, "local_kind": rec.local,
            "local_l": rec.local_l, "z_v": rec.z_v, "n_core": rec.n_core,
            "channels": [(rec.channel_l[k], rec.r_c[k], rec.eps[k], rec.t5[k]) for k in range(nc)],
            "projectors": [(rec.proj_l[k], rec.e_kb[k], rec.gss[k], rec.r_beta[k]) for k in range(npj)],
            "r_nlc": rec.r_nlc if rec.r_nlc > 0 else None, "r_core": rec.r_core, "r_cmp": rec.r_cmp,
            "r_pc_support": rec.r_pc_support, "h": rec.h, "provisional": bool(rec.provisional),
            "dimer": {0: None, 1: "pass", 2: "not applicable"}[rec.dimer],
            "dimer_pp": (rec.dimer_pp[0], rec.dimer_pp[1]), "dimer_ae": (rec.dimer_ae[0], rec.dimer_ae[1]),
            "dimer_solves": rec.dimer_solves,
            "h_req": None if math.isnan(rec.h_req) else rec.h_req, "h_eb": None if math.isnan(rec.h_eb) else rec.h_eb,
        }
        self.record["local"] = "polynomial" if rec.local == PSEUDO_LOCAL_POLYNOMIAL else "spdf"[rec.local_l] + "-local"


def generate_pseudo(z: int, h: float, f_index: int = -1, local: int = -1, gated: bool = False) -> PseudoSpecies:
    """qm::radial::generate(z) filtered for spacing h (lira_pseudo_generate); gated: with the dimer gate on f for
    Z <= 2 (lira_pseudo_generate_gated, solves on the current GPU). Raises Numerical with the reason when the element
    is refused."""
    handle = C.c_void_p()
    if gated:
        st = lib().lira_pseudo_generate_gated(int(z), float(h), C.byref(handle))
    else:
        st = lib().lira_pseudo_generate(int(z), float(h), int(f_index), int(local), C.byref(handle))
    if st != 0:
        why = lib().lira_pseudo_error().decode()
        raise {4: Numerical, 6: Invalid}.get(st, QmError)(f"lira_pseudo_generate: {why}")
    try:
        return PseudoSpecies(handle)
    finally:
        lib().lira_pseudo_free(handle)


def rayleigh_ritz(a, b, nk):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    m = a.shape[0]
    th, c = np.zeros(nk), np.zeros(m * nk)
    _raise(lib().lira_rayleigh_ritz(m, a, b, nk, th, c), None, "lira_rayleigh_ritz")
    return th, c.reshape(m, nk)


def rayleigh_ritz_device(a, b, nk):
    """lira_rayleigh_ritz_device: the solver's f64 single-workgroup Rayleigh-Ritz kernel on a host pair (A, B), m <= 36.
    Returns (theta, C, C_p) with C_p = C - [B11^-1 B1: C; 0]."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    m = a.shape[0]
    th, c, cp = np.zeros(nk), np.zeros(m * nk), np.zeros(m * nk)
    _raise(lib().lira_rayleigh_ritz_device(m, a, b, nk, th, c, cp), None, "lira_rayleigh_ritz_device")
    return th, c.reshape(m, nk), cp.reshape(m, nk)


def _ptr(t):
    """Device pointer of a torch tensor (torch is only the allocator here)."""
    return C.c_void_p(t.data_ptr())


def _dtype(t):
    return 4 if t.element_size() == 4 else 8


def hamiltonian(n, h, v, x, hx, nbands, stream=None):
    _raise(lib().lira_hamiltonian(_dtype(x), n, h, nbands, _ptr(v), _ptr(x), _ptr(hx), stream), None, "hamiltonian")


def laplacian(n, h, x, out, nbands, scale=1.0, stream=None):
    _raise(lib().lira_laplacian(_dtype(x), n, h, nbands, scale, _ptr(x), _ptr(out), stream), None, "laplacian")


def jacobi_sweep(n, h, v, w, r, lam, out, nbands, stream=None):
    _raise(lib().lira_jacobi_sweep(_dtype(w), n, h, nbands, _ptr(v), _ptr(w), _ptr(r), _ptr(lam), _ptr(out), stream),
           None, "jacobi_sweep")


def xc_lsda(nu, nd, e, vu, vd, stream=None):
    _raise(lib().lira_xc_lsda(_dtype(nu), nu.numel(), _ptr(nu), _ptr(nd), _ptr(e), _ptr(vu), _ptr(vd), stream), None,
           "xc_lsda")


def nuclear_potential(n, h, pos, z, sigma, out, stream=None):
    nuc = nuclei_array(pos, z)
    _raise(lib().lira_nuclear_potential(_dtype(out), n, h, nuc, len(nuc), sigma, _ptr(out), stream), None,
           "nuclear_potential")


def gram(n, h, s_block, m, u_block, q, stream=None):
    g = np.zeros(m * q)
    _raise(lib().lira_gram(_dtype(s_block), n, h, m, _ptr(s_block), q, _ptr(u_block), g, stream), None, "gram")
    return g.reshape(m, q)


def gram_noalloc(n, h, s_block, m, u_block, q, g_dev, stream=None):
    """lira_gram_noalloc: G = h^3 S^T U into the device tensor g_dev (m x q f64) with no device allocation; returns the
    status (LIRA_INVALID when the shape's partials exceed the fixed scratch array)."""
    return lib().lira_gram_noalloc(_dtype(s_block), n, h, m, _ptr(s_block), q, _ptr(u_block), _ptr(g_dev), stream)


STREAM_COPY, STREAM_READ, STREAM_WRITE = range(3)


def stream_bandwidth(kind, src, dst, nbytes, reps=20, nontemporal=False, stream=None):
    """lira_stream_bandwidth: GB/s of the hand-written streaming kernels (copy: 2 x nbytes moved per launch; read,
    write: nbytes) over device tensors src / dst, timed with HIP events on `stream`."""
    out = C.c_double(0.0)
    st = lib().lira_stream_bandwidth(kind, int(nontemporal), _ptr(src) if src is not None else None,
                                     _ptr(dst) if dst is not None else None, nbytes, reps, stream, C.byref(out))
    if st != OK:
        raise RuntimeError(f"lira_stream_bandwidth: status {st}")
    return out.value


MG_JACOBI, MG_RESIDUAL, MG_RESTRICT, MG_PROLONG_ADD = range(4)


def mg_level_op(op, n, h, d_low, d_high, x, b, out, stream=None):
    """lira_mg_level_op on torch tensors: one V-cycle level operator (Jacobi sweep, residual, restriction,
    prolongation-add)."""
    _raise(lib().lira_mg_level_op(_dtype(x), op, n, h, d_low, d_high, _ptr(x), None if b is None else _ptr(b),
                                  _ptr(out), stream), None, "mg_level_op")


def autotune_stencil(dtype_size, n, nbands):
    ms = C.c_double()
    v = lib().lira_autotune_stencil(dtype_size, n, nbands, C.byref(ms))
    return v, ms.value


def stencil_variants():
    return [lib().lira_stencil_variant_name(i).decode() for i in range(lib().lira_stencil_variant_count())]
