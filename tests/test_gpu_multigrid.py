"""GPU parity of the multigrid V-cycle itself and of the Poisson microbenchmark's workload.

The V-cycle's level operators against the oracle's f64 model (A2 with boundary-distance ghosts, full weighting,
trilinear prolongation; the reference's own checks, crates/qm-gpu/src/multigrid.rs:280-370), and the V-cycle as a
preconditioner: contraction <= 0.32 per cycle as a standalone solver at 33^3 and 65^3, M^-1 symmetric to 1e-5 and
positive (multigrid.rs:372-414). Then the Poisson workload (the 32-blob neutral density) against the oracle at 65^3
and, at its full 256^3, its relative residual recomputed in f64 on the host.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lira_mod():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    import lira
    lira.lib()
    return lira


def _field(n, h, fn):
    x, y, z = O.grid_points(n, h)
    return np.ascontiguousarray(fn(x, y, z), dtype=np.float64)


def _diag(n, d_low, d_high):
    """The diagonal of A2 times H^2: 6 minus the ghost factor (1 - 1/d) of every face a point touches."""
    i = np.arange(n)
    lo = (i == 0).astype(float)
    hi = (i == n - 1).astype(float)
    gl, gh = 1.0 - 1.0 / d_low, 1.0 - 1.0 / d_high
    per = lo * gl + hi * gh
    g = per[None, None, :] + per[None, :, None] + per[:, None, None]
    return (6.0 - g).reshape(-1)


@pytest.mark.parametrize("dtype_name", ["float64", "float32"])
@pytest.mark.parametrize("n, d_low, d_high", [(33, 1.0, 1.0), (16, 1.0, 1.5), (8, 1.5, 2.25)])
def test_level_operators_match_the_oracle(lira_mod, dtype_name, n, d_low, d_high):
    import torch
    lira = lira_mod
    h = 0.3
    dt = getattr(torch, dtype_name)
    tol = 1e-12 if dtype_name == "float64" else 2e-6
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n ** 3)
    b = rng.standard_normal(n ** 3)
    xd, bd = torch.tensor(x, dtype=dt, device="cuda"), torch.tensor(b, dtype=dt, device="cuda")
    xs, bs = xd.double().cpu().numpy(), bd.double().cpu().numpy()  # the values the kernels see
    out = torch.empty_like(xd)
    a2x = O.mg_apply_a2(n, h, d_low, d_high, xs)
    scale = np.abs(a2x).max()
    # residual b - A2 x
    lira.mg_level_op(lira.MG_RESIDUAL, n, h, d_low, d_high, xd, bd, out)
    want = bs - a2x
    assert np.abs(out.double().cpu().numpy() - want).max() <= 10 * tol * max(np.abs(want).max(), scale)
    # one damped Jacobi sweep x + (6/7) D^-1 (b - A2 x), D = diag / h^2
    lira.mg_level_op(lira.MG_JACOBI, n, h, d_low, d_high, xd, bd, out)
    want = xs + (6.0 / 7.0) * (h * h / _diag(n, d_low, d_high)) * (bs - a2x)
    assert np.abs(out.double().cpu().numpy() - want).max() <= tol * np.abs(want).max() * 10
    # full weighting onto (n/2)^3 and trilinear prolongation-add from it
    nc = n // 2
    coarse = torch.empty(nc ** 3, dtype=dt, device="cuda")
    lira.mg_level_op(lira.MG_RESTRICT, n, h, d_low, d_high, xd, None, coarse)
    want = O.mg_restrict(n, xs)
    assert np.abs(coarse.double().cpu().numpy() - want).max() <= tol * np.abs(want).max() * 10
    c = rng.standard_normal(nc ** 3)
    cd = torch.tensor(c, dtype=dt, device="cuda")
    fine = bd.clone()
    lira.mg_level_op(lira.MG_PROLONG_ADD, n, h, d_low, d_high, cd, None, fine)
    want = bs + O.mg_prolong(n, cd.double().cpu().numpy())
    assert np.abs(fine.double().cpu().numpy() - want).max() <= tol * np.abs(want).max() * 10


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("n", [33, 65])
def test_vcycle_contracts_and_is_symmetric_positive(lira_mod, precision, n):
    """x <- x + M (b - A2 x) contracts the residual by <= 0.32 per cycle (a grid-independent factor: the coarse
    correction works); <u, M v> = <v, M u> to 1e-5 of the scale and <u, M u> > 0."""
    lira = lira_mod
    h = 0.25
    s = lira.GpuSolver(n, h, precision)
    b = _field(n, h, lambda x, y, z: np.exp(-((x - 0.3) ** 2 + y * y + (z + 0.2) ** 2) / 0.5))
    x = np.zeros(n ** 3)
    norms = []
    for _ in range(8):
        r = b - O.mg_apply_a2(n, h, 1.0, 1.0, x)
        norms.append(np.linalg.norm(r))
        x = x + s.mg_vcycle(r)
    ratios = [norms[i + 1] / norms[i] for i in range(len(norms) - 1)]
    print(f"{n}^3 precision {precision}: V-cycle residual ratios {np.round(ratios, 3).tolist()}")
    assert max(ratios[1:]) < 0.32
    u = _field(n, h, lambda x, y, z: np.sin(0.7 * x + 0.3 * y - 0.5 * z) * np.exp(-(x * x + y * y + z * z) / 8.0))
    v = _field(n, h, lambda x, y, z: np.cos(x * y - z) * np.exp(-(x * x + 0.5 * y * y + z * z) / 6.0))
    mu, mv = s.mg_vcycle(u), s.mg_vcycle(v)
    d_umv, d_vmu, d_umu, d_vmv = u @ mv, v @ mu, u @ mu, v @ mv
    scale = np.sqrt((u @ u) * (v @ v)) * max(d_umu / (u @ u), d_vmv / (v @ v))
    assert abs(d_umv - d_vmu) < 1e-5 * scale
    assert d_umu > 0 and d_vmv > 0


@pytest.mark.parametrize("precision", [0, 1])
def test_poisson_neutral_blobs_65_matches_oracle(lira_mod, precision):
    """The Poisson workload's density (32 seeded blobs, net charge 0, so a zero monopole in the multipole boundary) on
    65^3 points over the full workload's box (h = 0.6 Bohr: 38.4 Bohr, as 256^3 x 0.15): V against the oracle's f64
    solution of the same discrete system."""
    import workloads
    lira = lira_mod
    n, h = 65, 0.6
    rho = workloads.neutral_blobs(n, h, seed=3)
    assert abs(rho.sum() * h ** 3) < 1e-3  # neutral to the grid's sampling of the blobs
    s = lira.GpuSolver(n, h, precision)
    v8, st8 = s.poisson(rho, tol=1e-8)
    v, st = s.poisson(rho, tol=1e-10)
    vo, _, res = O.poisson_solve(n, h, rho, tol=1e-10, max_iter=5000)
    print(f"65^3 blobs precision {precision}: to 1e-8 {st8.iterations} PCG iterations, residual {st8.residual:.2e}, "
          f"|V - V_oracle| {np.abs(v8 - vo).max():.2e}; to 1e-10 residual {st.residual:.2e}, "
          f"|V - V_oracle| {np.abs(v - vo).max():.2e}")
    assert st8.residual <= 1e-8 and st.residual <= 1e-10
    # both solve the same discrete system: the solutions agree to the system's conditioning times the residual
    assert np.abs(v - vo).max() < 5e-8
    assert np.abs(v8 - vo).max() < 1e-5


def test_poisson_neutral_blobs_256_relative_residual(lira_mod):
    """The full workload (256^3, mixed precision, to 1e-8): the relative residual ||b - A V|| / ||b|| recomputed in
    f64 on the host with the oracle's right-hand side (4 pi rho plus the multipole boundary terms) and 8th-order
    Laplacian."""
    import workloads
    lira = lira_mod
    n, h = 256, 0.15
    rho = workloads.neutral_blobs(n, h, seed=3)
    s = lira.GpuSolver(n, h, lira.MIXED)
    v, st = s.poisson(rho, tol=1e-8)
    b = O.poisson_rhs(n, h, rho)
    r = b + O.laplacian(n, h, v)
    rel = np.linalg.norm(r) / np.linalg.norm(b)
    print(f"256^3 blobs: {st.iterations} PCG iterations, {st.outer} corrections, {st.seconds * 1e3:.1f} ms, solver "
          f"residual {st.residual:.2e}, host f64 residual {rel:.2e}")
    assert rel <= 1e-8
