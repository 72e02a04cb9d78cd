"""GPU parity of liblira's kernels against the f64 oracle, called through the C ABI on device pointers.

FP64 kernels must match the oracle to rounding; FP32 kernels to the f32 storage rounding (the reference's own bounds
for its f32 kernels, crates/qm-gpu/src/stencil.rs:149-285, physics.rs:149-321). torch only allocates device memory.
"""
import math
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    import lira
    lira.lib()
    return torch


def _t(torch, a, dtype):
    return torch.tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda")


@pytest.mark.parametrize("n,h,nb", [(33, 0.25, 3), (37, 0.2, 2), (5, 0.5, 2), (70, 0.15, 1)])
def test_hamiltonian_fp64_matches_oracle(dev, n, h, nb):
    import lira
    torch = dev
    rng = np.random.default_rng(n)
    x = rng.standard_normal((nb, n ** 3))
    v = rng.uniform(-5, 0, n ** 3)
    xd, vd = _t(torch, x, torch.float64), _t(torch, v, torch.float64)
    out = torch.empty_like(xd)
    lira.hamiltonian(n, h, vd, xd, out, nb)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    for b in range(nb):
        want = O.hamiltonian(n, h, v, x[b])
        assert np.abs(got[b] - want).max() <= 1e-12 * np.abs(want).max(), b


def test_hamiltonian_fp32_kernel_only_and_end_to_end(dev):
    import lira
    torch = dev
    n, h, nb = 33, 0.25, 2
    rng = np.random.default_rng(1)
    x = rng.standard_normal((nb, n ** 3))
    v = rng.uniform(-5, 0, n ** 3)
    xd, vd = _t(torch, x, torch.float32), _t(torch, v, torch.float32)
    out = torch.empty_like(xd)
    lira.hamiltonian(n, h, vd, xd, out, nb)
    torch.cuda.synchronize()
    got = out.cpu().numpy().astype(np.float64)
    x32 = x.astype(np.float32).astype(np.float64)
    v32 = v.astype(np.float32).astype(np.float64)
    for b in range(nb):
        kernel_only = O.hamiltonian(n, h, v32, x32[b])
        end_to_end = O.hamiltonian(n, h, v, x[b])
        scale = np.abs(kernel_only).max()
        assert np.abs(got[b] - kernel_only).max() <= 1e-6 * scale
        assert np.abs(got[b] - end_to_end).max() <= 1e-5 * scale


def test_laplacian_dirichlet_closed_form_and_kinetic_scale(dev):
    import lira
    torch = dev
    n, h = 49, 0.25
    x, y, z = O.grid_points(n, h)
    r2 = x * x + y * y + z * z
    f = np.exp(-r2 / 2)
    fd = _t(torch, f, torch.float64)
    out = torch.empty_like(fd)
    lira.laplacian(n, h, fd, out, 1, 1.0)
    kin = torch.empty_like(fd)
    lira.laplacian(n, h, fd, kin, 1, -0.5)
    torch.cuda.synchronize()
    lap = out.cpu().numpy()
    np.testing.assert_allclose(lap, O.laplacian(n, h, f), rtol=0, atol=1e-12)
    inner = (lap - (r2 - 3) * f).reshape(n, n, n)[4:-4, 4:-4, 4:-4]
    assert np.abs(inner).max() < 1e-4
    # kinetic is bitwise -1/2 laplacian (crates/qm-gpu/src/stencil.rs:195-213)
    assert np.array_equal(kin.cpu().numpy(), -0.5 * lap)


@pytest.mark.parametrize("dtype_name", ["float32", "float64"])
@pytest.mark.parametrize("n,nb", [(41, 3), (44, 5), (64, 4)])
def test_every_stencil_variant_is_bitwise_identical(dev, dtype_name, n, nb):
    # n = 41 takes the scalar tiles only; n = 44 and 64 are multiples of the 16-byte vector width, so the vector-lane
    # and multi-band variants run too (44: ragged x/y tiles; nb = 5: a partial band group). Every mode (H.X, scaled
    # Laplacian, damped Jacobi) must give the same bits whichever variant the autotuner picks.
    import lira
    torch = dev
    dt = getattr(torch, dtype_name)
    h = 0.2
    rng = np.random.default_rng(7)
    xd = _t(torch, rng.standard_normal((nb, n ** 3)), dt)
    rd = _t(torch, rng.standard_normal((nb, n ** 3)), dt)
    vd = _t(torch, rng.uniform(-5, 0, n ** 3), dt)
    lam = _t(torch, rng.uniform(-1, 0, nb), torch.float64)
    size = 4 if dt == torch.float32 else 8
    outs = []
    try:
        for var in range(len(lira.stencil_variants())):
            lira.lib().lira_set_stencil_variant(size, var)
            o = [torch.empty_like(xd) for _ in range(3)]
            lira.hamiltonian(n, h, vd, xd, o[0], nb)
            lira.laplacian(n, h, xd, o[1], nb, scale=-0.5)
            lira.jacobi_sweep(n, h, vd, xd, rd, lam, o[2], nb)
            torch.cuda.synchronize()
            outs.append([t.cpu().numpy() for t in o])
    finally:
        lira.lib().lira_set_stencil_variant(size, -1)
    for var, o in enumerate(outs[1:], 1):
        for mode in range(3):
            assert np.array_equal(o[mode], outs[0][mode]), (lira.stencil_variants()[var], mode)
    if dt == torch.float64:
        x = xd.cpu().numpy()
        v = vd.cpu().numpy()
        for b in range(nb):
            want = O.hamiltonian(n, h, v, x[b])
            assert np.abs(outs[0][0][b] - want).max() < 1e-12 * np.abs(want).max()


def test_jacobi_sweep_matches_oracle_step(dev):
    import lira
    torch = dev
    n, h, nb = 25, 0.3, 2
    rng = np.random.default_rng(3)
    w = rng.standard_normal((nb, n ** 3))
    r = rng.standard_normal((nb, n ** 3))
    v = rng.uniform(-3, 1, n ** 3)
    lam = np.array([-0.7, 0.4])
    wd, rd, vd = _t(torch, w, torch.float64), _t(torch, r, torch.float64), _t(torch, v, torch.float64)
    ld = _t(torch, lam, torch.float64)
    out = torch.empty_like(wd)
    lira.jacobi_sweep(n, h, vd, wd, rd, ld, out, nb)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    omega = O.lib().or_jacobi_omega()
    diag_t = -1.5 * (-205.0 / 72.0) / (h * h)
    for b in range(nb):
        hw = O.hamiltonian(n, h, v, w[b])
        inv_d = 1.0 / np.maximum(diag_t + v - lam[b], 0.1)
        want = w[b] + omega * inv_d * (r[b] - (hw - lam[b] * w[b]))  # qm/src/eigen.rs:237
        assert np.abs(got[b] - want).max() <= 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("dtype_name,rtol", [("float64", 1e-13), ("float32", 3e-6)])
def test_xc_matches_oracle_across_densities(dev, dtype_name, rtol):
    # density sweep 1e-16 .. 3e3 and |zeta| = 1 (crates/qm-gpu/src/physics.rs:249-321)
    import lira
    torch = dev
    dt = getattr(torch, dtype_name)
    dens = np.logspace(-16, np.log10(3e3), 400)
    zeta = np.concatenate([np.linspace(-1, 1, 200), [1.0] * 100, [-1.0] * 100])
    nu, nd = dens * (1 + zeta) / 2, dens * (1 - zeta) / 2
    if dt == torch.float32:
        nu, nd = nu.astype(np.float32).astype(np.float64), nd.astype(np.float32).astype(np.float64)
    want = O.lsda(nu, nd)
    ins = [_t(torch, a, dt) for a in (nu, nd)]
    outs = [torch.empty_like(ins[0]) for _ in range(3)]
    lira.xc_lsda(ins[0], ins[1], *outs)
    torch.cuda.synchronize()
    for g, w in zip(outs, want):
        g = g.cpu().numpy().astype(np.float64)
        assert np.all(np.isfinite(g))
        assert np.all(np.abs(g - w) <= rtol * np.abs(w) + 1e-30)


def test_nuclear_potential_matches_oracle(dev):
    import lira
    torch = dev
    n, h, sigma = 33, 0.25, 0.25
    pos, z = [[0.0, 0.0, 0.0], [1.43, 1.11, 0.0], [-1.43, 1.11, 0.0]], [8, 1, 1]
    out = torch.empty(n ** 3, dtype=torch.float64, device="cuda")
    lira.nuclear_potential(n, h, pos, z, sigma, out)
    torch.cuda.synchronize()
    want = O.nuclear_potential(n, h, pos, z, sigma)
    assert np.abs(out.cpu().numpy() - want).max() <= 1e-13 * np.abs(want).max()


@pytest.mark.parametrize("m,q", [(5, 9), (24, 48), (40, 70)])
def test_gram_matches_numpy(dev, m, q):
    import lira
    torch = dev
    n, h = 21, 0.3
    rng = np.random.default_rng(m)
    s = rng.standard_normal((m, n ** 3))
    u = rng.standard_normal((q, n ** 3))
    g = lira.gram(n, h, _t(torch, s, torch.float64), m, _t(torch, u, torch.float64), q)
    want = s @ u.T * h ** 3
    assert np.abs(g - want).max() <= 1e-12 * np.abs(want).max()
    g32 = lira.gram(n, h, _t(torch, s, torch.float32), m, _t(torch, u, torch.float32), q)
    want32 = s.astype(np.float32).astype(np.float64) @ u.astype(np.float32).astype(np.float64).T * h ** 3
    assert np.abs(g32 - want32).max() <= 1e-12 * np.abs(want32).max()  # f32 inputs, f64 accumulation



@pytest.mark.parametrize("m,q", [(27, 54), (40, 70), (9, 9)])
def test_gram_with_u_beginning_with_s(dev, m, q):
    # U's first m columns are S's own memory (the LOBPCG shape U = [S, R, HW, HP]): those columns are read once
    import lira
    torch = dev
    n, h = 21, 0.3
    rng = np.random.default_rng(100 + m)
    x = rng.standard_normal((q, n ** 3))
    for dt in (torch.float64, torch.float32):
        xt = _t(torch, x, dt)
        g = lira.gram(n, h, xt, m, xt, q)
        xd = xt.double().cpu().numpy()
        want = xd[:m] @ xd.T * h ** 3
        assert np.abs(g - want).max() <= 1e-12 * np.abs(want).max(), dt


def test_stream_kernels_copy_and_fill(dev):
    """The streaming-ceiling kernels move exactly the bytes they are timed on: the copy (plain and nontemporal
    stores) leaves dst equal to src over a length that is not a multiple of a workgroup's span, the write fills every
    vector with 1.0, and each reports a positive rate."""
    import lira
    torch = dev
    count = 4 * (3 * 256 * 4 + 7)  # floats: 3 full workgroup spans plus 7 vectors
    src = torch.randn(count, device="cuda", dtype=torch.float32)
    for nt in (False, True):
        dst = torch.zeros(count + 8, device="cuda", dtype=torch.float32)
        gbs = lira.stream_bandwidth(lira.STREAM_COPY, src, dst, count * 4, reps=2, nontemporal=nt)
        torch.cuda.synchronize()
        assert gbs > 0.0
        assert torch.equal(dst[:count], src)
        assert torch.count_nonzero(dst[count:]) == 0  # nothing past the span
        gbs = lira.stream_bandwidth(lira.STREAM_WRITE, None, dst, count * 4, reps=2, nontemporal=nt)
        torch.cuda.synchronize()
        assert gbs > 0.0 and bool((dst[:count] == 1.0).all()) and torch.count_nonzero(dst[count:]) == 0
    assert lira.stream_bandwidth(lira.STREAM_READ, src, None, count * 4, reps=2) > 0.0
    with pytest.raises(RuntimeError):
        lira.stream_bandwidth(lira.STREAM_COPY, src, src, count * 4 - 4)  # not a multiple of 16 bytes
