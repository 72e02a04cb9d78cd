"""lira_gram_noalloc (the Gram with its partial sums in a fixed device array, for callers that cannot allocate) against
lira_gram (the same kernels on buffers it allocates): bitwise equal on two calls in a row, the second smaller and in
the other precision, and a shape whose partials exceed the array refused before any device work."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lira_mod():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    import lira
    lira.lib()
    return lira


def _blocks(torch, n, m, q, dtype, seed):
    g = torch.Generator(device="cpu").manual_seed(seed)
    N = n ** 3
    s = torch.randn(m * N, generator=g, dtype=torch.float64).to(dtype).to("cuda")
    u = torch.randn(q * N, generator=g, dtype=torch.float64).to(dtype).to("cuda")
    return s, u


def test_gram_noalloc_matches_gram_on_two_calls(lira_mod):
    import torch
    lira = lira_mod
    cases = [(48, 0.25, 24, 48, torch.float32, 1), (33, 0.3, 9, 18, torch.float64, 2)]
    for n, h, m, q, dtype, seed in cases:
        print(f"gram_noalloc n {n} m {m} q {q} {dtype}", flush=True)
        s, u = _blocks(torch, n, m, q, dtype, seed)
        ref = lira.gram(n, h, s, m, u, q)
        g_dev = torch.full((m * q,), float("nan"), dtype=torch.float64, device="cuda")
        st = lira.gram_noalloc(n, h, s, m, u, q, g_dev)
        assert st == lira.OK
        torch.cuda.synchronize()
        got = g_dev.cpu().numpy().reshape(m, q)
        assert np.array_equal(got, ref)


def test_gram_noalloc_refuses_a_shape_beyond_its_scratch(lira_mod):
    import torch
    lira = lira_mod
    tiny = torch.zeros(8, dtype=torch.float32, device="cuda")
    out = torch.zeros(8, dtype=torch.float64, device="cuda")
    # 2048 x 2048 outputs in one chunk need 4 M partial doubles, twice the 16 MiB array: refused before any launch
    assert lira.gram_noalloc(4, 0.3, tiny, 2048, tiny, 2048, out) == lira.INVALID
    torch.cuda.synchronize()
