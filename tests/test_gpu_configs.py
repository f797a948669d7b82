"""GPU parity on the larger workloads: 512 uneven real functions (bandlimits 8..64), l_max=255,
and 4096 S^2 fields at l_max=127.

Parity is checked against the FP64 oracle on a sample of functions (or of grid points, at
l_max=255 where a full oracle transform takes minutes), and on every function through a
size-independent property: the inverse-then-forward round trip returns the coefficients.
Tolerances: FP64 1e-10 relative (L2); FP32 storage and arithmetic 2e-6, and 1e-5 at l_max=255
(a 512-point beta sum and 512-point DFTs in FP32).
"""
import numpy as np
import pytest

from fftso3_amd import inputs

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.asarray(a).reshape(-1)
    b = np.asarray(b).reshape(-1)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_uneven_real_fp32_batch(fftso3, torch_cuda, oracle):
    """Config 4 shape: real functions with bandlimits drawn from {8,16,24,32,48,64} (seed 3).
    All 512 round trip; one function of every bandlimit is checked against the oracle."""
    torch = torch_cuda
    Bs, coefs = inputs.workload_uneven(512)
    print("uneven real fp32: 512 functions, bandlimits", sorted(set(Bs)), flush=True)
    plan = fftso3.Plan([B - 1 for B in Bs], domain="so3", kind="real", precision="fp32")
    flat = torch.from_numpy(np.concatenate(coefs)).to(torch.float32).cuda()
    g = plan.inverse(flat)
    c = plan.forward(g)
    torch.cuda.synchronize()
    g = g.cpu().numpy()
    c = c.cpu().numpy().astype(np.float64)
    worst = 0.0
    for i, B in enumerate(Bs):
        co = plan.coef_offset(i)
        worst = max(worst, rel_l2(c[co:co + len(coefs[i])], coefs[i]))
    assert worst < 2e-6
    seen = set()
    for i, B in enumerate(Bs):
        if B in seen:
            continue
        seen.add(B)
        go = plan.grid_offset(i)
        gi = g[go:go + (2 * B) ** 3].astype(np.float64)
        assert rel_l2(gi, oracle.so3_real_inverse(B, coefs[i]).reshape(-1)) < 2e-6, B


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-10), ("fp32", 1e-5), ("bf16x6", 1e-5), ("bf16x3", 5e-5)])
def test_so3_b256(fftso3, torch_cuda, oracle, prec, tol):
    """Config 5a: l_max=255 (512^3 grids), all 8 functions of the seed-4 workload.  Every function
    round trips on the GPU; the grids of the first and the last function at 12 sampled points each
    against the oracle's pointwise inverse.  The split precisions run the on-the-fly recursion
    kernels (d-on-the-fly), FP32 the Wigner-d table."""
    torch = torch_cuda
    B, n = 256, 8
    F = inputs.workload_b256(n)
    print("b256 %s: plan" % prec, flush=True)
    plan = fftso3.Plan(B - 1, n, "so3", "complex", prec)
    if prec.startswith("bf16"):
        assert "d-on-the-fly" in plan.describe()
    gdt, cdt = plan.dtypes()
    dF = torch.from_numpy(F.reshape(-1)).to(cdt).cuda()
    g = plan.inverse(dF)
    c = plan.forward(g)
    torch.cuda.synchronize()
    print("b256 %s: transforms done" % prec, flush=True)
    a = c.reshape(n, -1).to(torch.complex128)
    b = dF.reshape(n, -1).to(torch.complex128)
    err = (torch.linalg.norm(a - b, dim=1) / torch.linalg.norm(b, dim=1)).cpu().numpy()
    print("b256 %s: round trip per function %s" % (prec, np.array2string(err, precision=2)), flush=True)
    assert float(err.max()) < tol
    N = 2 * B
    rng = np.random.default_rng(256)
    grid = g.reshape(n, N, N, N)
    for i in (0, n - 1):
        pts = rng.integers(0, N, size=(12, 3))
        got = np.array([complex(grid[i, k, j1, j2].item()) for k, j1, j2 in pts])
        ref = np.array([oracle.so3_inverse_point(B, F[i], j1 * np.pi / B, (2 * k + 1) * np.pi / (4 * B),
                                                 j2 * np.pi / B) for k, j1, j2 in pts])
        assert rel_l2(got, ref) < tol, i


@pytest.mark.parametrize("prec,tol", [("bf16x6", 1e-5), ("bf16x3", 5e-5)])
def test_so3_b256_split_partial_chunk(fftso3, torch_cuda, prec, tol):
    """l_max=255 in split precision with 3 functions (a chunk of 8 with 5 empty column slots): the
    inverse grids against the FP64 GPU path's over every grid point, and the forward of the FP64
    grids against the coefficients (the FP64 path itself is held to the oracle in test_so3_b256)."""
    torch = torch_cuda
    B, n = 256, 3
    F = inputs.so3_coefficients(B, n, 4256, scale=lambda l: 1.0 / (1.0 + l))
    ref = fftso3.Plan(B - 1, n, "so3", "complex", "fp64")
    dF64 = torch.from_numpy(F.reshape(-1)).to(torch.complex128).cuda()
    g64 = ref.inverse(dF64)
    plan = fftso3.Plan(B - 1, n, "so3", "complex", prec)
    assert "d-on-the-fly" in plan.describe()
    g = plan.inverse(dF64.to(torch.complex64))
    c = plan.forward(g64.to(torch.complex64))
    torch.cuda.synchronize()
    eg = (torch.linalg.norm((g.to(torch.complex128) - g64).reshape(n, -1), dim=1) /
          torch.linalg.norm(g64.reshape(n, -1), dim=1)).cpu().numpy()
    ec = (torch.linalg.norm((c.to(torch.complex128) - dF64).reshape(n, -1), dim=1) /
          torch.linalg.norm(dF64.reshape(n, -1), dim=1)).cpu().numpy()
    print("b256 %s partial chunk: grid %s coef %s" % (prec, eg, ec), flush=True)
    assert float(eg.max()) < tol and float(ec.max()) < tol


def test_s2_b128_batch4096_fp32(fftso3, torch_cuda, oracle):
    """Config 5b: 4096 fields at l_max=127.  All round trip; the first 4 against the oracle."""
    torch = torch_cuda
    B, n = 128, 4096
    F = inputs.workload_s2(n)
    plan = fftso3.Plan(B - 1, n, "s2", "complex", "fp32")
    dF = torch.from_numpy(F.reshape(-1)).to(torch.complex64).cuda()
    g = plan.inverse(dF)
    c = plan.forward(g)
    torch.cuda.synchronize()
    c = c.cpu().numpy().reshape(n, -1).astype(np.complex128)
    err = np.linalg.norm(c - F, axis=1) / np.linalg.norm(F, axis=1)
    assert float(err.max()) < 2e-6
    g = g.cpu().numpy().reshape(n, 2 * B, 2 * B)
    for i in range(4):
        assert rel_l2(g[i], oracle.s2_inverse(B, F[i])) < 2e-6
