"""GPU parity: the HIP path (through the C ABI) against the FP64 CPU oracle on the same seeded inputs.

Tolerances: the FP64 path must match the oracle to 1e-10 relative (L2 over the array); FP32
storage with an FP32 contraction must stay within 1e-6 relative-L2 of the FP64 oracle at
l_max = 63 (the headline bar), and within 2e-6 at the other bandlimits.
"""
import numpy as np
import pytest

from fftso3_amd import inputs

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10
FP32_TOL = 2e-6
FP32_TOL_B64 = 1e-6   # l_max = 63: relative-L2 bar of the headline workload


def rel_l2(a, b):
    a = np.asarray(a).reshape(-1)
    b = np.asarray(b).reshape(-1)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _to_dev(torch, arr, dtype):
    return torch.from_numpy(np.ascontiguousarray(arr).reshape(-1)).to(dtype).cuda()


def so3_roundtrip(fftso3, torch, B, F, kind="complex", prec="fp64", grid_in=None):
    """F: [n_fn, ncoef] numpy.  Returns (GPU inverse grid, GPU forward of grid_in or of the GPU grid)."""
    n_fn = F.shape[0]
    plan = fftso3.Plan(B - 1, n_fn, "so3", kind, prec)
    gdt, cdt = plan.dtypes()
    g = plan.inverse(_to_dev(torch, F, cdt))
    src = g if grid_in is None else _to_dev(torch, grid_in, gdt)
    c = plan.forward(src)
    torch.cuda.synchronize()
    N = 2 * B
    return g.cpu().numpy().reshape(n_fn, N, N, N), c.cpu().numpy().reshape(n_fn, -1)


def test_so3_complex_fp64_b16(fftso3, torch_cuda, oracle):
    """Complex SO(3), l_max=15, seed 0: inverse to the 32^3 grid, forward back, FP64."""
    B = 16
    F = inputs.workload_complex_b16()
    g_ref = oracle.so3_inverse(B, F[0])
    F_ref = oracle.so3_forward(B, g_ref)
    g, c = so3_roundtrip(fftso3, torch_cuda, B, F, grid_in=g_ref[None])
    assert rel_l2(g[0], g_ref) < FP64_TOL
    assert rel_l2(c[0], F_ref) < FP64_TOL
    assert rel_l2(c[0], F[0]) < FP64_TOL


@pytest.mark.parametrize("B", [1, 2, 3, 6, 11, 21, 24, 32, 51, 80])
def test_so3_complex_fp64_sizes(fftso3, torch_cuda, oracle, B):
    """Bandlimits whose 2B is 2, 4, 6, 12, 22 (direct DFT), 42, 48, 64, 102 (two-pass FFT with a
    64-row tile and a partial last block) and 160 (transposed coefficient staging of the chunked
    kernels above l_max = 63)."""
    F = inputs.so3_coefficients(B, 1, 100 + B)
    g_ref = oracle.so3_inverse(B, F[0])
    g, c = so3_roundtrip(fftso3, torch_cuda, B, F, grid_in=g_ref[None])
    assert rel_l2(g[0], g_ref) < FP64_TOL
    assert rel_l2(c[0], oracle.so3_forward(B, g_ref)) < FP64_TOL


def test_so3_real_fp64_b64(fftso3, torch_cuda, oracle):
    """Real SO(3), l_max=63, seed 1: 128^3 grid, FP64 GPU vs oracle at 1e-10."""
    B = 64
    Fr = inputs.workload_real_b64()
    g_ref = oracle.so3_real_inverse(B, Fr[0])
    F_ref = oracle.so3_real_forward(B, g_ref)
    g, c = so3_roundtrip(fftso3, torch_cuda, B, Fr, kind="real", grid_in=g_ref[None])
    assert rel_l2(g[0], g_ref) < FP64_TOL
    assert rel_l2(c[0], F_ref) < FP64_TOL
    assert rel_l2(c[0], Fr[0]) < FP64_TOL


@pytest.mark.parametrize("B", [2, 5, 12])
def test_so3_real_fp64_small(fftso3, torch_cuda, oracle, B):
    Fr = inputs.so3_coefficients(B, 2, 200 + B, complex_=False)
    g, c = so3_roundtrip(fftso3, torch_cuda, B, Fr, kind="real")
    for i in range(2):
        assert rel_l2(g[i], oracle.so3_real_inverse(B, Fr[i])) < FP64_TOL
        assert rel_l2(c[i], Fr[i]) < FP64_TOL


def test_so3_fp32_b64_batch(fftso3, torch_cuda, oracle):
    """FP32 storage and contraction at l_max=63 on the first 3 functions of the seed-2 batch."""
    B = 64
    F = inputs.workload_batch_b64(3)
    g, c = so3_roundtrip(fftso3, torch_cuda, B, F, prec="fp32")
    for i in range(3):
        g_ref = oracle.so3_inverse(B, F[i])
        assert rel_l2(g[i], g_ref) < FP32_TOL_B64
        assert rel_l2(c[i], F[i]) < FP32_TOL_B64


def test_so3_fp32_b64_headline_batch(fftso3, torch_cuda, oracle):
    """The whole headline batch (256 functions, l_max=63, seed 2), FP32, function by function:
    the inverse grid against the FP64 GPU path's grid, the forward of the FP64 grid against the FP64
    GPU path's coefficients, and the round trip against the input, all at 1e-6 relative-L2; eight
    functions spread over the batch, the last included, against the CPU oracle's grid."""
    torch = torch_cuda
    B, n, N = 64, 256, 128
    F = inputs.workload_batch_b64(n)
    d64 = torch.from_numpy(F.reshape(-1)).cuda()
    p64 = fftso3.Plan(B - 1, n, "so3", "complex", "fp64")
    g64 = p64.inverse(d64)
    c64 = p64.forward(g64)
    p32 = fftso3.Plan(B - 1, n, "so3", "complex", "fp32")
    g32 = p32.inverse(d64.to(torch.complex64))
    rt32 = p32.forward(g32)
    f32 = p32.forward(g64.to(torch.complex64))
    torch.cuda.synchronize()
    print("headline batch: transforms done", flush=True)

    def per_fn(a, b):
        a = a.reshape(n, -1).to(torch.complex128)
        b = b.reshape(n, -1).to(torch.complex128)
        return (torch.linalg.norm(a - b, dim=1) / torch.linalg.norm(b, dim=1)).cpu().numpy()

    e_grid = per_fn(g32, g64)
    e_fwd = per_fn(f32, c64)
    e_rt = per_fn(rt32, d64)
    e_64 = per_fn(c64, d64)
    print("worst per function: grid %.2e forward %.2e round trip %.2e fp64 round trip %.2e"
          % (e_grid.max(), e_fwd.max(), e_rt.max(), e_64.max()), flush=True)
    assert e_grid.max() < FP32_TOL_B64
    assert e_fwd.max() < FP32_TOL_B64
    assert e_rt.max() < FP32_TOL_B64
    assert e_64.max() < FP64_TOL
    del g64, f32
    gv = g32.reshape(n, N ** 3)
    for i in (0, 37, 73, 110, 146, 183, 219, 255):
        g_ref = oracle.so3_inverse(B, F[i])
        assert rel_l2(gv[i].cpu().numpy(), g_ref) < FP32_TOL_B64, i


@pytest.mark.parametrize("B,kind", [(2, "complex"), (7, "complex"), (16, "real"), (33, "complex"), (64, "complex"),
                                    (64, "real")])
def test_so3_fp64_mfma_contractions(fftso3, torch_cuda, oracle, B, kind):
    """FP64 batches of 8 functions per pass contract on FP64 MFMA (mfma64.hpp): 11 functions (a full
    pass and a tail of 3), every function round trips at 1e-10 and the first and last match the
    oracle's inverse and forward."""
    n = 11
    cplx = kind == "complex"
    F = inputs.so3_coefficients(B, n, 700 + B, complex_=cplx)
    torch = torch_cuda
    plan = fftso3.Plan(B - 1, n, "so3", kind, "fp64")
    assert "mfma64" in plan.describe(), plan.describe()
    gdt, cdt = plan.dtypes()
    g = plan.inverse(_to_dev(torch, F, cdt))
    c = plan.forward(g)
    g = g.cpu().numpy().reshape(n, -1)
    c = c.cpu().numpy().reshape(n, -1)
    for i in range(n):
        assert rel_l2(c[i], F[i]) < FP64_TOL, i
    for i in (0, n - 1):
        g_ref = oracle.so3_inverse(B, F[i]) if cplx else oracle.so3_real_inverse(B, F[i])
        assert rel_l2(g[i], g_ref.reshape(-1)) < FP64_TOL, i
        c_ref = oracle.so3_forward(B, g_ref) if cplx else oracle.so3_real_forward(B, g_ref)
        assert rel_l2(c[i], c_ref) < FP64_TOL, i


def test_so3_chunk_tail(fftso3, torch_cuda, oracle):
    """5 functions through chunks of 4 (a full chunk and a tail of 1)."""
    B = 6
    F = inputs.so3_coefficients(B, 5, 7)
    plan = fftso3.Plan(B - 1, 5, "so3", "complex", "fp64", chunk=4)
    torch = torch_cuda
    g = plan.inverse(_to_dev(torch, F, torch.complex128))
    c = plan.forward(g)
    g = g.cpu().numpy().reshape(5, -1)
    c = c.cpu().numpy().reshape(5, -1)
    for i in range(5):
        assert rel_l2(g[i], oracle.so3_inverse(B, F[i]).reshape(-1)) < FP64_TOL
        assert rel_l2(c[i], F[i]) < FP64_TOL


def test_so3_uneven_batch(fftso3, torch_cuda, oracle):
    """Uneven bandlimits in one plan; functions packed back to back."""
    torch = torch_cuda
    Bs = [3, 8, 5, 8, 16, 2]
    Fs = [inputs.so3_coefficients(B, 1, 300 + i)[0] for i, B in enumerate(Bs)]
    plan = fftso3.Plan([B - 1 for B in Bs], domain="so3", kind="complex", precision="fp64")
    flat = np.concatenate(Fs)
    g = plan.inverse(_to_dev(torch, flat, torch.complex128))
    c = plan.forward(g)
    g = g.cpu().numpy()
    c = c.cpu().numpy()
    for i, B in enumerate(Bs):
        go, co = plan.grid_offset(i), plan.coef_offset(i)
        gi = g[go:go + (2 * B) ** 3]
        ci = c[co:co + len(Fs[i])]
        assert rel_l2(gi, oracle.so3_inverse(B, Fs[i]).reshape(-1)) < FP64_TOL
        assert rel_l2(ci, Fs[i]) < FP64_TOL


def s2_roundtrip(fftso3, torch, B, F, kind="complex", prec="fp64", grid_in=None):
    n_fn = F.shape[0]
    plan = fftso3.Plan(B - 1, n_fn, "s2", kind, prec)
    gdt, cdt = plan.dtypes()
    g = plan.inverse(_to_dev(torch, F, cdt))
    src = g if grid_in is None else _to_dev(torch, grid_in, gdt)
    c = plan.forward(src)
    torch.cuda.synchronize()
    N = 2 * B
    return g.cpu().numpy().reshape(n_fn, N, N), c.cpu().numpy().reshape(n_fn, -1)


@pytest.mark.parametrize("B", [1, 3, 6, 16, 21, 64, 97, 128])
def test_s2_complex_fp64(fftso3, torch_cuda, oracle, B):
    F = inputs.s2_coefficients(B, 2, 400 + B)
    g_ref = np.stack([oracle.s2_inverse(B, F[i]) for i in range(2)])
    g, c = s2_roundtrip(fftso3, torch_cuda, B, F, grid_in=g_ref)
    for i in range(2):
        assert rel_l2(g[i], g_ref[i]) < FP64_TOL
        assert rel_l2(c[i], oracle.s2_forward(B, g_ref[i])) < FP64_TOL
        assert rel_l2(c[i], F[i]) < FP64_TOL


@pytest.mark.parametrize("B", [2, 6, 32])
def test_s2_real_fp64(fftso3, torch_cuda, oracle, B):
    F = inputs.s2_coefficients(B, 2, 500 + B, complex_=False)
    g, c = s2_roundtrip(fftso3, torch_cuda, B, F, kind="real")
    for i in range(2):
        assert rel_l2(g[i], oracle.s2_real_inverse(B, F[i])) < FP64_TOL
        assert rel_l2(c[i], oracle.s2_real_forward(B, oracle.s2_real_inverse(B, F[i]))) < FP64_TOL


def test_s2_fp32_b128(fftso3, torch_cuda, oracle):
    """S^2 at l_max=127 (256x256 grid), first 4 fields of the seed-5 workload, FP32."""
    B = 128
    F = inputs.workload_s2(4)
    g, c = s2_roundtrip(fftso3, torch_cuda, B, F, prec="fp32")
    for i in range(4):
        assert rel_l2(g[i], oracle.s2_inverse(B, F[i])) < FP32_TOL
        assert rel_l2(c[i], F[i]) < FP32_TOL


# split-BF16 contraction on MFMA: BF16X6 (3 pieces, 6 products) is held to the FP32 bar;
# BF16X3 (2 pieces, 3 products) keeps ~16 significant bits per operand, bar 5e-5.
SPLIT_TOL = {"bf16x6": FP32_TOL, "bf16x3": 5e-5}


@pytest.mark.parametrize("prec", ["bf16x6", "bf16x3"])
def test_so3_split_bf16_b64(fftso3, torch_cuda, oracle, prec):
    B = 64
    F = inputs.workload_batch_b64(3)
    g, c = so3_roundtrip(fftso3, torch_cuda, B, F, prec=prec)
    for i in range(3):
        assert rel_l2(g[i], oracle.so3_inverse(B, F[i])) < SPLIT_TOL[prec]
        assert rel_l2(c[i], F[i]) < SPLIT_TOL[prec]


@pytest.mark.parametrize("B", [2, 7, 16, 33])
def test_so3_split_bf16_sizes(fftso3, torch_cuda, oracle, B):
    """Small and odd bandlimits, 10 functions (a full chunk of 8 and a tail of 2), real and complex."""
    F = inputs.so3_coefficients(B, 10, 600 + B)
    g, c = so3_roundtrip(fftso3, torch_cuda, B, F, prec="bf16x6")
    for i in (0, 7, 9):
        assert rel_l2(g[i], oracle.so3_inverse(B, F[i])) < FP32_TOL
        assert rel_l2(c[i], F[i]) < FP32_TOL
    Fr = inputs.so3_coefficients(B, 3, 700 + B, complex_=False)
    g, c = so3_roundtrip(fftso3, torch_cuda, B, Fr, kind="real", prec="bf16x6")
    for i in range(3):
        assert rel_l2(g[i], oracle.so3_real_inverse(B, Fr[i])) < FP32_TOL
        assert rel_l2(c[i], Fr[i]) < FP32_TOL


@pytest.mark.parametrize("prec", ["bf16x6", "bf16x3"])
def test_s2_split_bf16_b128(fftso3, torch_cuda, oracle, prec):
    B = 128
    F = inputs.workload_s2(4)
    g, c = s2_roundtrip(fftso3, torch_cuda, B, F, prec=prec)
    for i in range(4):
        assert rel_l2(g[i], oracle.s2_inverse(B, F[i])) < SPLIT_TOL[prec]
        assert rel_l2(c[i], F[i]) < SPLIT_TOL[prec]


@pytest.mark.parametrize("domain,B", [("s2", 75), ("s2", 100), ("s2", 128), ("so3", 70), ("so3", 100), ("so3", 128)])
def test_batchwide_large_n_fp32(fftso3, torch_cuda, oracle, domain, B):
    """FP32 batch-wide contraction with the d block resident in LDS (128 < N <= 256), including N
    that are not multiples of 32; 3 functions (S^2: 5) against the oracle, complex and real."""
    n = 5 if domain == "s2" else 3
    if domain == "s2":
        F = inputs.s2_coefficients(B, n, 800 + B)
        g, c = s2_roundtrip(fftso3, torch_cuda, B, F, prec="fp32")
        for i in range(n):
            assert rel_l2(g[i], oracle.s2_inverse(B, F[i])) < FP32_TOL
            assert rel_l2(c[i], F[i]) < FP32_TOL
        Fr = inputs.s2_coefficients(B, 2, 900 + B, complex_=False)
        g, c = s2_roundtrip(fftso3, torch_cuda, B, Fr, kind="real", prec="fp32")
        for i in range(2):
            assert rel_l2(g[i], oracle.s2_real_inverse(B, Fr[i])) < FP32_TOL
            assert rel_l2(c[i], Fr[i]) < FP32_TOL
    else:
        F = inputs.so3_coefficients(B, n, 800 + B, scale=lambda l: 1.0 / (1.0 + l))
        g, c = so3_roundtrip(fftso3, torch_cuda, B, F, prec="fp32")
        for i in range(n):
            assert rel_l2(g[i], oracle.so3_inverse(B, F[i])) < FP32_TOL
            assert rel_l2(c[i], F[i]) < FP32_TOL


@pytest.mark.parametrize("domain,B", [("s2", 150), ("s2", 152), ("so3", 136)])
def test_table_path_fp32(fftso3, torch_cuda, oracle, domain, B):
    """l_max above 127 on the chunked kernels with the plan's Wigner-d table (MFMA inverse in two
    launches split at 128 degrees; MFMA forward, half-staged when B is a multiple of 8: S^2 152,
    SO(3) 136), complex and real, against the oracle."""
    if domain == "s2":
        F = inputs.s2_coefficients(B, 3, 1600 + B, scale=lambda l: 1.0 / (1.0 + l))
        g, c = s2_roundtrip(fftso3, torch_cuda, B, F, prec="fp32")
        for i in range(3):
            assert rel_l2(g[i], oracle.s2_inverse(B, F[i])) < FP32_TOL
            assert rel_l2(c[i], F[i]) < FP32_TOL
        Fr = inputs.s2_coefficients(B, 2, 1700 + B, complex_=False, scale=lambda l: 1.0 / (1.0 + l))
        g, c = s2_roundtrip(fftso3, torch_cuda, B, Fr, kind="real", prec="fp32")
        for i in range(2):
            assert rel_l2(g[i], oracle.s2_real_inverse(B, Fr[i])) < FP32_TOL
            assert rel_l2(c[i], Fr[i]) < FP32_TOL
    else:
        Fr = inputs.so3_coefficients(B, 1, 1600 + B, complex_=False, scale=lambda l: 1.0 / (1.0 + l))
        g, c = so3_roundtrip(fftso3, torch_cuda, B, Fr, kind="real", prec="fp32")
        assert rel_l2(g[0], oracle.so3_real_inverse(B, Fr[0])) < 1e-5
        assert rel_l2(c[0], Fr[0]) < 1e-5


@pytest.mark.parametrize("klo,gil,split", [(0, 0, 1), (0, 1, 1), (0, 2, 1), (0, 1, 2), (3, 0, 1), (0, 0, 4)])
@pytest.mark.parametrize("kind", ["complex", "real"])
def test_batchwide_schedules_fp32(fftso3, torch_cuda, oracle, klo, gil, split, kind):
    """l_max=63 FP32 batch under each schedule autotune may pick (fftso3_plan_set_schedule):
    group interleave gil 0 (one-slice FFTs, one group per contraction workgroup), 1 (paired-slice
    FFTs, grouped 16x16 forward and 32x32 inverse contractions) and 2 (4 groups per workgroup);
    blocked workspace rows (klo 3) and group splits.  37 functions leave partial column blocks for
    every block width.  Round trip of all, and two functions against the oracle."""
    torch = torch_cuda
    B, n = 64, 37
    F = inputs.so3_coefficients(B, n, 900 + gil, complex_=(kind == "complex"),
                                scale=lambda l: 1.0 / (1.0 + l))
    plan = fftso3.Plan(B - 1, n, "so3", kind, "fp32")
    plan.set_schedule(klo, gil, split)
    assert f"klo={klo} gil={gil} split={split}" in plan.describe()
    gdt, cdt = plan.dtypes()
    g = plan.inverse(_to_dev(torch, F, cdt))
    c = plan.forward(g)
    torch.cuda.synchronize()
    g = g.cpu().numpy().reshape(n, -1)
    c = c.cpu().numpy().reshape(n, -1)
    assert rel_l2(c, F) < FP32_TOL_B64
    for i in (0, n - 1):
        ref = oracle.so3_inverse(B, F[i]) if kind == "complex" else oracle.so3_real_inverse(B, F[i])
        assert rel_l2(g[i], ref.reshape(-1)) < FP32_TOL_B64


@pytest.mark.parametrize("prec", ["bf16x6", "bf16x3"])
@pytest.mark.parametrize("domain,B,kind", [("so3", 64, "complex"), ("so3", 64, "real"), ("so3", 20, "complex"),
                                           ("s2", 64, "complex"), ("s2", 45, "real")])
def test_split_bf16_batchwide(fftso3, torch_cuda, oracle, prec, domain, B, kind):
    """Split-BF16 on the batch-wide kernels (contract_fwd_mxs / contract_inv_mxs): 37 functions
    leave a partial last column block; every function round trips, the first and the last
    against the oracle.  BF16X6 holds the FP32 bar (1e-6 at l_max = 63), BF16X3 its own 5e-5."""
    torch = torch_cuda
    n = 37
    cplx = kind == "complex"
    if domain == "so3":
        F = inputs.so3_coefficients(B, n, 7000 + B, complex_=cplx, scale=lambda l: 1.0 / (1.0 + l))
    else:
        F = inputs.s2_coefficients(B, n, 7000 + B, complex_=cplx, scale=lambda l: 1.0 / (1.0 + l))
    plan = fftso3.Plan(B - 1, n, domain, kind, prec)
    assert "batch-wide" in plan.describe()
    if domain == "so3" and B == 64 and kind == "real":
        plan.set_schedule(0, 0, 1)   # one group per workgroup and one-slice FFTs (the default interleaves 2)
    gdt, cdt = plan.dtypes()
    g = plan.inverse(_to_dev(torch, F, cdt))
    c = plan.forward(g)
    torch.cuda.synchronize()
    g = g.cpu().numpy().reshape(n, -1)
    c = c.cpu().numpy().reshape(n, -1)
    tol = (FP32_TOL_B64 if B == 64 else FP32_TOL) if prec == "bf16x6" else SPLIT_TOL["bf16x3"]
    assert rel_l2(c, F) < tol
    for i in (0, n - 1):
        if domain == "so3":
            ref = oracle.so3_inverse(B, F[i]) if cplx else oracle.so3_real_inverse(B, F[i])
        else:
            ref = oracle.s2_inverse(B, F[i]) if cplx else oracle.s2_real_inverse(B, F[i])
        assert rel_l2(g[i], ref) < tol, i
    plan.close()


@pytest.mark.parametrize("prec", ["bf16x6", "bf16x3"])
@pytest.mark.parametrize("domain,B,kind,n", [("s2", 75, "complex", 37), ("s2", 128, "complex", 37), ("s2", 100, "real", 37),
                                             ("so3", 70, "complex", 19), ("so3", 100, "real", 5)])
def test_split_bf16_batchwide_large_n(fftso3, torch_cuda, oracle, prec, domain, B, kind, n):
    """Split-BF16 on the batch-wide K-chunked kernels for 128 < N <= 256 (contract_fwd_mxks /
    contract_inv_mxks: items of 64 degrees and 128 beta values, N not a multiple of 32 at B = 75
    and 70, partial column blocks): every function round trips, the first and the last against
    the oracle.  BF16X6 holds the FP32 bar, BF16X3 its own."""
    torch = torch_cuda
    cplx = kind == "complex"
    if domain == "so3":
        F = inputs.so3_coefficients(B, n, 7300 + B, complex_=cplx, scale=lambda l: 1.0 / (1.0 + l))
    else:
        F = inputs.s2_coefficients(B, n, 7300 + B, complex_=cplx, scale=lambda l: 1.0 / (1.0 + l))
    plan = fftso3.Plan(B - 1, n, domain, kind, prec)
    assert "k-chunked" in plan.describe()
    gdt, cdt = plan.dtypes()
    g = plan.inverse(_to_dev(torch, F, cdt))
    c = plan.forward(g)
    torch.cuda.synchronize()
    g = g.cpu().numpy().reshape(n, -1)
    c = c.cpu().numpy().reshape(n, -1)
    tol = FP32_TOL if prec == "bf16x6" else SPLIT_TOL["bf16x3"]
    assert rel_l2(c, F) < tol
    for i in (0, n - 1):
        if domain == "so3":
            ref = oracle.so3_inverse(B, F[i]) if cplx else oracle.so3_real_inverse(B, F[i])
        else:
            ref = oracle.s2_inverse(B, F[i]) if cplx else oracle.s2_real_inverse(B, F[i])
        assert rel_l2(g[i], ref) < tol, i
    plan.close()


def test_split_bf16_chunked_kernels(fftso3, torch_cuda, oracle):
    """set_chunk(8) keeps the chunked split-BF16 kernels (mfma.hpp) reachable at l_max <= 63."""
    torch = torch_cuda
    B, n = 20, 11
    F = inputs.so3_coefficients(B, n, 7100)
    plan = fftso3.Plan(B - 1, n, "so3", "complex", "bf16x6", chunk=8)
    assert "chunked" in plan.describe()
    g = plan.inverse(_to_dev(torch, F, torch.complex64))
    c = plan.forward(g)
    torch.cuda.synchronize()
    g = g.cpu().numpy().reshape(n, -1)
    assert rel_l2(c.cpu().numpy().reshape(n, -1), F) < FP32_TOL
    assert rel_l2(g[n - 1], oracle.so3_inverse(B, F[n - 1])) < FP32_TOL
