"""fftso3_inverse_noalloc: the inverse transform with its workspaces in the library's static device
scratch instead of the plan's buffers.  Through the C ABI, its grids must equal those of
fftso3_inverse bit for bit (same kernels in the same order), on two calls in a row of which the
second has the smaller input, for FP32 and FP64, complex and real plans; a plan whose workspace
exceeds the scratch is refused with FFTSO3_EINVAL and leaves the output untouched.
"""
import ctypes

import numpy as np
import pytest

from fftso3_amd import inputs

pytestmark = pytest.mark.gpu


def _inverse_both(fftso3, torch, plan, F):
    L = fftso3.lib()
    gdt, cdt = plan.dtypes()
    dF = torch.from_numpy(np.ascontiguousarray(F).reshape(-1)).to(cdt).cuda()
    g_ref = torch.full((plan.grid_elems(),), 3.0, dtype=gdt, device="cuda")
    g_new = torch.full((plan.grid_elems(),), -5.0, dtype=gdt, device="cuda")
    s = torch.cuda.current_stream()
    st = L.fftso3_inverse(plan.handle, ctypes.c_void_p(dF.data_ptr()), ctypes.c_void_p(g_ref.data_ptr()),
                          ctypes.c_void_p(s.cuda_stream))
    assert st == 0, L.fftso3_last_error()
    st = L.fftso3_inverse_noalloc(plan.handle, ctypes.c_void_p(dF.data_ptr()), ctypes.c_void_p(g_new.data_ptr()),
                                  ctypes.c_void_p(s.cuda_stream))
    assert st == 0, L.fftso3_last_error()
    torch.cuda.synchronize()
    return g_ref, g_new


@pytest.mark.parametrize("prec,kind", [("fp32", "complex"), ("fp64", "complex"), ("fp32", "real")])
def test_inverse_noalloc_matches_inverse(fftso3, torch_cuda, oracle, prec, kind):
    torch = torch_cuda
    cplx = kind == "complex"
    # first call: l_max = 63, 4 functions (FP64: l_max = 31, whose workspace fits the scratch);
    # second call: l_max = 20, 2 functions (the smaller input)
    first = (64, 4, 5150) if prec == "fp32" else (32, 4, 5152)
    for B, n, seed in (first, (21, 2, 5151)):
        print("noalloc %s %s l_max=%d n=%d" % (prec, kind, B - 1, n), flush=True)
        F = inputs.so3_coefficients(B, n, seed, complex_=cplx, scale=lambda l: 1.0 / (1.0 + l))
        plan = fftso3.Plan(B - 1, n, "so3", kind, prec)
        g_ref, g_new = _inverse_both(fftso3, torch, plan, F)
        assert torch.equal(g_ref, g_new)
        # and the grid itself meets the parity bar against the oracle (last function)
        N = 2 * B
        got = g_new.reshape(n, -1)[n - 1].cpu().numpy().astype(np.complex128 if cplx else np.float64)
        ref = (oracle.so3_inverse(B, F[n - 1]) if cplx else oracle.so3_real_inverse(B, F[n - 1])).reshape(-1)
        assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < (1e-10 if prec == "fp64" else 1e-6)
        assert got.size == N ** 3
        plan.close()


def test_inverse_noalloc_s2(fftso3, torch_cuda):
    torch = torch_cuda
    for B, n, seed in ((128, 16, 5160), (33, 3, 5161)):
        F = inputs.s2_coefficients(B, n, seed)
        plan = fftso3.Plan(B - 1, n, "s2", "complex", "fp32")
        g_ref, g_new = _inverse_both(fftso3, torch, plan, F)
        assert torch.equal(g_ref, g_new)
        plan.close()


def test_inverse_noalloc_rejects_oversize_plan(fftso3, torch_cuda):
    torch = torch_cuda
    L = fftso3.lib()
    plan = fftso3.Plan(63, 16, "so3", "complex", "fp32")
    assert plan.workspace_bytes() > 128 << 20
    dF = torch.zeros(plan.coef_elems(), dtype=torch.complex64, device="cuda")
    g = torch.full((plan.grid_elems(),), 9.0, dtype=torch.complex64, device="cuda")
    st = L.fftso3_inverse_noalloc(plan.handle, ctypes.c_void_p(dF.data_ptr()), ctypes.c_void_p(g.data_ptr()), None)
    assert st == 1
    assert b"static scratch" in L.fftso3_last_error()
    torch.cuda.synchronize()
    assert bool((g == 9.0).all())
    plan.close()
