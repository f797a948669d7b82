"""Graph mode and schedule tuning of a plan.

Graph replays must reproduce the eager transforms bit for bit (same kernels, no atomics): each
replay runs into zeroed outputs, with eager work on the same stream between replays.  After
autotune the plan must still meet the parity bar against the oracle.
"""
import numpy as np
import pytest

from fftso3_amd import inputs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("B,n,prec,domain", [(16, 4, "fp64", "so3"), (64, 8, "fp32", "so3"), (40, 6, "fp32", "s2")])
def test_graph_replays_match_eager(fftso3, torch_cuda, B, n, prec, domain):
    torch = torch_cuda
    F = inputs.so3_coefficients(B, n, 1000 + B) if domain == "so3" else inputs.s2_coefficients(B, n, 1000 + B)
    plan = fftso3.Plan(B - 1, n, domain, "complex", prec)
    gdt, cdt = plan.dtypes()
    s = torch.cuda.Stream()
    dF = torch.from_numpy(F.reshape(-1)).to(cdt).cuda()
    with torch.cuda.stream(s):
        g0 = plan.inverse(dF, stream=s)
        c0 = plan.forward(g0, stream=s)
    s.synchronize()
    g = torch.empty_like(g0)
    c = torch.empty_like(c0)
    plan.set_graph(True)
    for rep in range(4):
        with torch.cuda.stream(s):
            g.zero_()
            c.zero_()
            junk = torch.randn(1 << 16, device="cuda") * 3.0   # eager work between replays
            plan.inverse(dF, g, stream=s)
            plan.forward(g, c, stream=s)
            junk += 1.0
        s.synchronize()
        assert torch.equal(g, g0), "grid differs on replay %d" % rep
        assert torch.equal(c, c0), "coefficients differ on replay %d" % rep
    plan.set_graph(False)
    plan.close()


def test_autotune_keeps_parity(fftso3, torch_cuda, oracle):
    torch = torch_cuda
    B, n = 32, 16
    F = inputs.so3_coefficients(B, n, 1234, scale=lambda l: 1.0 / (1.0 + l))
    plan = fftso3.Plan(B - 1, n, "so3", "complex", "fp32")
    plan.autotune(reps=1)
    gdt, cdt = plan.dtypes()
    g = plan.inverse(torch.from_numpy(F.reshape(-1)).to(cdt).cuda())
    c = plan.forward(g)
    torch.cuda.synchronize()
    g = g.cpu().numpy().reshape(n, -1)
    c = c.cpu().numpy().reshape(n, -1)
    for i in (0, n - 1):
        ref = oracle.so3_inverse(B, F[i]).reshape(-1)
        assert np.linalg.norm(g[i] - ref) / np.linalg.norm(ref) < 2e-6
        assert np.linalg.norm(c[i] - F[i]) / np.linalg.norm(F[i]) < 2e-6


def test_failed_reconfiguration_leaves_plan_unusable(fftso3, torch_cuda, oracle):
    """An allocation that fails inside set_chunk leaves the plan refusing to transform (EINVAL,
    no launch) instead of running kernels on freed buffers; a later successful set_chunk
    restores it, and its results meet the FP64 bar again."""
    import ctypes
    torch = torch_cuda
    L = fftso3.lib()
    B, n = 12, 5
    F = inputs.so3_coefficients(B, n, 4321)
    plan = fftso3.Plan(B - 1, n, "so3", "complex", "fp64")
    dF = torch.from_numpy(F.reshape(-1)).to(torch.complex128).cuda()
    g0 = plan.inverse(dF)
    torch.cuda.synchronize()
    assert L.fftso3_debug_fail_alloc(1) == 0
    with pytest.raises(fftso3.FFTSO3Error) as e:
        plan.set_chunk(2)
    assert e.value.status.startswith("FFTSO3_ENOMEM")
    assert L.fftso3_debug_fail_alloc(0) == 0
    g = torch.full_like(g0, 7.0)
    st = L.fftso3_inverse(plan.handle, ctypes.c_void_p(dF.data_ptr()), ctypes.c_void_p(g.data_ptr()), None)
    assert st == 1 and b"not usable" in L.fftso3_last_error()
    torch.cuda.synchronize()
    assert bool((g == 7.0).all()), "a kernel ran on the unusable plan"
    plan.set_chunk(2)
    g = plan.inverse(dF)
    c = plan.forward(g)
    torch.cuda.synchronize()
    g = g.cpu().numpy().reshape(n, -1)
    c = c.cpu().numpy().reshape(n, -1)
    for i in (0, n - 1):
        ref = oracle.so3_inverse(B, F[i]).reshape(-1)
        assert np.linalg.norm(g[i] - ref) / np.linalg.norm(ref) < 1e-10
        assert np.linalg.norm(c[i] - F[i]) / np.linalg.norm(F[i]) < 1e-10
    plan.close()
