"""Batched pointwise evaluation (fftso3_evaluate) against the oracle's pointwise inverse and
against the grid inverse at grid nodes.  FP64 plans: 1e-10 relative; FP32 storage: 2e-6.
Angles outside [0, pi] for beta (theta) are checked through the identities
f(a, b, g) = f(a + pi, 2 pi - b, g - pi) (SO(3)) and f(t, p) = f(2 pi - t, p + pi) (S^2)."""
import numpy as np
import pytest

from fftso3_amd import inputs

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _eval(torch, plan, F, ang, fn=0):
    gdt, cdt = plan.dtypes()
    dF = torch.from_numpy(np.ascontiguousarray(F).reshape(-1)).to(cdt).cuda()
    dA = torch.from_numpy(np.ascontiguousarray(ang, dtype=np.float64).reshape(-1)).cuda()
    v = plan.evaluate(dF, dA, fn=fn)
    torch.cuda.synchronize()
    return v.cpu().numpy()


def _so3_ref_point(oracle, B, F, a, b, g, real):
    b2 = np.mod(b, 2 * np.pi)
    if b2 > np.pi:
        a, b2, g = a + np.pi, 2 * np.pi - b2, g - np.pi
    return oracle.so3_real_inverse_point(B, F, a, b2, g) if real else oracle.so3_inverse_point(B, F, a, b2, g)


@pytest.mark.parametrize("kind", ["complex", "real"])
def test_so3_evaluate_fp64(fftso3, torch_cuda, oracle, kind):
    B, n = 12, 40
    real = kind == "real"
    F = inputs.so3_coefficients(B, 1, 1300, complex_=not real)
    rng = np.random.default_rng(13)
    ang = np.stack([rng.uniform(-4, 8, n), rng.uniform(-np.pi, 3 * np.pi, n), rng.uniform(-4, 8, n)], axis=1)
    ang[0] = (0.3, 0.0, 1.1)       # beta = 0 and beta = pi exactly
    ang[1] = (0.3, np.pi, -0.7)
    plan = fftso3.Plan(B - 1, 1, "so3", kind, "fp64")
    v = _eval(torch_cuda, plan, F, ang)
    ref = np.array([_so3_ref_point(oracle, B, F[0], *p, real) for p in ang])
    assert rel(v, ref) < 1e-10


@pytest.mark.parametrize("kind", ["complex", "real"])
def test_s2_evaluate_fp64(fftso3, torch_cuda, oracle, kind):
    B, n = 20, 40
    real = kind == "real"
    F = inputs.s2_coefficients(B, 1, 1400, complex_=not real)
    rng = np.random.default_rng(14)
    ang = np.stack([rng.uniform(-np.pi, 3 * np.pi, n), rng.uniform(-4, 8, n)], axis=1)
    ang[0] = (0.0, 0.4)
    ang[1] = (np.pi, 2.0)
    plan = fftso3.Plan(B - 1, 1, "s2", kind, "fp64")
    v = _eval(torch_cuda, plan, F, ang)
    ref = []
    for t, p in ang:
        t2 = np.mod(t, 2 * np.pi)
        if t2 > np.pi:
            t2, p = 2 * np.pi - t2, p + np.pi
        ref.append(oracle.s2_real_inverse_point(B, F[0], t2, p) if real else oracle.s2_inverse_point(B, F[0], t2, p))
    assert rel(v, np.array(ref)) < 1e-10


def test_evaluate_at_grid_nodes_matches_grid_inverse_fp32(fftso3, torch_cuda):
    """Function 5 of an FP32 batch of 8 at l_max=63, evaluated at 500 grid nodes, against the
    batch's grid inverse at the same nodes."""
    torch = torch_cuda
    B, n = 64, 8
    F = inputs.workload_batch_b64(n)
    plan = fftso3.Plan(B - 1, n, "so3", "complex", "fp32")
    gdt, cdt = plan.dtypes()
    dF = torch.from_numpy(F.reshape(-1)).to(cdt).cuda()
    g = plan.inverse(dF).cpu().numpy().reshape(n, 2 * B, 2 * B, 2 * B)
    rng = np.random.default_rng(15)
    idx = rng.integers(0, 2 * B, size=(500, 3))
    ang = np.stack([idx[:, 1] * np.pi / B, (2 * idx[:, 0] + 1) * np.pi / (4 * B), idx[:, 2] * np.pi / B], axis=1)
    dA = torch.from_numpy(ang.reshape(-1)).cuda()
    v = plan.evaluate(dF, dA, fn=5).cpu().numpy()
    ref = g[5, idx[:, 0], idx[:, 1], idx[:, 2]]
    assert rel(v, ref) < 2e-6


def test_evaluate_uneven_plan(fftso3, torch_cuda, oracle):
    """Each function of an uneven plan is evaluated with its own bandlimit."""
    torch = torch_cuda
    Bs = [3, 9, 6]
    Fs = [inputs.so3_coefficients(B, 1, 1500 + i)[0] for i, B in enumerate(Bs)]
    plan = fftso3.Plan([B - 1 for B in Bs], domain="so3", kind="complex", precision="fp64")
    flat = torch.from_numpy(np.concatenate(Fs)).cuda()
    ang = np.array([[0.2, 1.0, -0.5], [2.0, 0.3, 4.0], [5.0, 2.9, 1.0]])
    dA = torch.from_numpy(ang.reshape(-1)).cuda()
    for i, B in enumerate(Bs):
        v = plan.evaluate(flat, dA, fn=i).cpu().numpy()
        ref = [oracle.so3_inverse_point(B, Fs[i], *p) for p in ang]
        assert rel(v, ref) < 1e-10
