"""Uneven batches through the unified schedule (one persistent launch per stage over every
bandlimit class, multi.hpp) and the N = 16R slice FFTs (fft_r16.hpp), against the oracle.

The unified launches run the same device functions per unit as the per-class launches; the
l_max = 63 slices are compiled with an opaque thread index inside the persistent loop, so the
last bits may differ (instruction selection), and the two schedules must agree to 1e-6 relative
per function with a plan created with FFTSO3_FLAG_PER_CLASS_LAUNCHES (the bandlimits below
l_max = 63 bit for bit); every bandlimit is also checked against the FP64 oracle (FP32 bar 2e-6,
1e-6 at l_max = 63).
"""
import numpy as np
import pytest

from fftso3_amd import inputs

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.asarray(a).reshape(-1)
    b = np.asarray(b).reshape(-1)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("B", [8, 16, 24, 32, 48])
@pytest.mark.parametrize("kind", ["complex", "real"])
def test_r16_slice_fft_uniform(fftso3, torch_cuda, oracle, B, kind):
    """Uniform FP32 plans of B = 8R run the 16R slice FFTs (several slices per workgroup; 13
    functions leave partial workgroups); every function round trips, two against the oracle."""
    torch = torch_cuda
    n = 13
    cplx = kind == "complex"
    F = inputs.so3_coefficients(B, n, 9000 + B, complex_=cplx, scale=lambda l: 1.0 / (1.0 + l))
    plan = fftso3.Plan(B - 1, n, "so3", kind, "fp32")
    assert "fft16R" in plan.describe()
    gdt, cdt = plan.dtypes()
    g = plan.inverse(torch.from_numpy(F.reshape(-1)).to(cdt).cuda())
    c = plan.forward(g)
    torch.cuda.synchronize()
    g = g.cpu().numpy().reshape(n, -1)
    c = c.cpu().numpy().reshape(n, -1)
    assert rel_l2(c, F) < 2e-6
    for i in (0, n - 1):
        ref = oracle.so3_inverse(B, F[i]) if cplx else oracle.so3_real_inverse(B, F[i])
        assert rel_l2(g[i], ref) < 2e-6, i
    plan.close()


@pytest.mark.parametrize("kind,precision", [("complex", "fp32"), ("real", "fp32"), ("real", "bf16x6"),
                                            ("complex", "bf16x3")])
def test_unified_matches_per_class_and_oracle(fftso3, torch_cuda, oracle, kind, precision):
    """Every bandlimit of config 4 ({8,16,24,32,48,64}) in one plan, 3 to 5 functions each in
    mixed order: unified launches (one per stage) equal the per-class launches (FP32: bit for bit
    below l_max = 63, where both run the same kernel bodies; the split precisions run the
    one-group bodies in the unified launches and the K-chunked split kernels per class, so those
    agree to the precision's bar), and one function of every bandlimit matches the oracle
    (bf16x3 bar 5e-5)."""
    torch = torch_cuda
    cplx = kind == "complex"
    Bs = [64, 8, 48, 16, 16, 32, 24, 64, 8, 48, 24, 32, 8, 16, 64, 48, 24, 32, 8, 64, 16, 24]
    Fs = [inputs.so3_coefficients(B, 1, 9100 + i, complex_=cplx, scale=lambda l: 1.0 / (1.0 + l))[0]
          for i, B in enumerate(Bs)]
    flat = np.concatenate(Fs)
    outs = {}
    for flags in (0, fftso3.FLAG_PER_CLASS_LAUNCHES, fftso3.FLAG_PER_CLASS_CONTRACTIONS):
        plan = fftso3.Plan([B - 1 for B in Bs], domain="so3", kind=kind, precision=precision, flags=flags)
        desc = plan.describe()
        assert ("unified" in desc) == (flags != fftso3.FLAG_PER_CLASS_LAUNCHES), desc
        assert ("contractions per class" in desc) == (flags == fftso3.FLAG_PER_CLASS_CONTRACTIONS), desc
        gdt, cdt = plan.dtypes()
        dF = torch.from_numpy(flat).to(cdt).cuda()
        plan.timing(True)
        g = plan.inverse(dF)
        c = plan.forward(g)
        torch.cuda.synchronize()
        st = plan.timing_read()
        plan.timing(False)
        if flags == 0:
            assert all(v[1] == 1 for k, v in st.items() if k in ("fft_fwd", "contract_fwd", "contract_inv", "fft_inv"))
        if flags == fftso3.FLAG_PER_CLASS_CONTRACTIONS:
            assert all(v[1] == 1 for k, v in st.items() if k in ("fft_fwd", "fft_inv"))
        outs[flags] = (g.cpu().numpy(), c.cpu().numpy(), [plan.grid_offset(i) for i in range(len(Bs))],
                       [plan.coef_offset(i) for i in range(len(Bs))])
        plan.close()
    g0, c0, go, co = outs[0]
    g1, c1, _, _ = outs[fftso3.FLAG_PER_CLASS_LAUNCHES]
    g2, c2, go2, co2 = outs[fftso3.FLAG_PER_CLASS_CONTRACTIONS]
    # unified FFT stages with per-class contractions: the per-class results (the same kernels at
    # every stage but the FFTs, whose unified and per-class forms agree bit for bit below l_max = 63)
    for i, B in enumerate(Bs):
        gs = slice(go[i], go[i] + (2 * B) ** 3)
        cs = slice(co[i], co[i] + len(Fs[i]))
        assert go2[i] == go[i] and co2[i] == co[i]
        if B < 64:
            assert np.array_equal(g1[gs], g2[gs]) and np.array_equal(c1[cs], c2[cs]), (i, B)
        else:
            assert rel_l2(g1[gs], g2[gs]) < 1e-6 and rel_l2(c1[cs], c2[cs]) < 1e-6, (i, B)
    for i, B in enumerate(Bs):
        gs = slice(go[i], go[i] + (2 * B) ** 3)
        cs = slice(co[i], co[i] + len(Fs[i]))
        if B < 64 and precision == "fp32":
            assert np.array_equal(g0[gs], g1[gs]) and np.array_equal(c0[cs], c1[cs]), (i, B)
        else:
            bar = 2e-5 if precision == "bf16x3" else 1e-6
            assert rel_l2(g0[gs], g1[gs]) < bar and rel_l2(c0[cs], c1[cs]) < bar, (i, B)
    seen = set()
    for i, B in enumerate(Bs):
        ci = c0[co[i]:co[i] + len(Fs[i])]
        tol = 5e-5 if precision == "bf16x3" else 1e-6 if B == 64 else 2e-6
        assert rel_l2(ci, Fs[i]) < tol, (i, B)
        if B in seen:
            continue
        seen.add(B)
        gi = g0[go[i]:go[i] + (2 * B) ** 3]
        ref = oracle.so3_inverse(B, Fs[i]) if cplx else oracle.so3_real_inverse(B, Fs[i])
        assert rel_l2(gi, ref) < tol, B


def test_unified_graph_replays(fftso3, torch_cuda):
    """The unified launches under graph capture: replays into zeroed outputs, with eager work
    between them, equal the eager run bit for bit."""
    torch = torch_cuda
    Bs = [8, 16, 24, 32, 48, 64, 16, 32]
    Fs = [inputs.so3_coefficients(B, 1, 9200 + i, complex_=False)[0] for i, B in enumerate(Bs)]
    plan = fftso3.Plan([B - 1 for B in Bs], domain="so3", kind="real", precision="fp32")
    dF = torch.from_numpy(np.concatenate(Fs)).to(torch.float32).cuda()
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        g0 = plan.inverse(dF, stream=s)
        c0 = plan.forward(g0, stream=s)
    s.synchronize()
    g = torch.empty_like(g0)
    c = torch.empty_like(c0)
    plan.set_graph(True)
    for rep in range(3):
        with torch.cuda.stream(s):
            g.zero_()
            c.zero_()
            junk = torch.randn(1 << 16, device="cuda") * 2.0
            plan.inverse(dF, g, stream=s)
            plan.forward(g, c, stream=s)
            junk += 1.0
        s.synchronize()
        assert torch.equal(g, g0) and torch.equal(c, c0), rep
    plan.set_graph(False)
    plan.close()
