"""GPU parity of the alternative kernels behind the schedule switches.

The FFTSO3_* switches are read once per process, so one child process runs with every switch set
to its alternative (one-slice N = 128 FFTs, 32x32 grouped forward and one-group inverse
contractions, 4-wave mxk inverse, 1024-thread whole-slice FFTs, whole-block table forward) and
reports round-trip and oracle errors as JSON.  Tolerances as in test_gpu_parity: 2e-6 relative-L2
for FP32, 1e-10 for FP64.
"""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SWITCHES = {
    "FFTSO3_FFT128_PAIRS": "0",
    "FFTSO3_FFT128_PAIRS_INV": "0",
    "FFTSO3_FWD_MX16": "0",
    "FFTSO3_INV_GROUPED": "0",
    "FFTSO3_MXK_INV8": "0",
    "FFTSO3_FUSED_THREADS": "0",
    "FFTSO3_FWD_TAB2": "0",
}

CHILD = r"""
import json, sys
import numpy as np
import torch
import fftso3_amd
from fftso3_amd import inputs
from oracle import oracle

def rel(a, b):
    a = np.asarray(a).reshape(-1); b = np.asarray(b).reshape(-1)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))

def run(domain, B, n, kind, prec, F):
    plan = fftso3_amd.Plan(B - 1, n, domain, kind, prec)
    gdt, cdt = plan.dtypes()
    g = plan.inverse(torch.from_numpy(np.ascontiguousarray(F).reshape(-1)).to(cdt).cuda())
    c = plan.forward(g)
    torch.cuda.synchronize()
    return g.cpu().numpy().reshape(n, -1), c.cpu().numpy().reshape(n, -1)

out = {}
# SO(3) l_max = 63, FP32: one-slice FFTs, 32x32 grouped forward, one-group inverse (group interleave 1)
B, n = 64, 37
F = inputs.so3_coefficients(B, n, 4100, scale=lambda l: 1.0 / (1.0 + l))
g, c = run("so3", B, n, "complex", "fp32", F)
out["so3_b64_fp32_roundtrip"] = rel(c, F)
out["so3_b64_fp32_grid"] = max(rel(g[i], oracle.so3_inverse(B, F[i])) for i in (0, n - 1))
# S^2 l_max = 127, FP32: 4-wave mxk inverse
B, n = 128, 5
F = inputs.s2_coefficients(B, n, 4200)
g, c = run("s2", B, n, "complex", "fp32", F)
out["s2_b128_fp32_roundtrip"] = rel(c, F)
out["s2_b128_fp32_grid"] = max(rel(g[i], oracle.s2_inverse(B, F[i])) for i in range(n))
# SO(3) l_max = 20, FP64: 1024-thread whole-slice FFTs
B, n = 21, 3
F = inputs.so3_coefficients(B, n, 4300)
g, c = run("so3", B, n, "complex", "fp64", F)
out["so3_b21_fp64_roundtrip"] = rel(c, F)
out["so3_b21_fp64_grid"] = max(rel(g[i], oracle.so3_inverse(B, F[i])) for i in range(n))
# S^2 l_max = 151, FP32: forward on the d table with the whole spectrum block in LDS
B, n = 152, 3
F = inputs.s2_coefficients(B, n, 4400, scale=lambda l: 1.0 / (1.0 + l))
g, c = run("s2", B, n, "complex", "fp32", F)
out["s2_b152_fp32_roundtrip"] = rel(c, F)
out["s2_b152_fp32_grid"] = max(rel(g[i], oracle.s2_inverse(B, F[i])) for i in range(n))
print("RESULT " + json.dumps(out), flush=True)
"""


def test_schedule_switch_alternatives(fftso3, torch_cuda, oracle):
    env = dict(os.environ, **SWITCHES)
    env["PYTHONPATH"] = ROOT + os.pathsep + env.get("PYTHONPATH", "")
    r = subprocess.run([sys.executable, "-c", CHILD], cwd=ROOT, env=env, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    line = [l for l in r.stdout.splitlines() if l.startswith("RESULT ")][-1]
    res = json.loads(line[len("RESULT "):])
    for k, v in res.items():
        tol = 1e-10 if "fp64" in k else 2e-6
        assert v < tol, (k, v)
