"""The fdcl:: C++ class layer (fftso3_amd/cpp): host FP64 functions against the oracle and the
reference-style self checks (tests/cpp/test_fdcl.cpp); GPU round trips through the class layer.
"""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "test_fdcl")
SOURCES = [os.path.join(ROOT, "tests", "cpp", "test_fdcl.cpp")] + [
    os.path.join(ROOT, "fftso3_amd", "cpp", d, f)
    for d in ("include", "src")
    for f in sorted(os.listdir(os.path.join(ROOT, "fftso3_amd", "cpp", d)))
]


@pytest.fixture(scope="module")
def exe():
    from fftso3_amd import _build
    lib = _build.FDCL_LIB
    newest = max(os.path.getmtime(s) for s in SOURCES)
    if not os.path.exists(lib) or os.path.getmtime(lib) < newest:
        _build.build_fdcl()
    if not os.path.exists(EXE) or os.path.getmtime(EXE) < max(newest, os.path.getmtime(lib)):
        _build.build_fdcl_test()
    return EXE


def _run(exe, *args, timeout=300):
    r = subprocess.run([exe] + [str(a) for a in args], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=timeout)
    return r.returncode, r.stdout


def test_class_layer_host_checks(exe):
    rc, out = _run(exe, "cpu")
    assert rc == 0, out
    assert "FAIL" not in out


@pytest.mark.parametrize("beta,L", [(0.3, 12), (1.5707963267948966, 30), (2.9, 64), (0.01, 40)])
def test_class_layer_wigner_d_vs_oracle(exe, oracle, beta, L):
    """wigner_d of the class layer (its own recursion) against the oracle's long-double recursion."""
    rc, out = _run(exe, "dump", beta, L)
    assert rc == 0
    d = np.array([float(x) for x in out.split()])
    ref = oracle.wigner_d_stable(beta, L)
    assert d.shape == ref.shape
    assert np.max(np.abs(d - ref)) < 1e-12


@pytest.mark.parametrize("l1,l2", [(1, 1), (2, 5), (3, 3), (6, 4), (0, 3)])
def test_clebsch_gordan_vs_sympy(exe, l1, l2):
    """Clebsch-Gordan coefficients of the class layer against sympy's exact CG (Condon-Shortley
    phase): every coefficient of the (l1, l2) block, zeros included."""
    from sympy import S
    from sympy.physics.quantum.cg import CG
    rc, out = _run(exe, "cg", l1, l2)
    assert rc == 0, out
    rows = [line.split() for line in out.strip().splitlines()]
    expected = sum(1 for l in range(abs(l1 - l2), l1 + l2 + 1) for m in range(-l, l + 1)
                   for m1 in range(-l1, l1 + 1) if abs(m - m1) <= l2)
    assert len(rows) == expected
    err = 0.0
    for l, m, m1, m2, v in rows:
        ref = float(CG(S(l1), S(m1), S(l2), S(m2), S(l), S(m)).doit())
        err = max(err, abs(float(v) - ref))
    assert err < 1e-13


@pytest.mark.parametrize("x,L", [(0.37, 20), (-0.93, 64), (0.0, 15)])
def test_class_layer_legendre_vs_oracle(exe, oracle, x, L):
    rc, out = _run(exe, "dumpnp", x, L)
    assert rc == 0
    nP = np.array([float(v) for v in out.split()])
    ref = oracle.legendre(x, L)
    assert np.max(np.abs(nP - ref)) < 1e-12


@pytest.mark.gpu
def test_class_layer_gpu_round_trips(exe, torch_cuda):
    """check_transform at l_max = 5 (the reference's test), grid vs pointwise inverse, batches, FP32."""
    rc, out = _run(exe, "gpu")
    print(out)
    assert rc == 0, out
    assert "FAIL" not in out
