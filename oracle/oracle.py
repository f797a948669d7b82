"""ctypes binding of the FP64 CPU oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, never by the fftso3_amd package.  Arrays are numpy; complex arrays are
complex128 and are passed to C as interleaved doubles.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)


def build():
    subprocess.run(["make", "-s", "-B", "-C", _HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        i, d, c_long = ctypes.c_int, ctypes.c_double, ctypes.c_long
        sig = {
            "or_so3_ncoef": (c_long, [i]),
            "or_s2_ncoef": (c_long, [i]),
            "or_so3_index": (i, [i, i, i]),
            "or_so3_weights": (None, [i, _dp]),
            "or_s2_weights": (None, [i, _dp]),
            "or_wigner_d_ref": (None, [d, i, _dp]),
            "or_wigner_d_stable": (None, [d, i, _dp]),
            "or_legendre": (None, [d, i, _dp]),
            "or_fft": (None, [i, _dp, _dp]),
            "or_so3_forward": (None, [i, _dp, _dp, i]),
            "or_so3_inverse": (None, [i, _dp, _dp, i]),
            "or_so3_inverse_point": (None, [i, _dp, d, d, d, _dp]),
            "or_so3_forward_direct": (None, [i, _dp, _dp]),
            "or_so3_real_forward": (None, [i, _dp, _dp, i]),
            "or_so3_real_forward_T": (None, [i, _dp, _dp, i]),
            "or_so3_real_inverse": (None, [i, _dp, _dp, i]),
            "or_so3_real_inverse_point": (d, [i, _dp, d, d, d]),
            "or_so3_real_harmonics": (None, [i, d, d, d, _dp]),
            "or_so3_T": (None, [i, _dp]),
            "or_so3_c2r": (None, [i, _dp, _dp]),
            "or_so3_r2c": (None, [i, _dp, _dp]),
            "or_s2_forward": (None, [i, _dp, _dp, i]),
            "or_s2_inverse": (None, [i, _dp, _dp, i]),
            "or_s2_inverse_point": (None, [i, _dp, d, d, _dp]),
            "or_s2_real_forward": (None, [i, _dp, _dp, i]),
            "or_s2_real_inverse": (None, [i, _dp, _dp, i]),
            "or_s2_real_inverse_point": (d, [i, _dp, d, d]),
            "or_euler323_to_R": (None, [d, d, d, _dp]),
            "or_R_to_euler323": (None, [_dp, _dp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def _c(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def _r(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def so3_ncoef(B):
    return B * (4 * B * B - 1) // 3


def s2_ncoef(B):
    return B * B


def so3_index(l, m, n):
    return l * (4 * l * l - 1) // 3 + (m + l) + (n + l) * (2 * l + 1)


def so3_weights(B):
    w = np.zeros(2 * B)
    lib().or_so3_weights(B, _p(w))
    return w


def s2_weights(B):
    w = np.zeros(2 * B)
    lib().or_s2_weights(B, _p(w))
    return w


def wigner_d_ref(beta, L):
    d = np.zeros(so3_ncoef(L + 1))
    lib().or_wigner_d_ref(float(beta), L, _p(d))
    return d


def wigner_d_stable(beta, L):
    d = np.zeros(so3_ncoef(L + 1))
    lib().or_wigner_d_stable(float(beta), L, _p(d))
    return d


def legendre(x, L):
    nP = np.zeros((L + 1) ** 2)
    lib().or_legendre(float(x), L, _p(nP))
    return nP


def fft(x):
    x = _c(x)
    y = np.zeros_like(x)
    lib().or_fft(x.size, _p(x), _p(y))
    return y


def so3_forward(B, grid, nthreads=0):
    g = _c(grid).reshape(-1)
    assert g.size == (2 * B) ** 3
    F = np.zeros(so3_ncoef(B), np.complex128)
    lib().or_so3_forward(B, _p(g), _p(F), nthreads)
    return F


def so3_inverse(B, F, nthreads=0):
    F = _c(F)
    assert F.size == so3_ncoef(B)
    g = np.zeros((2 * B) ** 3, np.complex128)
    lib().or_so3_inverse(B, _p(F), _p(g), nthreads)
    return g.reshape(2 * B, 2 * B, 2 * B)


def so3_inverse_point(B, F, a, b, g):
    F = _c(F)
    out = np.zeros(2)
    lib().or_so3_inverse_point(B, _p(F), a, b, g, _p(out))
    return complex(out[0], out[1])


def so3_forward_direct(B, grid):
    g = _c(grid).reshape(-1)
    F = np.zeros(so3_ncoef(B), np.complex128)
    lib().or_so3_forward_direct(B, _p(g), _p(F))
    return F


def so3_real_forward(B, grid, nthreads=0):
    g = _r(grid).reshape(-1)
    F = np.zeros(so3_ncoef(B))
    lib().or_so3_real_forward(B, _p(g), _p(F), nthreads)
    return F


def so3_real_forward_T(B, grid, nthreads=0):
    g = _r(grid).reshape(-1)
    F = np.zeros(so3_ncoef(B))
    lib().or_so3_real_forward_T(B, _p(g), _p(F), nthreads)
    return F


def so3_real_inverse(B, Fr, nthreads=0):
    Fr = _r(Fr)
    g = np.zeros((2 * B) ** 3)
    lib().or_so3_real_inverse(B, _p(Fr), _p(g), nthreads)
    return g.reshape(2 * B, 2 * B, 2 * B)


def so3_real_inverse_point(B, Fr, a, b, g):
    return lib().or_so3_real_inverse_point(B, _p(_r(Fr)), a, b, g)


def so3_real_harmonics(L, a, b, g):
    U = np.zeros(so3_ncoef(L + 1))
    lib().or_so3_real_harmonics(L, a, b, g, _p(U))
    return U


def so3_T(L):
    T = np.zeros(so3_ncoef(L + 1), np.complex128)
    lib().or_so3_T(L, _p(T))
    return T


def so3_c2r(B, Fc):
    Fc = _c(Fc)
    Fr = np.zeros(so3_ncoef(B))
    lib().or_so3_c2r(B, _p(Fc), _p(Fr))
    return Fr


def so3_r2c(B, Fr):
    Fr = _r(Fr)
    Fc = np.zeros(so3_ncoef(B), np.complex128)
    lib().or_so3_r2c(B, _p(Fr), _p(Fc))
    return Fc


def s2_forward(B, grid):
    g = _c(grid).reshape(-1)
    F = np.zeros(B * B, np.complex128)
    lib().or_s2_forward(B, _p(g), _p(F), 0)
    return F


def s2_inverse(B, F):
    F = _c(F)
    g = np.zeros((2 * B) ** 2, np.complex128)
    lib().or_s2_inverse(B, _p(F), _p(g), 0)
    return g.reshape(2 * B, 2 * B)


def s2_inverse_point(B, F, theta, phi):
    out = np.zeros(2)
    lib().or_s2_inverse_point(B, _p(_c(F)), theta, phi, _p(out))
    return complex(out[0], out[1])


def s2_real_forward(B, grid):
    g = _r(grid).reshape(-1)
    F = np.zeros(B * B)
    lib().or_s2_real_forward(B, _p(g), _p(F), 0)
    return F


def s2_real_inverse(B, Fr):
    g = np.zeros((2 * B) ** 2)
    lib().or_s2_real_inverse(B, _p(_r(Fr)), _p(g), 0)
    return g.reshape(2 * B, 2 * B)


def s2_real_inverse_point(B, Fr, theta, phi):
    return lib().or_s2_real_inverse_point(B, _p(_r(Fr)), theta, phi)


def euler323_to_R(a, b, g):
    R = np.zeros(9)
    lib().or_euler323_to_R(a, b, g, _p(R))
    return R.reshape(3, 3)


def R_to_euler323(R):
    R = _r(R).reshape(-1)
    out = np.zeros(3)
    lib().or_R_to_euler323(_p(R), _p(out))
    return out
