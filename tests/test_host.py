"""CPU tests of the host side: the C ABI library loads and exports every declared symbol, argument
validation fails cleanly before any device work, and the seeded input generator is stable."""
import ctypes
import os
import re

import numpy as np
import pytest

import fftso3_amd
from fftso3_amd import inputs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "fftso3_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fftso3_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = fftso3_amd.lib()
    names = declared_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(fftso3_amd.EXPORTED_SYMBOLS)


def test_invalid_arguments_return_einval():
    L = fftso3_amd.lib()
    h = ctypes.c_void_p()
    assert L.fftso3_plan_create(ctypes.byref(h), 7, 5, 1, 0, 0, 0) == 1          # bad domain
    assert L.fftso3_plan_create(ctypes.byref(h), 0, 5, 0, 0, 0, 0) == 1          # no functions
    assert L.fftso3_plan_create(ctypes.byref(h), 0, -1, 1, 0, 0, 0) == 1         # negative l_max
    assert L.fftso3_plan_create(ctypes.byref(h), 0, 300, 1, 0, 0, 0) == 1        # beyond B = 256
    assert L.fftso3_plan_create(ctypes.byref(h), 0, 5, 1, 3, 0, 0) == 1          # bad kind
    assert L.fftso3_plan_create(ctypes.byref(h), 0, 5, 1, 0, 9, 0) == 1          # bad precision
    assert L.fftso3_plan_create(None, 0, 5, 1, 0, 0, 0) == 1                     # NULL out pointer
    assert L.fftso3_plan_create(ctypes.byref(h), 0, 5, 1, 0, 0, 6) == 1          # unknown flags
    assert b"precision" in L.fftso3_last_error() or len(L.fftso3_last_error()) > 0
    assert L.fftso3_forward(None, None, None, None) == 1
    assert L.fftso3_plan_destroy(None) == 1
    assert L.fftso3_status_string(0) == b"FFTSO3_OK"
    assert L.fftso3_so3_ncoef(63) == 349504
    assert L.fftso3_s2_ncoef(127) == 16384
    assert L.fftso3_plan_grid_elems(None) == -1


def test_python_plan_rejects_bad_arguments():
    with pytest.raises(fftso3_amd.FFTSO3Error):
        fftso3_amd.Plan(-3, 1)


def test_splitmix64_known_values():
    """SplitMix64 reference outputs for seed 0 (the published test vector of the generator:
    0xe220a8397b1dcdaf, 0x6e789e6aa1b965f4, 0x06c45d188009454f)."""
    z = inputs.splitmix64(0, 3)
    assert [int(x) for x in z] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_normals_are_standard():
    x = inputs.normal(42, 200000)
    assert abs(x.mean()) < 0.01 and abs(x.std() - 1) < 0.01


def test_draw_permutation_is_packed_layout():
    B = 5
    perm = inputs.so3_draw_permutation(B)
    assert sorted(perm) == list(range(inputs.so3_ncoef(B)))
    i = 0
    for l in range(B):
        for m in range(-l, l + 1):
            for n in range(-l, l + 1):
                assert perm[i] == l * (4 * l * l - 1) // 3 + (m + l) + (n + l) * (2 * l + 1)
                i += 1


def test_workload_shapes_and_scales():
    F = inputs.workload_complex_b16()
    assert F.shape == (1, 5456) and F.dtype == np.complex128
    Fr = inputs.workload_real_b64()
    assert Fr.shape == (1, 349504) and Fr.dtype == np.float64
    Bs, coefs = inputs.workload_uneven(16)
    assert all(B in inputs.UNEVEN_BANDLIMITS for B in Bs)
    assert all(len(c) == inputs.so3_ncoef(B) for B, c in zip(Bs, coefs))
    S = inputs.s2_coefficients(4, 2, 5)
    assert S.shape == (2, 16)
