"""Host code under sanitizers (CPU only): the FP64 oracle and the class layer's OpenMP sampling
step under AddressSanitizer + UndefinedBehaviorSanitizer and under ThreadSanitizer (OpenMP runtime
through its Archer tool), and the class layer's host checks (test_fdcl cpu) under ASan + UBSan.

A ThreadSanitizer setup that reports nothing proves nothing by itself, so the same build flags and
run options are first shown to catch a race planted in a scratch program of this test.
"""
import os
import subprocess

import pytest

from fftso3_amd import _build

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LLVM_LIB = os.path.join(_build.ROCM, "llvm", "lib")
SUPP = os.path.join(ROOT, "tests", "cpp", "lsan.supp")

pytestmark = pytest.mark.skipif(not os.path.exists(_build.CLANGXX), reason="ROCm clang++ not installed")


def _env(kind):
    env = dict(os.environ, OMP_NUM_THREADS="4")
    if kind == "tsan":
        env["TSAN_OPTIONS"] = "ignore_noninstrumented_modules=1 halt_on_error=1 exitcode=66"
        env["OMP_TOOL_LIBRARIES"] = os.path.join(LLVM_LIB, "libarcher.so")
    else:
        env["LSAN_OPTIONS"] = "suppressions=" + SUPP
        env["ASAN_OPTIONS"] = "halt_on_error=1"
    return env


@pytest.mark.parametrize("kind", ["asan", "tsan"])
def test_oracle_and_sampling_under_sanitizer(kind):
    exe = _build.build_sanitized(kind)
    r = subprocess.run([exe], env=_env(kind), capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-3000:])
    assert "FAIL" not in r.stdout
    assert "WARNING: ThreadSanitizer" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr


def test_class_layer_host_checks_under_asan():
    exe = _build.build_fdcl_asan()
    r = subprocess.run([exe, "cpu"], env=_env("asan"), capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-3000:])
    assert "FAIL" not in r.stdout and r.stdout.count("PASS") >= 20


def test_tsan_setup_catches_a_planted_race(tmp_path):
    src = tmp_path / "race.cpp"
    src.write_text("#include <cstdio>\nint main() { double s = 0;\n#pragma omp parallel for\n"
                   "for (int i = 0; i < 100000; i++) s += 1.0;\nstd::printf(\"%g\\n\", s); }\n")
    exe = tmp_path / "race"
    subprocess.run([_build.CLANGXX, "-g", "-O1", "-fopenmp"] + _build.SANITIZERS["tsan"] +
                   ["-o", str(exe), str(src), "-Wl,-rpath," + LLVM_LIB, "-L", LLVM_LIB], check=True)
    r = subprocess.run([str(exe)], env=_env("tsan"), capture_output=True, text=True, timeout=120)
    assert r.returncode == 66 and "data race" in r.stderr
