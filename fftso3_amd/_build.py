"""Build recipes for the in-tree shared libraries (hipcc for gfx950, g++ for host code)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libfftso3_amd.so")
FDCL_LIB = os.path.join(HERE, "libfftso3_fdcl.so")

HIP_SOURCES = [os.path.join(HERE, "csrc", "api.hip")]   # includes every csrc/*.hpp
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # keep the compiler from forming packed-FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build_hip(verbose=False):
    """Compile the HIP kernels and the C ABI into fftso3_amd/libfftso3_amd.so."""
    out = _run([HIPCC] + HIP_FLAGS + ["-I", os.path.join(ROOT, "include"), "-o", LIB] + HIP_SOURCES)
    if verbose and out.strip():
        print(out)
    return LIB


def build_fdcl(verbose=False):
    """Compile the C++ class layer (fdcl:: drop-in) into fftso3_amd/libfftso3_fdcl.so."""
    src_dir = os.path.join(HERE, "cpp", "src")
    srcs = sorted(os.path.join(src_dir, f) for f in os.listdir(src_dir) if f.endswith(".cpp"))
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-D__HIP_PLATFORM_AMD__",
           "-I", os.path.join(HERE, "cpp", "include"), "-I", os.path.join(ROOT, "include"),
           "-I", os.path.join(ROCM, "include"),
           "-o", FDCL_LIB] + srcs + ["-L", HERE, "-lfftso3_amd", "-L", os.path.join(ROCM, "lib"), "-lamdhip64",
                                     "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + os.path.join(ROCM, "lib")]
    out = _run(cmd)
    if verbose and out.strip():
        print(out)
    return FDCL_LIB


FDCL_TEST = os.path.join(ROOT, "tests", "cpp", "test_fdcl")


def build_fdcl_test(verbose=False):
    """Compile the class layer's check program tests/cpp/test_fdcl (links libfftso3_fdcl.so)."""
    cmd = ["g++", "-O2", "-std=c++17", "-fopenmp", "-D__HIP_PLATFORM_AMD__",
           "-I", os.path.join(HERE, "cpp", "include"), "-I", os.path.join(ROOT, "include"),
           "-I", os.path.join(ROCM, "include"), "-o", FDCL_TEST,
           os.path.join(ROOT, "tests", "cpp", "test_fdcl.cpp"),
           "-L", HERE, "-lfftso3_fdcl", "-lfftso3_amd", "-L", os.path.join(ROCM, "lib"), "-lamdhip64",
           "-Wl,-rpath," + HERE, "-Wl,-rpath," + os.path.join(ROCM, "lib")]
    out = _run(cmd)
    if verbose and out.strip():
        print(out)
    return FDCL_TEST


def build_all(verbose=False):
    build_hip(verbose)
    if os.path.isdir(os.path.join(HERE, "cpp", "src")) and any(
            f.endswith(".cpp") for f in os.listdir(os.path.join(HERE, "cpp", "src"))):
        build_fdcl(verbose)
        build_fdcl_test(verbose)


# ---------------------------------------------------------------------------------------------
# sanitizer builds of the host code (CPU only; clang from the ROCm toolchain, no device code)
# ---------------------------------------------------------------------------------------------
CLANGXX = os.path.join(ROCM, "llvm", "bin", "clang++")
CLANG = os.path.join(ROCM, "llvm", "bin", "clang")
SAN_DIR = os.path.join(ROOT, "tests", "cpp")
SANITIZERS = {"asan": ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"],
              "tsan": ["-fsanitize=thread"]}


def build_sanitized(kind):
    """tests/cpp/test_sanitize_<kind>: the oracle and the class layer's sampling step under the
    sanitizer (asan: AddressSanitizer + UndefinedBehaviorSanitizer; tsan: ThreadSanitizer)."""
    flags = SANITIZERS[kind] + ["-g", "-O1", "-fopenmp"]
    out = os.path.join(SAN_DIR, "test_sanitize_" + kind)
    obj = out + "_oracle.o"
    _run([CLANG, "-std=gnu11", "-c", "-o", obj, os.path.join(ROOT, "oracle", "oracle.c")] + flags)
    _run([CLANGXX, "-std=c++17", "-o", out, os.path.join(SAN_DIR, "test_sanitize.cpp"),
          os.path.join(HERE, "cpp", "src", "fdcl_sampling.cpp"), obj, "-I", os.path.join(HERE, "cpp", "include"),
          "-lm", "-Wl,-rpath," + os.path.join(ROCM, "llvm", "lib"), "-L", os.path.join(ROCM, "llvm", "lib")] + flags)
    return out


def build_fdcl_asan():
    """tests/cpp/test_fdcl_asan: the class layer's host checks (test_fdcl cpu) under
    AddressSanitizer + UndefinedBehaviorSanitizer."""
    src_dir = os.path.join(HERE, "cpp", "src")
    srcs = sorted(os.path.join(src_dir, f) for f in os.listdir(src_dir) if f.endswith(".cpp"))
    out = os.path.join(SAN_DIR, "test_fdcl_asan")
    _run([CLANGXX, "-std=c++17", "-g", "-O1", "-fopenmp", "-D__HIP_PLATFORM_AMD__"] + SANITIZERS["asan"] +
         ["-I", os.path.join(HERE, "cpp", "include"), "-I", os.path.join(ROOT, "include"),
          "-I", os.path.join(ROCM, "include"), "-o", out, os.path.join(SAN_DIR, "test_fdcl.cpp")] + srcs +
         ["-L", HERE, "-lfftso3_amd", "-L", os.path.join(ROCM, "lib"), "-lamdhip64",
          "-L", os.path.join(ROCM, "llvm", "lib"),
          "-Wl,-rpath," + HERE, "-Wl,-rpath," + os.path.join(ROCM, "lib"), "-Wl,-rpath," + os.path.join(ROCM, "llvm", "lib")])
    return out
