"""fftso3_amd -- SO(3) and S^2 harmonic transforms on AMD MI355X (gfx950).

Python binding of the C ABI in include/fftso3_amd.h (libfftso3_amd.so, built in-tree).  Device
memory and streams come from PyTorch (ROCm build); every transform runs in the hand-written HIP
kernels of the library.  There is no CPU fallback: if the library is missing the import of
`lib()` raises.
"""
import ctypes
import os

from . import inputs  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfftso3_amd.so")

DOMAIN = {"so3": 0, "s2": 1}
KIND = {"complex": 0, "real": 1}
PRECISION = {"fp64": 0, "fp32": 1, "bf16x3": 2, "bf16x6": 3}
FLAG_PER_CLASS_LAUNCHES = 1
FLAG_PER_CLASS_CONTRACTIONS = 2
STAGES = ("fft_fwd", "contract_fwd", "basis_fwd", "basis_inv", "contract_inv", "fft_inv")

_lib = None

_SIGS = {
    "fftso3_plan_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint]),
    "fftso3_plan_create_uneven": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int,
                                                 ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int,
                                                 ctypes.c_int, ctypes.c_uint]),
    "fftso3_plan_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "fftso3_plan_set_chunk": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "fftso3_plan_get_chunk": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]),
    "fftso3_plan_describe": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int]),
    "fftso3_plan_set_schedule": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "fftso3_plan_set_graph": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "fftso3_rotate_s2": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long,
                                        ctypes.c_void_p, ctypes.c_void_p]),
    "fftso3_evaluate": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long,
                                       ctypes.c_void_p, ctypes.c_void_p]),
    "fftso3_plan_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_int]),
    "fftso3_plan_grid_offset": (ctypes.c_long, [ctypes.c_void_p, ctypes.c_int]),
    "fftso3_plan_coef_offset": (ctypes.c_long, [ctypes.c_void_p, ctypes.c_int]),
    "fftso3_plan_grid_elems": (ctypes.c_long, [ctypes.c_void_p]),
    "fftso3_plan_coef_elems": (ctypes.c_long, [ctypes.c_void_p]),
    "fftso3_plan_workspace_bytes": (ctypes.c_long, [ctypes.c_void_p]),
    "fftso3_forward": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "fftso3_inverse": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "fftso3_inverse_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "fftso3_status_string": (ctypes.c_char_p, [ctypes.c_int]),
    "fftso3_last_error": (ctypes.c_char_p, []),
    "fftso3_debug_fail_alloc": (ctypes.c_int, [ctypes.c_int]),
    "fftso3_timing_enable": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "fftso3_timing_read": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(ctypes.c_long)]),
    "fftso3_so3_ncoef": (ctypes.c_long, [ctypes.c_int]),
    "fftso3_s2_ncoef": (ctypes.c_long, [ctypes.c_int]),
}

EXPORTED_SYMBOLS = tuple(_SIGS)


class FFTSO3Error(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("%s: %s" % (msg, status))
        self.status = status


def lib():
    """Load libfftso3_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libfftso3_amd.so is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(status):
    if status != 0:
        L = lib()
        raise FFTSO3Error(L.fftso3_status_string(status).decode(), L.fftso3_last_error().decode())


def so3_ncoef(l_max):
    return (l_max + 1) * (4 * (l_max + 1) ** 2 - 1) // 3


def s2_ncoef(l_max):
    return (l_max + 1) ** 2


class Plan:
    """A batch of functions on SO(3) (or S^2) with their bandlimits, kind and precision.

    forward(grid) -> coefficients, inverse(coefficients) -> grid, on torch CUDA tensors laid out
    as include/fftso3_amd.h describes (batch packed back to back; complex interleaved)."""

    def __init__(self, l_max, n_fn=1, domain="so3", kind="complex", precision="fp64", chunk=None, flags=0):
        L = lib()
        self.domain, self.kind, self.precision = domain, kind, precision
        h = ctypes.c_void_p()
        if isinstance(l_max, (list, tuple)):
            arr = (ctypes.c_int * len(l_max))(*l_max)
            _check(L.fftso3_plan_create_uneven(ctypes.byref(h), DOMAIN[domain], arr, len(l_max), KIND[kind],
                                               PRECISION[precision], int(flags)))
            self.l_max = list(l_max)
            self.n_fn = len(l_max)
        else:
            _check(L.fftso3_plan_create(ctypes.byref(h), DOMAIN[domain], int(l_max), int(n_fn), KIND[kind],
                                        PRECISION[precision], int(flags)))
            self.l_max = int(l_max)
            self.n_fn = int(n_fn)
        self._h = h
        if chunk is not None:
            self.set_chunk(chunk)

    # --- properties ---
    @property
    def handle(self):
        return self._h

    def grid_elems(self):
        return lib().fftso3_plan_grid_elems(self._h)

    def coef_elems(self):
        return lib().fftso3_plan_coef_elems(self._h)

    def grid_offset(self, fn):
        return lib().fftso3_plan_grid_offset(self._h, fn)

    def coef_offset(self, fn):
        return lib().fftso3_plan_coef_offset(self._h, fn)

    def workspace_bytes(self):
        return lib().fftso3_plan_workspace_bytes(self._h)

    def set_chunk(self, c):
        _check(lib().fftso3_plan_set_chunk(self._h, int(c)))

    def chunk(self):
        c = ctypes.c_int()
        _check(lib().fftso3_plan_get_chunk(self._h, ctypes.byref(c)))
        return c.value

    def dtypes(self):
        """(grid dtype, coefficient dtype) as torch dtypes."""
        import torch
        wide = self.precision == "fp64"
        if self.kind == "real":
            t = torch.float64 if wide else torch.float32
            return t, t
        t = torch.complex128 if wide else torch.complex64
        return t, t

    # --- transforms ---
    @staticmethod
    def _stream(stream):
        import torch
        s = stream if stream is not None else torch.cuda.current_stream()
        return ctypes.c_void_p(s.cuda_stream)

    def _validate(self, t, n, dtype, what):
        if not t.is_cuda:
            raise ValueError("%s must be a device tensor" % what)
        if not t.is_contiguous():
            raise ValueError("%s must be contiguous" % what)
        if t.dtype != dtype:
            raise TypeError("%s has dtype %s, expected %s" % (what, t.dtype, dtype))
        if t.numel() != n:
            raise ValueError("%s has %d elements, expected %d" % (what, t.numel(), n))

    def forward(self, grid, coef=None, stream=None):
        import torch
        gdt, cdt = self.dtypes()
        self._validate(grid, self.grid_elems(), gdt, "grid")
        if coef is None:
            coef = torch.empty(self.coef_elems(), dtype=cdt, device=grid.device)
        self._validate(coef, self.coef_elems(), cdt, "coefficients")
        _check(lib().fftso3_forward(self._h, ctypes.c_void_p(grid.data_ptr()), ctypes.c_void_p(coef.data_ptr()),
                                    self._stream(stream)))
        return coef

    def inverse(self, coef, grid=None, stream=None):
        import torch
        gdt, cdt = self.dtypes()
        self._validate(coef, self.coef_elems(), cdt, "coefficients")
        if grid is None:
            grid = torch.empty(self.grid_elems(), dtype=gdt, device=coef.device)
        self._validate(grid, self.grid_elems(), gdt, "grid")
        _check(lib().fftso3_inverse(self._h, ctypes.c_void_p(coef.data_ptr()), ctypes.c_void_p(grid.data_ptr()),
                                    self._stream(stream)))
        return grid

    def evaluate(self, coef, angles, fn=0, out=None, stream=None):
        """Values of function fn at arbitrary points: angles is a float64 CUDA tensor of shape
        [n, 3] (alpha, beta, gamma) for SO(3) or [n, 2] (theta, phi) for S^2."""
        import torch
        gdt, cdt = self.dtypes()
        self._validate(coef, self.coef_elems(), cdt, "coefficients")
        per = 2 if self.domain == "s2" else 3
        if angles.dtype != torch.float64 or not angles.is_cuda or not angles.is_contiguous() or angles.numel() % per:
            raise ValueError("angles must be a contiguous float64 CUDA tensor with %d values per point" % per)
        n = angles.numel() // per
        vdt = gdt
        if out is None:
            out = torch.empty(n, dtype=vdt, device=coef.device)
        self._validate(out, n, vdt, "values")
        _check(lib().fftso3_evaluate(self._h, int(fn), ctypes.c_void_p(coef.data_ptr()),
                                     ctypes.c_void_p(angles.data_ptr()), n, ctypes.c_void_p(out.data_ptr()),
                                     self._stream(stream)))
        return out

    def rotate_s2(self, F, angles, out=None, stream=None):
        """S^2 coefficients F ((l_max+1)^2 float64, complex128 for complex plans) rotated by each
        rotation of angles (float64 [n, 3]: alpha, beta, gamma): G_i^l = U^l(R_i) F^l (real plans)
        or D^l(R_i) F^l (complex plans).  Needs an SO(3) plan of one bandlimit."""
        import torch
        dt = torch.float64 if self.kind == "real" else torch.complex128
        nc = (self.l_max + 1) ** 2
        self._validate(F, nc, dt, "coefficients")
        if angles.dtype != torch.float64 or not angles.is_cuda or not angles.is_contiguous() or angles.numel() % 3:
            raise ValueError("angles must be a contiguous float64 CUDA tensor with 3 values per rotation")
        n = angles.numel() // 3
        if out is None:
            out = torch.empty(n * nc, dtype=dt, device=F.device)
        self._validate(out, n * nc, dt, "rotated coefficients")
        _check(lib().fftso3_rotate_s2(self._h, ctypes.c_void_p(F.data_ptr()), ctypes.c_void_p(angles.data_ptr()), n,
                                      ctypes.c_void_p(out.data_ptr()), self._stream(stream)))
        return out

    def set_graph(self, on=True):
        """Capture each (direction, input, output, stream) into a hipGraph and replay it."""
        _check(lib().fftso3_plan_set_graph(self._h, 1 if on else 0))

    def set_schedule(self, klo=0, gil=1, split=1):
        """Workspace row blocking, group interleave and group split of the batch-wide kernels."""
        _check(lib().fftso3_plan_set_schedule(self._h, int(klo), int(gil), int(split)))

    def describe(self):
        """The current schedule, one line per bandlimit class."""
        buf = ctypes.create_string_buffer(4096)
        _check(lib().fftso3_plan_describe(self._h, buf, len(buf)))
        return buf.value.decode()

    def autotune(self, grid=None, coef=None, stream=None, reps=2):
        """Pick the fastest schedule for this plan; grid / coef (device tensors) are clobbered."""
        import torch
        gdt, cdt = self.dtypes()
        if grid is None:
            grid = torch.empty(self.grid_elems(), dtype=gdt, device="cuda")
        if coef is None:
            coef = torch.zeros(self.coef_elems(), dtype=cdt, device="cuda")
        self._validate(grid, self.grid_elems(), gdt, "grid")
        self._validate(coef, self.coef_elems(), cdt, "coefficients")
        _check(lib().fftso3_plan_autotune(self._h, ctypes.c_void_p(grid.data_ptr()), ctypes.c_void_p(coef.data_ptr()),
                                          self._stream(stream), int(reps)))

    # --- timing ---
    def timing(self, on=True):
        _check(lib().fftso3_timing_enable(self._h, 1 if on else 0))

    def timing_read(self):
        ms = (ctypes.c_double * 6)()
        cnt = (ctypes.c_long * 6)()
        _check(lib().fftso3_timing_read(self._h, ms, cnt))
        return {s: (ms[i], cnt[i]) for i, s in enumerate(STAGES)}

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().fftso3_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
