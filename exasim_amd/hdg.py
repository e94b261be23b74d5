"""ctypes mirror of the C ABI in include/exasim_hdg.h.

The solver runs in libexahip.so on the GPU; this module only marshals
arguments.  It raises when the library is missing or a call fails -- there is
no CPU fallback.  Stage methods carry the names of the reference functions they
replace (``hdgAssembleLinearSystem`` -> ``assemble``, ``hdgMatVec`` ->
``matvec``, ``NewtonSolver`` -> ``solve``).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libexahip.so")
_lib = None

# (name, restype, argtypes) for every entry point of include/exasim_hdg.h
_P = ctypes.c_void_p
_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_L = ctypes.POINTER(ctypes.c_long)
SIGNATURES = [
    ("exa_create", ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int,
                                  ctypes.POINTER(_P)]),
    ("exa_destroy", None, [_P]),
    ("exa_last_error", ctypes.c_char_p, [_P]),
    ("exa_sizes", ctypes.c_int, [_P, _I, ctypes.c_int]),
    ("exa_set_option", ctypes.c_int, [_P, ctypes.c_char_p, ctypes.c_double]),
    ("exa_solve", ctypes.c_int, [_P, _I, _I, _D]),
    ("exa_reset", ctypes.c_int, [_P]),
    ("exa_get_q", ctypes.c_int, [_P]),
    ("exa_assemble", ctypes.c_int, [_P]),
    ("exa_residual", ctypes.c_int, [_P, _D, _D]),
    ("exa_matvec", ctypes.c_int, [_P, _P, _P]),
    ("exa_matvec_noalloc", ctypes.c_int, [_P, _P, _P]),
    ("exa_precond", ctypes.c_int, [_P, _P]),
    ("exa_gmres", ctypes.c_int, [_P, _P, _P, _I]),
    ("exa_get_array", ctypes.c_long, [_P, ctypes.c_char_p, _D, ctypes.c_long]),
    ("exa_set_array", ctypes.c_int, [_P, ctypes.c_char_p, _D, ctypes.c_long]),
    ("exa_device_ptr", ctypes.c_int, [_P, ctypes.c_char_p, ctypes.POINTER(_P), _L]),
    ("exa_write_output", ctypes.c_int, [_P]),
    ("exa_timings", ctypes.c_int, [_P, _D, _L, ctypes.c_int]),
    ("exa_reset_timings", ctypes.c_int, [_P]),
    ("exa_sync", ctypes.c_int, [_P]),
    ("exa_gemm_tuning", ctypes.c_int, [ctypes.c_char_p, ctypes.c_int]),
    ("exa_batch_create", ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p), ctypes.c_char_p,
                                        ctypes.c_int, ctypes.POINTER(_P)]),
    ("exa_batch_size", ctypes.c_int, [_P]),
    ("exa_batch_stats", ctypes.c_int, [_P, _I, _I, _D, ctypes.c_int]),
    ("exa_batch_get", ctypes.c_long, [_P, ctypes.c_int, ctypes.c_char_p, _D, ctypes.c_long]),
    ("exa_batch_step", ctypes.c_int, [_P, _I, ctypes.c_int, _I]),
    ("exa_batch_refill", ctypes.c_int, [_P, ctypes.c_int, ctypes.c_char_p]),
    ("libmain", ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p)]),
]
TIMER_CLASSES = ["matvec", "precond", "ortho", "assembly", "lu", "other"]


def gemm_tuning():
    """The GEMM tile autotuner's choices so far (``exa_gemm_tuning``): a list of dicts with the
    product shape, the chosen tile, the number of products in the timed call and each tile's time
    in microseconds."""
    lib = load()
    n = lib.exa_gemm_tuning(ctypes.create_string_buffer(1), 1)
    buf = ctypes.create_string_buffer(n + 1)
    lib.exa_gemm_tuning(buf, n + 1)
    out = []
    for line in buf.value.decode().splitlines():
        f = line.split()
        out.append(dict(M=int(f[0]), N=int(f[1]), K=int(f[2]), log2_batch=int(f[3]), fp32=int(f[4]),
                        storage=int(f[5]), tile=f[6], batch=int(f[7]), us=[float(x) for x in f[8:]]))
    return out


class ExaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("libexahip error %d: %s" % (code, msg))
        self.code = code


def load(path=LIB_PATH):
    """Load libexahip.so (raises when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError("libexahip.so not found at %s; build it with "
                           "`python -m exasim_amd.build`" % path)
    lib = ctypes.CDLL(path)
    for name, res, args in SIGNATURES:
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _dp(a):
    return a.ctypes.data_as(_D)


# options applied to every handle this process creates (exa_set_option), before any other
DEFAULT_OPTIONS = {}


class HDGSolver:
    """One HDG problem resident on the GPU (``exa_handle``)."""

    def __init__(self, datain, model, dataout=None, device=-1):
        lib = load()
        self._lib = lib
        h = _P()
        rc = lib.exa_create(datain.encode(), (dataout or "").encode() if dataout else None,
                            model.encode(), device, ctypes.byref(h))
        if rc != 0:
            raise ExaError(rc, lib.exa_last_error(None).decode())
        self._h = h
        for k, v in DEFAULT_OPTIONS.items():
            self.set_option(k, v)
        self.sizes = self.state()

    @classmethod
    def batch(cls, datains, model, device=-1):
        """A batch of independent instances (``exa_batch_create``): one handle, shared launches,
        per-instance Newton/GMRES.  ``solve()`` solves them all; ``instance_stats()`` and
        ``get_instance()`` give each instance's counts and arrays."""
        lib = load()
        self = cls.__new__(cls)
        self._lib = lib
        arr = (ctypes.c_char_p * len(datains))(*[d.encode() for d in datains])
        h = _P()
        rc = lib.exa_batch_create(len(datains), arr, model.encode(), device, ctypes.byref(h))
        if rc != 0:
            raise ExaError(rc, lib.exa_last_error(None).decode())
        self._h = h
        for k, v in DEFAULT_OPTIONS.items():
            self.set_option(k, v)
        self.sizes = self.state()
        return self

    @property
    def ninst(self):
        return self._lib.exa_batch_size(self._h)

    def instance_stats(self):
        """(newton_iters[ninst], gmres_iters[ninst], final_norm[ninst]) of the last solve."""
        n = self.ninst
        a = np.zeros(n, dtype=np.int32)
        b = np.zeros(n, dtype=np.int32)
        c = np.zeros(n)
        self._check(self._lib.exa_batch_stats(self._h, a.ctypes.data_as(_I), b.ctypes.data_as(_I), _dp(c), n))
        return a, b, c

    def get_instance(self, inst, name):
        """Instance inst's "uh" or "udg" (flattened, column-major as the merged arrays)."""
        n = self._lib.exa_batch_get(self._h, inst, name.encode(), None, 0)
        if n < 0:
            self._check(-n)
        out = np.empty(n, dtype=np.float64)
        r = self._lib.exa_batch_get(self._h, inst, name.encode(), _dp(out), n)
        if r < 0:
            self._check(-r)
        return out

    def step(self):
        """One Newton iteration (steady) or one DIRK time step of every instance still iterating
        (``exa_batch_step``): returns the status per instance (1 iterating, 0 finished, -1 broken
        down)."""
        n = self.ninst
        st = np.zeros(n, dtype=np.int32)
        left = ctypes.c_int(0)
        self._check(self._lib.exa_batch_step(self._h, st.ctypes.data_as(_I), n, ctypes.byref(left)))
        return st

    def refill(self, inst, datain):
        """Put the problem of ``datain`` into the finished slot ``inst`` (``exa_batch_refill``)."""
        self._check(self._lib.exa_batch_refill(self._h, int(inst), datain.encode()))

    SIZE_KEYS = ["nd", "ncu", "nc", "npe", "npf", "nfe", "ne", "nf", "n", "m", "N", "nge", "ngf",
                 "nco", "restart", "preconditioner", "precision", "refinements", "neb",
                 "ilu_nse", "ilu_nfse", "ilu_nnz"]

    def state(self):
        """exa_sizes: the sizes plus the current preconditioner, precision and the number of
        refinement corrections of the last mixed-precision Newton step."""
        s = (ctypes.c_int * len(self.SIZE_KEYS))()
        self._lib.exa_sizes(self._h, s, len(self.SIZE_KEYS))
        return dict(zip(self.SIZE_KEYS, list(s)))

    def _check(self, rc):
        if rc != 0:
            raise ExaError(rc, self._lib.exa_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            self._lib.exa_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # options and stages
    def set_option(self, name, value):
        self._check(self._lib.exa_set_option(self._h, name.encode(), float(value)))

    def solve(self):
        """NewtonSolver: returns (newton_iters, gmres_iters, final_residual_norm)."""
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        self._check(self._lib.exa_solve(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    def reset(self):
        """Restore udg/uh to the sol.bin input state on the device."""
        self._check(self._lib.exa_reset(self._h))

    def get_q(self):
        self._check(self._lib.exa_get_q(self._h))

    def assemble(self):
        """hdgAssembleLinearSystem: H, DinvF, Ru, Rh, b and the preconditioner."""
        self._check(self._lib.exa_assemble(self._h))

    def residual(self):
        """hdgAssembleResidual: returns (||b||, ||Ru||); arrays "res_b", "res_Ru"."""
        a, b = ctypes.c_double(), ctypes.c_double()
        self._check(self._lib.exa_residual(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def matvec(self, v_dev, w_dev):
        self._check(self._lib.exa_matvec(self._h, ctypes.c_void_p(v_dev), ctypes.c_void_p(w_dev)))

    def precond(self, x_dev):
        self._check(self._lib.exa_precond(self._h, ctypes.c_void_p(x_dev)))

    def gmres(self, b_dev, x_dev):
        it = ctypes.c_int()
        self._check(self._lib.exa_gmres(self._h, ctypes.c_void_p(b_dev), ctypes.c_void_p(x_dev),
                                        ctypes.byref(it)))
        return it.value

    def get(self, name):
        n = self._lib.exa_get_array(self._h, name.encode(), None, 0)
        if n < 0:
            self._check(-n)
        out = np.empty(n, dtype=np.float64)
        r = self._lib.exa_get_array(self._h, name.encode(), _dp(out), n)
        if r < 0:
            self._check(-r)
        return out

    def set(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).ravel()
        self._check(self._lib.exa_set_array(self._h, name.encode(), _dp(a), a.size))

    def device_ptr(self, name):
        p, n = _P(), ctypes.c_long()
        self._check(self._lib.exa_device_ptr(self._h, name.encode(), ctypes.byref(p), ctypes.byref(n)))
        return p.value, n.value

    def write_output(self):
        self._check(self._lib.exa_write_output(self._h))

    def timings(self):
        ms = np.zeros(8)
        cnt = np.zeros(8, dtype=np.int64)
        k = self._lib.exa_timings(self._h, _dp(ms), cnt.ctypes.data_as(_L), 8)
        if k < 0:
            self._check(-k)
        return {TIMER_CLASSES[i]: (float(ms[i]), int(cnt[i])) for i in range(min(k, len(TIMER_CLASSES)))}

    def reset_timings(self):
        self._check(self._lib.exa_reset_timings(self._h))

    def sync(self):
        self._check(self._lib.exa_sync(self._h))
