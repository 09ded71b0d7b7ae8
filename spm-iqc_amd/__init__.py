"""spm-iqc_amd: MI355X (gfx950) implementation of the pseudo-spectral Landau
free-energy path of JKGroupCUDA/SPM-IQC.

The compute path is the HIP library ``lib/libspm_iqc.so`` behind the C ABI
declared in ``include/spm_iqc.h``; this module binds it with ctypes.  PyTorch
provides device memory and the stream (every call is enqueued on torch's
current stream).  There is no CPU fallback: if the library is missing, or no
GPU is present, the calls raise.

Import it by path, e.g. ``importlib.import_module("spm-iqc_amd")`` with the
repository root on ``sys.path``.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libspm_iqc.so")

FP64 = 0
FP32 = 1
SUPPORTED_N = (4, 8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256)

# LPsystParameter (Initialization.cpp:7-42)
REFERENCE_PARAMS = dict(L=26.0, kscale=1.0, Tau=-0.0061, Gamma=0.30, dt=0.3)

_lib = None


class SpmError(RuntimeError):
    pass


class _Params(ctypes.Structure):
    _fields_ = [("N", ctypes.c_int), ("precision", ctypes.c_int), ("L", ctypes.c_double),
                ("kscale", ctypes.c_double), ("Tau", ctypes.c_double), ("Gamma", ctypes.c_double),
                ("dt", ctypes.c_double), ("shape", ctypes.c_int), ("step_passes", ctypes.c_int),
                ("step_variant", ctypes.c_int)]


class _Layout(ctypes.Structure):
    _fields_ = [("N", ctypes.c_int), ("nh", ctypes.c_int), ("nhp", ctypes.c_int), ("precision", ctypes.c_int),
                ("R", ctypes.c_longlong), ("C", ctypes.c_longlong), ("real_bytes", ctypes.c_size_t),
                ("spec_bytes", ctypes.c_size_t), ("ctx_bytes", ctypes.c_size_t)]


_c = ctypes.c_void_p
_i = ctypes.c_int
_d = ctypes.c_double
_pd = ctypes.POINTER(ctypes.c_double)
_pi = ctypes.POINTER(ctypes.c_int)

# every symbol declared in include/spm_iqc.h: name -> (restype, argtypes)
SIGNATURES = {
    "spm_create": (_i, [ctypes.POINTER(_Params), ctypes.POINTER(_c)]),
    "spm_destroy": (_i, [_c]),
    "spm_set_stream": (_i, [_c, _c]),
    "spm_get_layout": (_i, [_c, ctypes.POINTER(_Layout)]),
    "spm_last_error": (ctypes.c_char_p, []),
    "spm_get_GK": (_d, [_d]),
    "spm_my_fft": (_i, [_c, _c, _c]),
    "spm_my_ifft": (_i, [_c, _c, _c]),
    "spm_r2c": (_i, [_c, _c, _c]),
    "spm_c2r": (_i, [_c, _c, _c]),
    "spm_energy": (_i, [_c, _c, _pd]),
    "spm_energy_uf": (_i, [_c, _c, _c, _pd]),
    "spm_ngrad": (_i, [_c, _c, _c, _pd]),
    "spm_gradflow": (_i, [_c, _c, _i, _d, _d, _i, _c, _pi]),
    "spm_gf_begin": (_i, [_c, _c, _i, _d, _d]),
    "spm_gf_steps": (_i, [_c, _i]),
    "spm_gf_end": (_i, [_c, _c, _c, _pi]),
    "spm_hv": (_i, [_c, _c, _c, _c]),
    "spm_lobpcg_init": (_i, [_c, _c, _c, _i]),
    "spm_lobpcg": (_i, [_c, _c, _c, _i, _i, _c]),
    "spm_hisd": (_i, [_c, _c, _c, _i, _d, _d, _d, _d, _i, _c, _c]),
    "spm_gosd": (_i, [_c, _c, _c, _i, _c, _i, _d, _d, _d, _d, _i, _c, _c]),
    "spm_rayleighq": (_i, [_c, _c, _c, _pd]),
    "spm_gram": (_i, [_c, _c, _i, _c, _i, _c]),
    "spm_get_tv": (_i, [_c, _c, _c, _c]),
    "spm_natural_force": (_i, [_c, _c, _c]),
    "spm_minmax_force": (_i, [_c, _c, _c]),
    "spm_spring_pairing": (_i, [_c, _c, _c, _i, _d, _pi, _pd, _c]),
    "spm_precond": (_i, [_c, _c]),
    "spm_inpfp": (_i, [_c, _c, _c, _pd]),
    "spm_nrmp": (_i, [_c, _c, _pd]),
    "spm_lin2": (_i, [_c, _c, _d, _c, _d, _c]),
    "spm_ngrad_noalloc": (_i, [ctypes.POINTER(_Params), _c, _c, _pd, _c]),
    "spm_cplx_to_real": (_i, [_c, _c, _c]),
    "spm_real_to_cplx": (_i, [_c, _c, _c]),
    "spm_batch_create": (_i, [ctypes.POINTER(_Params), _i, _i, ctypes.POINTER(_c)]),
    "spm_batch_destroy": (_i, [_c]),
    "spm_batch_stream": (_i, [_c, ctypes.POINTER(_c)]),
    "spm_batch_info": (_i, [_c, _pi, ctypes.POINTER(ctypes.c_size_t)]),
    "spm_batch_load": (_i, [_c, _i, _c, _d, _d, _i, _d, _d]),
    "spm_batch_steps": (_i, [_c, _i]),
    "spm_batch_poll": (_i, [_c, _c, _c]),
    "spm_batch_read": (_i, [_c, _i, _c, _c, _pi]),
    "spm_batch_release": (_i, [_c, _i]),
    "spm_shapes": (_i, [_i, _pi, _c, _i]),
    "spm_tune": (_i, [ctypes.POINTER(_Params), _i, _i, _pi, _c]),
    "spm_tune_step": (_i, [ctypes.POINTER(_Params), _i, _i, _pi, _c]),
}


def lib():
    """The loaded HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SpmError("libspm_iqc.so is not built: run `python spm-iqc_amd/build.py`")
        # torch brings its own HIP runtime; it must be the one in the process
        # before the library binds to libamdhip64, or the two runtimes clash
        import torch  # noqa: F401

        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc: int, what: str):
    if rc != 0:
        raise SpmError("%s failed (%d): %s" % (what, rc, lib().spm_last_error().decode()))


def get_GK(k2: float) -> float:
    """Initialization.cpp:150-166 (host)."""
    return lib().spm_get_GK(float(k2))


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise SpmError("no GPU: the spm-iqc_amd path runs only on the device")
    torch.cuda.init()  # the HIP runtime is torch's: initialise it before the library's first call
    return torch


def _params(N, precision, L, kscale, Tau, Gamma, dt, shape, step_passes=0, step_variant=0):
    return _Params(int(N), int(precision), float(L), float(kscale), float(Tau), float(Gamma), float(dt), int(shape),
                   int(step_passes), int(step_variant))


def shapes(N):
    """FFT pass shapes of N: list of (row, j, i) threads per pencil; index 0 is the default."""
    n = ctypes.c_int()
    _check(lib().spm_shapes(int(N), ctypes.byref(n), None, 0), "spm_shapes")
    out = np.zeros(3 * n.value, dtype=np.int32)
    _check(lib().spm_shapes(int(N), ctypes.byref(n), out.ctypes.data_as(_c), n.value), "spm_shapes")
    return [tuple(int(x) for x in out[3 * k:3 * k + 3]) for k in range(n.value)]


def tune(N, precision=FP64, nmem=1, steps=20, L=26.0, kscale=1.0, Tau=-0.0061, Gamma=0.30, dt=0.3):
    """Time every shape of N (batch of nmem members); returns (best, us_per_step list).
    The choice is recorded in the tuning cache read for shape=-1."""
    _torch()
    p = _params(N, precision, L, kscale, Tau, Gamma, dt, 0)
    best = ctypes.c_int()
    us = np.zeros(len(shapes(N)))
    _check(lib().spm_tune(ctypes.byref(p), int(nmem), int(steps), ctypes.byref(best), us.ctypes.data_as(_c)),
           "spm_tune")
    return best.value, us.tolist()


# Two-pass step variants (spm_params.step_variant): 0 = built-in, -1 = tuned,
# else STEP_EXPLICIT | flags
STEP_EXPLICIT, STEP_GATHER_U, STEP_U_CACHED, STEP_U_STREAM, STEP_VARIANTS = 8, 1, 2, 4, 6


def tune_step(N, precision=FP64, nmem=1, steps=20, L=26.0, kscale=1.0, Tau=-0.0061, Gamma=0.30, dt=0.3):
    """Time every two-pass step variant (flag sets 0 .. STEP_VARIANTS-1) of N;
    returns (best flags, us_per_step list).  Recorded for step_variant=-1."""
    _torch()
    p = _params(N, precision, L, kscale, Tau, Gamma, dt, 0, 2)
    best = ctypes.c_int()
    us = np.zeros(STEP_VARIANTS)
    _check(lib().spm_tune_step(ctypes.byref(p), int(nmem), int(steps), ctypes.byref(best), us.ctypes.data_as(_c)),
           "spm_tune_step")
    return best.value, us.tolist()


class Grid:
    """One device context: an N^3 grid, the model parameters and a precision.

    Fields are torch tensors on the GPU: real fields of shape (N, N, N) in the
    context dtype (float64 / float32), half spectra of shape (N, N, nhp)
    complex.  Mirrors the reference's process-global state (Data.h:4-28).
    """

    def __init__(self, N, precision=FP64, L=26.0, kscale=1.0, Tau=-0.0061, Gamma=0.30, dt=0.3, shape=0,
                 step_passes=0, step_variant=0):
        torch = _torch()
        self.torch = torch
        p = _params(N, precision, L, kscale, Tau, Gamma, dt, shape, step_passes, step_variant)
        h = _c()
        _check(lib().spm_create(ctypes.byref(p), ctypes.byref(h)), "spm_create")
        self.h = h
        lay = _Layout()
        _check(lib().spm_get_layout(self.h, ctypes.byref(lay)), "spm_get_layout")
        self.N, self.nh, self.nhp = lay.N, lay.nh, lay.nhp
        self.R, self.C = lay.R, lay.C
        self.ctx_bytes = lay.ctx_bytes
        self.precision = int(precision)
        self.dtype = torch.float64 if precision == FP64 else torch.float32
        self.cdtype = torch.complex128 if precision == FP64 else torch.complex64
        self.params = dict(L=L, kscale=kscale, Tau=Tau, Gamma=Gamma, dt=dt)

    # -- lifetime -----------------------------------------------------------
    def close(self):
        if getattr(self, "h", None):
            lib().spm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- helpers ------------------------------------------------------------
    def _stream(self):
        _check(lib().spm_set_stream(self.h, _c(self.torch.cuda.current_stream().cuda_stream)), "spm_set_stream")

    def _ptr(self, t, dtype, shape=None):
        if not t.is_cuda:
            raise SpmError("field must be a GPU tensor")
        if t.dtype != dtype:
            raise SpmError("field dtype %s, expected %s" % (t.dtype, dtype))
        if not t.is_contiguous():
            raise SpmError("field must be contiguous")
        if shape is not None and t.numel() != int(np.prod(shape)):
            raise SpmError("field has %d elements, expected %s" % (t.numel(), shape))
        return _c(t.data_ptr())

    def _real(self, t):
        return self._ptr(t, self.dtype, (self.N,) * 3)

    def _spec(self, t):
        return self._ptr(t, self.cdtype, (self.N, self.N, self.nhp))

    def _full(self, t):
        return self._ptr(t, self.cdtype, (self.N,) * 3)

    def real_field(self, values=None):
        t = self.torch.zeros((self.N,) * 3, dtype=self.dtype, device="cuda")
        if values is not None:
            t.copy_(self.torch.as_tensor(np.asarray(values)).reshape(t.shape))
        return t

    def spectrum(self):
        return self.torch.zeros((self.N, self.N, self.nhp), dtype=self.cdtype, device="cuda")

    # -- My_FFT -------------------------------------------------------------
    def my_fft(self, x_R, x_F=None):
        """My_FFT.cpp:4-14: c2c forward / N^3 on a full complex (N,N,N) field."""
        self._stream()
        x_F = self.torch.empty_like(x_R) if x_F is None else x_F
        _check(lib().spm_my_fft(self.h, self._full(x_R), self._full(x_F)), "spm_my_fft")
        return x_F

    def my_ifft(self, x_F, x_R=None):
        """My_FFT.cpp:16-21: c2c inverse, unnormalised."""
        self._stream()
        x_R = self.torch.empty_like(x_F) if x_R is None else x_R
        _check(lib().spm_my_ifft(self.h, self._full(x_F), self._full(x_R)), "spm_my_ifft")
        return x_R

    def r2c(self, x, S=None):
        self._stream()
        S = self.spectrum() if S is None else S
        _check(lib().spm_r2c(self.h, self._real(x), self._spec(S)), "spm_r2c")
        return S

    def c2r(self, S, x=None):
        self._stream()
        x = self.real_field() if x is None else x
        _check(lib().spm_c2r(self.h, self._spec(S), self._real(x)), "spm_c2r")
        return x

    # -- LPsrc --------------------------------------------------------------
    def energy(self, phi) -> float:
        """get_Ene (LPsrc.cpp:8-38) of the real field phi."""
        self._stream()
        E = ctypes.c_double()
        _check(lib().spm_energy(self.h, self._real(phi), ctypes.byref(E)), "spm_energy")
        return E.value

    def energy_uf(self, phi, u_F) -> float:
        """get_Ene (LPsrc.cpp:8-38) with the non-local term from the full complex
        spectrum u_F (N, N, N) rather than from the transform of phi."""
        self._stream()
        E = ctypes.c_double()
        _check(lib().spm_energy_uf(self.h, self._real(phi), self._full(u_F), ctypes.byref(E)), "spm_energy_uf")
        return E.value

    def ngrad(self, phi, grad=None, want_norm=True):
        """ngrad_cam (LPsrc.cpp:119-173); returns (grad, nrmp(grad))."""
        self._stream()
        grad = self.real_field() if grad is None else grad
        nrm = ctypes.c_double()
        _check(lib().spm_ngrad(self.h, self._real(phi), self._real(grad),
                               ctypes.byref(nrm) if want_norm else None), "spm_ngrad")
        return grad, (nrm.value if want_norm else None)

    def gradflow(self, phi, max_iter, res_tol=1e-16, gn_tol=0.0, check_every=64):
        """get_min_u_F (LPsrc.cpp:40-117) on phi in place.

        Returns (iters, hist) with hist[n] = (E, ||grad F||, res) of iteration n.
        """
        self._stream()
        hist = np.zeros((max(int(max_iter), 1), 3))
        it = ctypes.c_int()
        _check(lib().spm_gradflow(self.h, self._real(phi), int(max_iter), float(res_tol), float(gn_tol),
                                  int(check_every), hist.ctypes.data_as(_c), ctypes.byref(it)), "spm_gradflow")
        return it.value, hist[: it.value]

    def gf_begin(self, phi, max_iter, res_tol=1e-16, gn_tol=0.0):
        self._stream()
        _check(lib().spm_gf_begin(self.h, self._real(phi), int(max_iter), float(res_tol), float(gn_tol)),
               "spm_gf_begin")

    def gf_steps(self, n):
        """Enqueue n gradient-flow iterations (asynchronous)."""
        _check(lib().spm_gf_steps(self.h, int(n)), "spm_gf_steps")

    def gf_end(self, phi=None, max_iter=0):
        """Final field into phi (if given); returns (iters, hist).  max_iter must be
        the gf_begin limit for the history to be returned (0: no history)."""
        self._stream()
        hist = np.zeros((max(int(max_iter), 1), 3))
        it = ctypes.c_int()
        _check(lib().spm_gf_end(self.h, self._real(phi) if phi is not None else None,
                                hist.ctypes.data_as(_c) if max_iter > 0 else None, ctypes.byref(it)), "spm_gf_end")
        return it.value, hist[: min(it.value, len(hist))] if max_iter > 0 else hist[:0]

    # -- LOBPCG operators ----------------------------------------------------
    def hv(self, phi, v, hv=None):
        """hv_cam (LOBPCG.cpp:113-166)."""
        self._stream()
        hv = self.real_field() if hv is None else hv
        _check(lib().spm_hv(self.h, self._real(phi), self._real(v), self._real(hv)), "spm_hv")
        return hv

    def precond(self, w):
        """precond_camnew (LOBPCG.cpp:168-199), in place."""
        self._stream()
        _check(lib().spm_precond(self.h, self._real(w)), "spm_precond")
        return w

    def _block(self, t, k=None):
        k = t.shape[0] if k is None else k
        return self._ptr(t, self.dtype, (k, self.N, self.N, self.N))

    def lobpcg_init(self, start, V=None):
        """hiosd_lobpcg_initialization (LOBPCG.cpp:10-109) from a (k, N, N, N) start block."""
        self._stream()
        V = self.torch.empty_like(start) if V is None else V
        _check(lib().spm_lobpcg_init(self.h, self._block(start), self._block(V), int(start.shape[0])),
               "spm_lobpcg_init")
        return V

    def lobpcg(self, phi, V, maxiter=21):
        """hiosd_lobpcg_inip (LOBPCG.cpp:207-487) on V in place; returns the last lowest-k Ritz values."""
        self._stream()
        k = int(V.shape[0])
        ritz = np.zeros(k)
        _check(lib().spm_lobpcg(self.h, self._real(phi), self._block(V), k, int(maxiter), ritz.ctypes.data_as(_c)),
               "spm_lobpcg")
        return ritz

    def hisd(self, x, V, dt=0.8, betat=8.0, betau=0.02, epsf=1e-7, maxiter=10000):
        """hiosd_lobpcg_sieh (HISD.cpp:86-485) on x and V (k, N, N, N) in place; returns
        (steps, last residual)."""
        self._stream()
        k = int(V.shape[0])
        it = np.zeros(1, dtype=np.int32)
        res = np.zeros(1)
        _check(lib().spm_hisd(self.h, self._real(x), self._block(V), k, float(dt), float(betat), float(betau),
                              float(epsf), int(maxiter), it.ctypes.data_as(_c), res.ctypes.data_as(_c)), "spm_hisd")
        return int(it[0]), float(res[0])

    def gosd(self, x, V, V0, dt=0.8, betat=8.0, betau=1.3, epsf=1e-9, maxiter=40000):
        """hiosd_lobpcg_sieh_V0 (GOSD.cpp:123-531) on x and V (k, N, N, N) in place,
        orthogonal to the fixed fields V0 (k0, N, N, N); returns (steps, last residual)."""
        self._stream()
        k, k0 = int(V.shape[0]), int(V0.shape[0])
        it = np.zeros(1, dtype=np.int32)
        res = np.zeros(1)
        _check(lib().spm_gosd(self.h, self._real(x), self._block(V), k, self._block(V0) if k0 else None, k0,
                              float(dt), float(betat), float(betau), float(epsf), int(maxiter),
                              it.ctypes.data_as(_c), res.ctypes.data_as(_c)), "spm_gosd")
        return int(it[0]), float(res[0])

    def gram(self, X, Y):
        """inpfp of every pair of fields of the blocks X (ma, N, N, N), Y (mb, N, N, N)."""
        self._stream()
        ma, mb = int(X.shape[0]), int(Y.shape[0])
        out = np.zeros((ma, mb))
        _check(lib().spm_gram(self.h, self._block(X), ma, self._block(Y), mb, out.ctypes.data_as(_c)), "spm_gram")
        return out

    def rayleighq(self, phi, v) -> float:
        self._stream()
        out = ctypes.c_double()
        _check(lib().spm_rayleighq(self.h, self._real(phi), self._real(v), ctypes.byref(out)), "spm_rayleighq")
        return out.value

    # -- SPM ------------------------------------------------------------------
    def get_tv(self, phi0, phi1, v=None):
        self._stream()
        v = self.real_field() if v is None else v
        _check(lib().spm_get_tv(self.h, self._real(phi0), self._real(phi1), self._real(v)), "spm_get_tv")
        return v

    def natural_force(self, x, v):
        self._stream()
        _check(lib().spm_natural_force(self.h, self._real(x), self._real(v)), "spm_natural_force")
        return x

    def minmax_force(self, x, v):
        self._stream()
        _check(lib().spm_minmax_force(self.h, self._real(x), self._real(v)), "spm_minmax_force")
        return x

    def spring_pairing(self, phi0, phi1, maxiter, tol):
        """Spring_Pairing (SPM.cpp:170-263) in place; returns (iters, res, res_hist)."""
        self._stream()
        it = ctypes.c_int()
        res = ctypes.c_double()
        hist = np.zeros(max(int(maxiter), 1))
        _check(lib().spm_spring_pairing(self.h, self._real(phi0), self._real(phi1), int(maxiter), float(tol),
                                        ctypes.byref(it), ctypes.byref(res), hist.ctypes.data_as(_c)),
               "spm_spring_pairing")
        return it.value, res.value, hist[: it.value]

    # -- BasicOperators -------------------------------------------------------
    def inpfp(self, p, q) -> float:
        self._stream()
        out = ctypes.c_double()
        _check(lib().spm_inpfp(self.h, self._real(p), self._real(q), ctypes.byref(out)), "spm_inpfp")
        return out.value

    def nrmp(self, p) -> float:
        self._stream()
        out = ctypes.c_double()
        _check(lib().spm_nrmp(self.h, self._real(p), ctypes.byref(out)), "spm_nrmp")
        return out.value

    def lin2(self, r, a1, F1, a2, F2):
        self._stream()
        _check(lib().spm_lin2(self.h, self._real(r), float(a1), self._real(F1), float(a2), self._real(F2)),
               "spm_lin2")
        return r


class Batch:
    """A batch of independent gradient flows (get_min_u_F, LPsrc.cpp:40-117)
    of one grid size and precision, each member with its own (Tau, Gamma) and
    limits; members are loaded and read while the others run.  Fields are
    torch tensors on the GPU in the batch precision."""

    def __init__(self, N, precision=FP64, capacity=8, hist_cap=4096, L=26.0, kscale=1.0, Tau=-0.0061,
                 Gamma=0.30, dt=0.3, shape=0, step_passes=0, step_variant=0):
        torch = _torch()
        self.torch = torch
        p = _params(N, precision, L, kscale, Tau, Gamma, dt, shape, step_passes, step_variant)
        h = _c()
        _check(lib().spm_batch_create(ctypes.byref(p), int(capacity), int(hist_cap), ctypes.byref(h)),
               "spm_batch_create")
        self.h = h
        self.N, self.capacity, self.hist_cap = int(N), int(capacity), int(hist_cap)
        self.precision = int(precision)
        self.dtype = torch.float64 if precision == FP64 else torch.float32
        cap, nb = ctypes.c_int(), ctypes.c_size_t()
        _check(lib().spm_batch_info(self.h, ctypes.byref(cap), ctypes.byref(nb)), "spm_batch_info")
        self.dev_bytes = nb.value
        s = _c()
        _check(lib().spm_batch_stream(self.h, ctypes.byref(s)), "spm_batch_stream")
        self.stream = torch.cuda.ExternalStream(s.value) if s.value else None
        self._keep = {}  # inputs read asynchronously: held until the next poll / read

    def close(self):
        if getattr(self, "h", None):
            lib().spm_batch_destroy(self.h)
            self.h = None
            self._keep = {}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _field(self, t):
        if not t.is_cuda or t.dtype != self.dtype or not t.is_contiguous() or t.numel() != self.N ** 3:
            raise SpmError("field must be a contiguous (N, N, N) %s GPU tensor" % self.dtype)
        return _c(t.data_ptr())

    def load(self, slot, phi, Tau, Gamma, max_iter, res_tol=1e-16, gn_tol=0.0):
        """Start member `slot` from phi (copied into the member's state)."""
        self.torch.cuda.current_stream().synchronize()  # phi is read on the batch stream
        _check(lib().spm_batch_load(self.h, int(slot), self._field(phi), float(Tau), float(Gamma), int(max_iter),
                                    float(res_tol), float(gn_tol)), "spm_batch_load")
        self._keep[slot] = phi

    def steps(self, n):
        """Enqueue n iterations of every running member (asynchronous)."""
        _check(lib().spm_batch_steps(self.h, int(n)), "spm_batch_steps")

    def poll(self):
        """(running count, iters array, stopped array); synchronises."""
        it = np.zeros(self.capacity, dtype=np.int32)
        st = np.zeros(self.capacity, dtype=np.int32)
        rc = lib().spm_batch_poll(self.h, it.ctypes.data_as(_c), st.ctypes.data_as(_c))
        if rc < 0:
            _check(rc, "spm_batch_poll")
        self._keep = {}
        return rc, it, st

    def read(self, slot, phi=None, want_hist=True):
        """Field (into phi, or a new tensor) and history of a stopped member."""
        if phi is None:
            phi = self.torch.empty((self.N,) * 3, dtype=self.dtype, device="cuda")
        hist = np.zeros((self.hist_cap, 3))
        it = ctypes.c_int()
        self.torch.cuda.current_stream().synchronize()  # phi is written on the batch stream
        _check(lib().spm_batch_read(self.h, int(slot), self._field(phi),
                                    hist.ctypes.data_as(_c) if want_hist else None, ctypes.byref(it)),
               "spm_batch_read")
        self._keep.pop(slot, None)
        return phi, hist[: it.value]

    def release(self, slot):
        _check(lib().spm_batch_release(self.h, int(slot)), "spm_batch_release")
