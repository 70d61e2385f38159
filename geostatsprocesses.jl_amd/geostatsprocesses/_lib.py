"""ctypes binding of libgsp_hip.so (include/gsp.h).

The shared library is the product: every computation of SEQSIM and LUSIM runs
in it on the GPU. There is no CPU fallback; loading fails loudly when the
library has not been built.
"""
from __future__ import annotations

import ctypes
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.normpath(os.path.join(_HERE, "..", "lib", "libgsp_hip.so"))

GSP_OK, GSP_ERR_ARGUMENT, GSP_ERR_ASSERTION, GSP_ERR_POSDEF = 0, 1, 2, 3
GSP_ERR_DEVICE, GSP_ERR_MEMORY, GSP_ERR_UNSUPPORTED, GSP_ERR_INTERNAL = 4, 5, 6, 7
SPHERICAL, EXPONENTIAL, GAUSSIAN = 0, 1, 2
VARIOGRAM, COVARIANCE = 0, 1
KNN_THEN_MASK, MASKED_KNN, DATA_ONLY = 0, 1, 2
FP64, FP32_REFINE = 0, 1
LU_FP64, LU_MIXED = 0, 1
FLAG_FEW, FLAG_FAIL, FLAG_FP64_FALLBACK, FLAG_LEARNED = 1, 2, 4, 8


class ArgumentError(ValueError):
    """Julia's ArgumentError (lusim.jl:46, field.jl:99)."""


class PosDefException(np.linalg.LinAlgError):
    """LinearAlgebra.PosDefException raised by `cholesky` (lusim.jl:92,98,103)."""


class GspError(RuntimeError):
    """Device, memory or internal failure of the backend."""


class UnsupportedError(GspError):
    """A configuration this backend does not implement."""


class GspModel(ctypes.Structure):
    _fields_ = [("family", ctypes.c_int32), ("kind", ctypes.c_int32),
                ("sill", ctypes.c_double), ("nugget", ctypes.c_double),
                ("ranges", ctypes.c_double * 3), ("rotation", ctypes.c_double * 9),
                ("has_rotation", ctypes.c_int32), ("pad_", ctypes.c_int32)]


class GspGrid(ctypes.Structure):
    _fields_ = [("ndim", ctypes.c_int32), ("pad_", ctypes.c_int32),
                ("dims", ctypes.c_int64 * 3), ("origin", ctypes.c_double * 3),
                ("spacing", ctypes.c_double * 3)]


class GspSeqsimProblem(ctypes.Structure):
    _fields_ = [("grid", GspGrid), ("model", GspModel), ("mean", ctypes.c_double),
                ("n_data", ctypes.c_int64), ("data_coords", ctypes.c_void_p),
                ("n_path", ctypes.c_int64), ("path", ctypes.c_void_p),
                ("minneighbors", ctypes.c_int32), ("maxneighbors", ctypes.c_int32),
                ("radius", ctypes.c_double), ("search_mode", ctypes.c_int32),
                ("precision", ctypes.c_int32), ("device_pointers", ctypes.c_int32),
                ("pad_", ctypes.c_int32)]


class GspLusimProblem(ctypes.Structure):
    _fields_ = [("grid", GspGrid), ("model", GspModel), ("mean", ctypes.c_double),
                ("n_data", ctypes.c_int64), ("dinds", ctypes.c_void_p),
                ("z1", ctypes.c_void_p), ("precision", ctypes.c_int32),
                ("device_pointers", ctypes.c_int32)]


class GspFftsimProblem(ctypes.Structure):
    _fields_ = [("grid", GspGrid), ("model", GspModel), ("mean", ctypes.c_double),
                ("n_data", ctypes.c_int64), ("data_coords", ctypes.c_void_p),
                ("data_vals", ctypes.c_void_p), ("n_cells", ctypes.c_int64),
                ("cells", ctypes.c_void_p), ("minneighbors", ctypes.c_int32),
                ("maxneighbors", ctypes.c_int32), ("radius", ctypes.c_double),
                ("device_pointers", ctypes.c_int32), ("pad_", ctypes.c_int32)]


class GspSeqsimInfo(ctypes.Structure):
    _fields_ = [("n_nodes", ctypes.c_int64), ("n_data", ctypes.c_int64),
                ("minneighbors", ctypes.c_int32), ("maxneighbors", ctypes.c_int32),
                ("depth", ctypes.c_int32), ("pad_", ctypes.c_int32),
                ("alpha", ctypes.c_double), ("n_prior", ctypes.c_int64),
                ("n_fallback", ctypes.c_int64), ("n_learned", ctypes.c_int64)]


(K_PLAN, K_LEVELS, K_SWEEP, K_TRANSPOSE, K_LU_FACTOR, K_LU_SAMPLE, K_SEARCH, K_LEARNED, K_FFT,
 K_SPECTRAL, K_STATS, K_LU_TRAIL, K_COUNT) = 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12
STAT_MEAN, STAT_VAR, STAT_CDF, STAT_CCDF, STAT_QUANTILE = 0, 1, 2, 3, 4


class GspTrainOptions(ctypes.Structure):
    _fields_ = [("max_samples", ctypes.c_int64), ("epochs", ctypes.c_int32),
                ("batch", ctypes.c_int32), ("lr", ctypes.c_double), ("var_weight", ctypes.c_double),
                ("seed", ctypes.c_uint64), ("n_slots", ctypes.c_int32),
                ("data_slots", ctypes.c_int32)]


class GspTrainReport(ctypes.Structure):
    _fields_ = [("n_samples", ctypes.c_int64), ("n_candidates", ctypes.c_int64),
                ("steps", ctypes.c_int64), ("rms_weight_first_epoch", ctypes.c_double),
                ("rms_weight_last_epoch", ctypes.c_double),
                ("rms_logvar_last_epoch", ctypes.c_double), ("ms", ctypes.c_double)]


class GspSurrogateInfo(ctypes.Structure):
    _fields_ = [("family", ctypes.c_int32), ("ndim", ctypes.c_int32), ("n_slots", ctypes.c_int32),
                ("n_train_requests", ctypes.c_int32), ("sill", ctypes.c_double),
                ("nugget_ratio", ctypes.c_double * 2), ("scale", (ctypes.c_double * 2) * 3),
                ("n_samples", ctypes.c_int64), ("n_data_slots", ctypes.c_int32),
                ("pad_", ctypes.c_int32)]


class GspTimings(ctypes.Structure):
    _fields_ = [("ms", ctypes.c_double * K_COUNT), ("launches", ctypes.c_int64 * K_COUNT)]


# every symbol include/gsp.h declares, with (restype, argtypes)
_VP, _I32, _I64, _D = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
SIGNATURES = {
    "gsp_abi_version": (_I32, []),
    "gsp_ctx_create": (_I32, [_I32, ctypes.POINTER(_VP)]),
    "gsp_ctx_destroy": (_I32, [_VP]),
    "gsp_ctx_stream": (_I32, [_VP, ctypes.POINTER(_VP)]),
    "gsp_ctx_synchronize": (_I32, [_VP]),
    "gsp_last_error": (ctypes.c_char_p, []),
    "gsp_ctx_enable_timing": (_I32, [_VP, _I32]),
    "gsp_ctx_timings": (_I32, [_VP, ctypes.POINTER(GspTimings), _I32]),
    "gsp_device_alloc": (_I32, [_VP, _I64, ctypes.POINTER(_VP)]),
    "gsp_device_free": (_I32, [_VP, _VP]),
    "gsp_memcpy_to_device": (_I32, [_VP, _VP, _VP, _I64]),
    "gsp_memcpy_to_host": (_I32, [_VP, _VP, _VP, _I64]),
    "gsp_seqsim_plan_create": (_I32, [_VP, ctypes.POINTER(GspSeqsimProblem), ctypes.POINTER(_VP)]),
    "gsp_seqsim_plan_info": (_I32, [_VP, ctypes.POINTER(GspSeqsimInfo)]),
    "gsp_seqsim_plan_export": (_I32, [_VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP]),
    "gsp_seqsim_sweep": (_I32, [_VP, _VP, _I32, _VP, _VP, _VP, _I32]),
    "gsp_seqsim_plan_free": (_I32, [_VP]),
    "gsp_seqsim_sweep_noalloc": (_I32, [_VP, _VP, _I32, _VP, _VP, _VP]),
    "gsp_seqsim_sweep_indicator": (_I32, [_VP, _VP, _I32, _I32, _VP, _VP, _VP, _VP, _I32]),
    "gsp_seqsim_sweep_multivariate": (_I32, [_VP, _VP, _I32, _I32, _VP, _VP, _VP, _VP, _VP,
                                             _I32]),
    "gsp_seqsim_plan_create_batch": (_I32, [_VP, _I32, ctypes.POINTER(GspSeqsimProblem),
                                            ctypes.POINTER(_VP)]),
    "gsp_seqsim_plan_layout": (_I32, [_VP, ctypes.POINTER(_I32), _VP, _VP]),
    "gsp_surrogate_train": (_I32, [_VP, _I32, _VP, ctypes.POINTER(GspTrainOptions),
                                   ctypes.POINTER(_VP), ctypes.POINTER(GspTrainReport)]),
    "gsp_surrogate_free": (_I32, [_VP]),
    "gsp_surrogate_export": (_I32, [_VP, _VP, _VP, ctypes.POINTER(_I64), _VP]),
    "gsp_seqsim_plan_create_learned": (_I32, [_VP, _I32, ctypes.POINTER(GspSeqsimProblem), _VP,
                                              ctypes.c_double, ctypes.POINTER(_VP)]),
    "gsp_seqsim_plan_create_learned_multi": (_I32, [_VP, _I32, ctypes.POINTER(GspSeqsimProblem),
                                                    _I32, _VP, ctypes.c_double,
                                                    ctypes.POINTER(_VP)]),
    "gsp_surrogate_get_info": (_I32, [_VP, ctypes.POINTER(GspSurrogateInfo)]),
    "gsp_lusim_factor": (_I32, [_VP, ctypes.POINTER(GspLusimProblem), ctypes.POINTER(_VP)]),
    "gsp_lusim_sample": (_I32, [_VP, _VP, _I32, _VP, _VP, _I32]),
    "gsp_lusim_dims": (_I32, [_VP, ctypes.POINTER(_I64), ctypes.POINTER(_I64)]),
    "gsp_lusim_export": (_I32, [_VP, _VP, _VP, _VP]),
    "gsp_lusim_plan_free": (_I32, [_VP]),
    "gsp_fftsim_plan_create": (_I32, [_VP, ctypes.POINTER(GspFftsimProblem), ctypes.POINTER(_VP)]),
    "gsp_fftsim_sample": (_I32, [_VP, _VP, _I32, _VP, _VP, _I32]),
    "gsp_fftsim_export": (_I32, [_VP, _VP, _VP, _VP]),
    "gsp_fftsim_plan_free": (_I32, [_VP]),
    "gsp_ensemble_stats": (_I32, [_VP, _I32, _I64, _VP, _I32, _VP, _VP, _VP, _I32]),
}

# test hooks (include/gsp_testing.h)
TESTING_SIGNATURES = {
    "gsp_selftest_mfma": (_I32, [_VP, _I32, _VP]),
    "gsp_selftest_copy": (_I32, [_VP, _VP, _VP, _I64]),
}

_lib = None
_lock = threading.Lock()


def load():
    """Load libgsp_hip.so (raises if it is missing: there is no fallback)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise ImportError(
                    f"libgsp_hip.so not found at {LIB_PATH}; build it with "
                    "`make -C geostatsprocesses.jl_amd/csrc` (hipcc, gfx950)")
            L = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in {**SIGNATURES, **TESTING_SIGNATURES}.items():
                fn = getattr(L, name)
                fn.restype = res
                fn.argtypes = args
            _lib = L
    return _lib


def check(rc: int):
    if rc == GSP_OK:
        return
    msg = load().gsp_last_error().decode(errors="replace")
    if rc == GSP_ERR_ARGUMENT:
        raise ArgumentError(msg)
    if rc == GSP_ERR_ASSERTION:
        raise AssertionError(msg)
    if rc == GSP_ERR_POSDEF:
        raise PosDefException(msg)
    if rc == GSP_ERR_UNSUPPORTED:
        raise UnsupportedError(msg)
    raise GspError(f"gsp status {rc}: {msg}")


def ptr(a) -> int:
    """Address of a numpy array, a DeviceBuffer, or any object with data_ptr()."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    if isinstance(a, DeviceBuffer):
        return a.ptr
    return int(a.data_ptr())


class DeviceBuffer:
    """A device allocation made by the library (gsp_device_alloc)."""

    def __init__(self, ctx, shape, dtype=np.float64):
        self.ctx = ctx
        self.shape = tuple(np.atleast_1d(shape).tolist()) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        h = ctypes.c_void_p()
        check(load().gsp_device_alloc(ctx.handle, self.nbytes, ctypes.byref(h)))
        self._p = h

    @classmethod
    def from_array(cls, ctx, arr):
        arr = np.ascontiguousarray(arr)
        buf = cls(ctx, arr.shape, arr.dtype)
        buf.upload(arr)
        return buf

    @property
    def ptr(self) -> int:
        return self._p.value

    def upload(self, arr):
        arr = np.ascontiguousarray(arr, dtype=self.dtype)
        assert arr.nbytes == self.nbytes
        check(load().gsp_memcpy_to_device(self.ctx.handle, self._p, arr.ctypes.data, self.nbytes))

    def download(self):
        out = np.empty(self.shape, dtype=self.dtype)
        check(load().gsp_memcpy_to_host(self.ctx.handle, out.ctypes.data, self._p, self.nbytes))
        return out

    def free(self):
        if self._p and self._p.value:
            load().gsp_device_free(self.ctx.handle, self._p)
            self._p = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def selftest_mfma(ctx, fp32):
    """fp32: False/True for the FP64/FP32 16x16x4 probe, "bf16" for the
    16x16x32 BF16 probe (three setups)."""
    mode = 2 if fp32 == "bf16" else int(bool(fp32))
    out = np.zeros((3 if mode == 2 else 2, 64, 4))
    check(load().gsp_selftest_mfma(ctx.handle, mode, out.ctypes.data))
    return out


class Context:
    """One HIP stream + workspace on one device (gsp_ctx)."""

    def __init__(self, device: int = 0):
        self._h = ctypes.c_void_p()
        check(load().gsp_ctx_create(int(device), ctypes.byref(self._h)))
        self.device = device

    @property
    def handle(self):
        return self._h

    @property
    def stream(self) -> int:
        s = ctypes.c_void_p()
        check(load().gsp_ctx_stream(self._h, ctypes.byref(s)))
        return s.value or 0

    def synchronize(self):
        check(load().gsp_ctx_synchronize(self._h))

    def enable_timing(self, on=True):
        check(load().gsp_ctx_enable_timing(self._h, int(bool(on))))

    def timings(self, reset=False):
        """{kernel slot: (milliseconds, launches)} accumulated with HIP events."""
        t = GspTimings()
        check(load().gsp_ctx_timings(self._h, ctypes.byref(t), int(bool(reset))))
        return {k: (t.ms[k], t.launches[k]) for k in range(K_COUNT)}

    def close(self):
        if self._h and self._h.value:
            load().gsp_ctx_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default = {}


def default_context(device: int | None = None) -> Context:
    if device is None:
        device = int(os.environ.get("GSP_DEVICE", "0"))
    with _lock:
        ctx = _default.get(device)
    if ctx is None:
        ctx = Context(device)
        with _lock:
            _default[device] = ctx
    return ctx


class SeqsimPlan:
    """gsp_seqsim_plan: realization-invariant neighbours, weights and sigmas.

    `problem` is one GspSeqsimProblem, or a sequence of them for a batch of
    independent requests sharing one plan (gsp_seqsim_plan_create_batch)."""

    def __init__(self, ctx: Context, problem, keep=(), surrogate=None, tolerance=0.05):
        self.ctx = ctx
        self._keep = keep
        self._h = ctypes.c_void_p()
        if surrogate is not None:
            # one surrogate, or a sequence of them (each request takes the
            # first whose domain holds it)
            probs = [problem] if isinstance(problem, GspSeqsimProblem) else list(problem)
            arr = (GspSeqsimProblem * len(probs))(*probs)
            surs = list(surrogate) if isinstance(surrogate, (list, tuple)) else [surrogate]
            handles = (ctypes.c_void_p * len(surs))(*[x.handle.value for x in surs])
            self._keep = (keep, arr, surs, handles)
            check(load().gsp_seqsim_plan_create_learned_multi(
                ctx.handle, len(probs), arr, len(surs), handles, float(tolerance),
                ctypes.byref(self._h)))
        elif isinstance(problem, GspSeqsimProblem):
            check(load().gsp_seqsim_plan_create(ctx.handle, ctypes.byref(problem),
                                                ctypes.byref(self._h)))
        else:
            arr = (GspSeqsimProblem * len(problem))(*problem)
            self._keep = (keep, arr)
            check(load().gsp_seqsim_plan_create_batch(ctx.handle, len(problem), arr,
                                                      ctypes.byref(self._h)))
        info = GspSeqsimInfo()
        check(load().gsp_seqsim_plan_info(self._h, ctypes.byref(info)))
        self.info = info
        n = ctypes.c_int32()
        check(load().gsp_seqsim_plan_layout(self._h, ctypes.byref(n), None, None))
        self.node_offset = np.zeros(n.value + 1, np.int64)
        self.data_offset = np.zeros(n.value + 1, np.int64)
        check(load().gsp_seqsim_plan_layout(self._h, ctypes.byref(n), ptr(self.node_offset),
                                            ptr(self.data_offset)))

    @property
    def n_requests(self):
        return len(self.node_offset) - 1

    @property
    def n_nodes(self):
        return self.info.n_nodes

    @property
    def kmax(self):
        return self.info.maxneighbors

    def export(self, fields=("nfound", "nbr", "lam", "sigma", "flags", "level")):
        N, K = self.info.n_nodes, self.info.maxneighbors
        make = dict(nfound=lambda: np.zeros(N, np.int32), nbr=lambda: np.zeros((N, K), np.int32),
                    lam=lambda: np.zeros((N, K)), sigma=lambda: np.zeros(N),
                    flags=lambda: np.zeros(N, np.uint8), level=lambda: np.zeros(N, np.int32))
        out = {k: make[k]() for k in fields}
        check(load().gsp_seqsim_plan_export(self.ctx.handle, self._h, *(ptr(out.get(k)) for k in
                                            ("nfound", "nbr", "lam", "sigma", "flags", "level"))))
        return out

    def sweep_indicator(self, u, prob, data_cat=None, out=None, device_pointers=False):
        """u: (R, N) uniforms in path order; prob: a priori probabilities;
        data_cat: 0-based data categories. Returns (R, N) int32 categories."""
        prob = np.ascontiguousarray(prob, dtype=np.float64)
        if device_pointers:
            R = int(u.shape[0])
            check(load().gsp_seqsim_sweep_indicator(self.ctx.handle, self._h, R, len(prob),
                                                    ptr(prob), ptr(u), ptr(data_cat), ptr(out), 1))
            return out
        u = np.ascontiguousarray(u, dtype=np.float64)
        if u.ndim == 1:
            u = u[None, :]
        R = u.shape[0]
        if u.shape[1] != self.info.n_nodes:
            raise ArgumentError("u must hold one uniform per path node")
        dc = None if data_cat is None else np.ascontiguousarray(data_cat, dtype=np.int32)
        res = np.empty((R, self.info.n_nodes), np.int32) if out is None else out
        check(load().gsp_seqsim_sweep_indicator(self.ctx.handle, self._h, R, len(prob), ptr(prob),
                                                ptr(u), ptr(dc), ptr(res), 0))
        return res

    def sweep_multivariate(self, z, A, mean, data_vals=None):
        """z: (R, N, nvars) standard normals in draw order; A: (nvars, nvars)
        coefficient matrix of A * c(h); mean: nvars means; data_vals:
        (nvars, n_data). Returns (nvars, R, N)."""
        z = np.ascontiguousarray(z, dtype=np.float64)
        R, N, nv = z.shape
        if N != self.info.n_nodes:
            raise ArgumentError("z must hold nvars standard normals per path node")
        A = np.ascontiguousarray(A, dtype=np.float64).reshape(nv, nv)
        mean = np.ascontiguousarray(np.broadcast_to(np.asarray(mean, dtype=np.float64), (nv,)))
        dv = None
        if self.info.n_data:
            dv = np.ascontiguousarray(data_vals, dtype=np.float64)
            if dv.shape != (nv, self.info.n_data):
                raise ArgumentError("data_vals must hold one row of data values per variable")
        res = np.empty((nv, R, N))
        check(load().gsp_seqsim_sweep_multivariate(self.ctx.handle, self._h, R, nv, ptr(A),
                                                   ptr(mean), ptr(z), ptr(dv), ptr(res), 0))
        return res

    def sweep(self, z, data_vals=None, out=None, device_pointers=False):
        """z: (R, N) standard normals in path order. Returns/out: (R, N)."""
        if device_pointers:
            R = int(z.shape[0])
            check(load().gsp_seqsim_sweep(self.ctx.handle, self._h, R, ptr(z), ptr(data_vals),
                                          ptr(out), 1))
            return out
        z = np.ascontiguousarray(z, dtype=np.float64)
        if z.ndim == 1:
            z = z[None, :]
        R = z.shape[0]
        if z.shape[1] != self.info.n_nodes:
            raise ArgumentError("z must hold one standard normal per path node")
        dv = None
        if self.info.n_data:
            dv = np.ascontiguousarray(data_vals, dtype=np.float64)
            if dv.shape != (self.info.n_data,):
                raise ArgumentError("data_vals must hold one value per datum")
        res = np.empty((R, self.info.n_nodes)) if out is None else out
        check(load().gsp_seqsim_sweep(self.ctx.handle, self._h, R, ptr(z), ptr(dv), ptr(res), 0))
        return res

    def close(self):
        if self._h and self._h.value:
            load().gsp_seqsim_plan_free(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Surrogate:
    """gsp_surrogate: the learned kriging-weight MLP, trained on the systems of
    exact plans (gsp_surrogate_train)."""

    def __init__(self, ctx: Context, plans, max_samples=0, epochs=20, batch=1024, lr=2e-3,
                 var_weight=0.01, seed=0, n_slots=0, data_slots=0):
        self.ctx = ctx
        self._h = ctypes.c_void_p()
        opts = GspTrainOptions(max_samples=int(max_samples), epochs=int(epochs), batch=int(batch),
                               lr=float(lr), var_weight=float(var_weight), seed=int(seed),
                               n_slots=int(n_slots), data_slots=int(data_slots))
        handles = (ctypes.c_void_p * len(plans))(*[p._h.value for p in plans])
        self.report = GspTrainReport()
        check(load().gsp_surrogate_train(ctx.handle, len(plans), handles, ctypes.byref(opts),
                                         ctypes.byref(self._h), ctypes.byref(self.report)))

    @property
    def handle(self):
        return self._h

    def export(self):
        """(params FP32 flat array, slot offsets in grid cells (S, 3))."""
        n = ctypes.c_int64()
        check(load().gsp_surrogate_export(self.ctx.handle, self._h, None, ctypes.byref(n), None))
        params = np.zeros(n.value, np.float32)
        slots = np.zeros((32, 3))
        check(load().gsp_surrogate_export(self.ctx.handle, self._h, params.ctypes.data,
                                          ctypes.byref(n), slots.ctypes.data))
        return params, slots[:self.info["n_slots"]]

    @property
    def info(self):
        """The domain it serves: family, ndim, slots, the sill (gaussian), and
        the trained spans of nugget / sill and spacing / range per axis."""
        i = GspSurrogateInfo()
        check(load().gsp_surrogate_get_info(self._h, ctypes.byref(i)))
        return {"family": i.family, "ndim": i.ndim, "n_slots": i.n_slots,
                "n_train_requests": i.n_train_requests, "sill": i.sill,
                "nugget_ratio": tuple(i.nugget_ratio),
                "scale": [tuple(i.scale[d]) for d in range(3)], "n_samples": i.n_samples,
                "n_data_slots": i.n_data_slots}

    def close(self):
        if self._h and self._h.value:
            load().gsp_surrogate_free(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class LusimPlan:
    """gsp_lusim_plan: d2 and L22 of one variable."""

    def __init__(self, ctx: Context, problem: GspLusimProblem, keep=()):
        self.ctx = ctx
        self._keep = keep
        self._h = ctypes.c_void_p()
        check(load().gsp_lusim_factor(ctx.handle, ctypes.byref(problem), ctypes.byref(self._h)))
        ns, nd = ctypes.c_int64(), ctypes.c_int64()
        check(load().gsp_lusim_dims(self._h, ctypes.byref(ns), ctypes.byref(nd)))
        self.n_sim, self.n_data = ns.value, nd.value
        self.n_nodes = self.n_sim + self.n_data

    def export(self):
        L = np.zeros((self.n_sim, self.n_sim))
        d2 = np.zeros(self.n_sim)
        check(load().gsp_lusim_export(self.ctx.handle, self._h, ptr(L), ptr(d2)))
        return L, d2

    def sample(self, w, out=None, device_pointers=False):
        """w: (R, n_sim) standard normals. Returns (R, N)."""
        if device_pointers:
            R = int(w.shape[0])
            check(load().gsp_lusim_sample(self.ctx.handle, self._h, R, ptr(w), ptr(out), 1))
            return out
        w = np.ascontiguousarray(w, dtype=np.float64)
        if w.ndim == 1:
            w = w[None, :]
        if w.shape[1] != self.n_sim:
            raise ArgumentError("w must hold one standard normal per simulated cell")
        R = w.shape[0]
        res = np.empty((R, self.n_nodes)) if out is None else out
        check(load().gsp_lusim_sample(self.ctx.handle, self._h, R, ptr(w), ptr(res), 0))
        return res

    def close(self):
        if self._h and self._h.value:
            load().gsp_lusim_plan_free(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FftsimPlan:
    """gsp_fftsim_plan: amplitude spectrum and, with data, the kriging mean and
    the kriging weights of the data cells."""

    def __init__(self, ctx: Context, problem: GspFftsimProblem, keep=()):
        self.ctx = ctx
        self._keep = keep
        self._h = ctypes.c_void_p()
        check(load().gsp_fftsim_plan_create(ctx.handle, ctypes.byref(problem), ctypes.byref(self._h)))
        g = problem.grid
        self.dims = tuple(int(g.dims[d]) for d in range(g.ndim))
        self.n_nodes = int(np.prod(self.dims))
        self.conditional = problem.n_data > 0

    def export(self):
        """(amplitude on the half spectrum, shape (nz, ny, nx // 2 + 1) in C
        order for 3D, zbar or None)."""
        half = list(self.dims[::-1])
        half[-1] = half[-1] // 2 + 1
        F = np.zeros(half)
        zbar = np.zeros(self.n_nodes) if self.conditional else None
        check(load().gsp_fftsim_export(self.ctx.handle, self._h, ptr(F), ptr(zbar)))
        return F, zbar

    def sample(self, w, out=None, device_pointers=False):
        """w: (R, N) uniform [0, 1) noise, column-major grid order. Returns (R, N)."""
        if device_pointers:
            R = int(w.shape[0])
            check(load().gsp_fftsim_sample(self.ctx.handle, self._h, R, ptr(w), ptr(out), 1))
            return out
        w = np.ascontiguousarray(w, dtype=np.float64)
        if w.ndim == 1:
            w = w[None, :]
        if w.shape[1] != self.n_nodes:
            raise ArgumentError("w must hold one value per grid cell")
        R = w.shape[0]
        res = np.empty((R, self.n_nodes)) if out is None else out
        check(load().gsp_fftsim_sample(self.ctx.handle, self._h, R, ptr(w), ptr(res), 0))
        return res

    def close(self):
        if self._h and self._h.value:
            load().gsp_fftsim_plan_free(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ensemble_stats(ctx: Context, reals, stats, out=None, device_pointers=False, shape=None):
    """gsp_ensemble_stats: `stats` is a list of (kind, param) pairs; `reals` an
    (R, N) array (or a device buffer with `shape` = (R, N)). Returns
    (len(stats), N)."""
    kinds = np.ascontiguousarray([k for k, _ in stats], dtype=np.int32)
    params = np.ascontiguousarray([float(p) for _, p in stats], dtype=np.float64)
    if device_pointers:
        R, N = shape
        check(load().gsp_ensemble_stats(ctx.handle, R, N, ptr(reals), len(kinds), ptr(kinds),
                                        ptr(params), ptr(out), 1))
        return out
    reals = np.ascontiguousarray(reals, dtype=np.float64)
    if reals.ndim != 2:
        raise ArgumentError("reals must be (realizations, cells)")
    R, N = reals.shape
    res = np.empty((len(kinds), N)) if out is None else out
    check(load().gsp_ensemble_stats(ctx.handle, R, N, ptr(reals), len(kinds), ptr(kinds),
                                    ptr(params), ptr(res), 0))
    return res
