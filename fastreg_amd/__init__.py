"""fastreg_amd — MI355X (gfx950) fastreg registration solver.

The product is the native library ``libfastreg_hip.so`` (HIP kernels + the C-ABI declared in
``include/fastreg_c.h``) and the ``fastreg_app`` CLI.  This module is a thin ctypes binding of
that C-ABI for Python callers (tests, bench.py): it holds no solver logic, and it raises if the
native library has not been built — there is no CPU fallback.

Device memory is handled through the library's own plumbing (``fr_dev_malloc`` and friends), so
no other GPU runtime is needed in the process.
"""
from __future__ import annotations

import ctypes as C
import os
import re
from dataclasses import dataclass

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT_DIR = os.path.dirname(PKG_DIR)
LIB_PATH = os.path.join(PKG_DIR, "libfastreg_hip.so")
HEADER_PATH = os.path.join(ROOT_DIR, "include", "fastreg_c.h")

FR_OK = 0
FR_ERR_INVALID = 1
FR_ERR_HIP = 2
STORAGE_F32, STORAGE_BF16, STORAGE_F16 = 0, 1, 2
ARITH_REFERENCE, ARITH_FMA = 0, 1
FR_ERR_ALLOC = 3
FR_ERR_IO = 4
FR_ERR_PARSE = 5
FR_ERR_STATE = 6


class FastRegError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"fastreg error {code}: {msg}")
        self.code = code


class FrParams(C.Structure):
    _fields_ = [("vol_dims", C.c_int * 3), ("vol_size", C.c_float * 3), ("no_labels", C.c_int),
                ("max_iter", C.c_int), ("max_update_norm", C.c_float), ("rho_0", C.c_float), ("tau", C.c_float),
                ("alpha", C.c_float), ("w_reg", C.c_float), ("p", C.c_float), ("C", C.c_float),
                ("reset_iter", C.c_int), ("verbosity", C.c_int)]


class FrStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("stop_reason", C.c_int), ("last_max_update", C.c_float),
                ("last_max_voxel", C.c_longlong), ("last_energy", C.c_double), ("min_det_true", C.c_double),
                ("min_det_ref", C.c_double), ("dice_all", C.c_double), ("solver_ms", C.c_float),
                ("total_ms", C.c_float)]


_lib = None


def header_functions():
    """Names of the functions declared in include/fastreg_c.h."""
    txt = open(HEADER_PATH).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fr_[a-z0-9_]+)\s*\(", txt)))


def lib():
    """Load libfastreg_hip.so (raises if it is missing: build with `python -m fastreg_amd.build`)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"native library missing: {LIB_PATH}; build it with `python -m fastreg_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp, i, f, d, ll, sz = C.c_void_p, C.c_int, C.c_float, C.c_double, C.c_longlong, C.c_size_t
    P = C.POINTER

    def sig(name, *args, res=C.c_int):
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = list(args)

    sig("fr_last_error", res=C.c_char_p)
    sig("fr_params_default", P(FrParams))
    sig("fr_params_load_ini", C.c_char_p, P(FrParams))
    sig("fr_create", P(FrParams), P(vp))
    sig("fr_destroy", vp)
    sig("fr_get_params", vp, P(FrParams))
    sig("fr_set_stream", vp, vp)
    sig("fr_set_chunking", vp, i, i)
    sig("fr_set_storage", vp, i)
    sig("fr_get_storage", vp, P(i))
    sig("fr_set_keep_state", vp, i)
    sig("fr_get_keep_state", vp, P(i))
    sig("fr_set_arith", vp, i)
    sig("fr_get_arith", vp, P(i))
    sig("fr_batch_create", P(vp), i, P(vp))
    sig("fr_batch_destroy", vp)
    sig("fr_batch_relayout", vp)
    sig("fr_batch_run", vp)
    sig("fr_batch_time", vp, P(f))
    sig("fr_batch_time_iterations", vp, i, P(f))
    sig("fr_get_chunking", vp, P(i), P(i))
    sig("fr_register", vp, vp, vp, vp, vp)
    sig("fr_get_deformation", vp, vp)
    sig("fr_get_inverse_deformation", vp, vp)
    sig("fr_get_stats", vp, P(FrStats))
    sig("fr_get_dice", vp, i, vp, vp, vp)
    sig("fr_get_history", vp, i, vp, vp)
    sig("fr_get_log", vp, C.c_char_p, sz)
    sig("fr_solver_init", vp, vp, vp, vp, vp)
    sig("fr_solver_step", vp, f, P(f), P(ll))
    sig("fr_solver_run", vp)
    sig("fr_solver_finalize", vp, vp, vp, vp, vp)
    sig("fr_solver_set_state", vp, vp, vp, vp, vp)
    sig("fr_solver_get_state", vp, vp, vp, vp, vp)
    sig("fr_solver_energy", vp, P(d), P(d))
    sig("fr_warp", i, i, i, vp, vp, vp, vp, vp, vp)
    sig("fr_field_identity", i, i, i, vp, vp)
    sig("fr_time_kernels", vp, i, P(f), P(f))
    sig("fr_get_inverse", i, i, i, vp, vp, vp)
    sig("fr_interpolate_gradient", i, i, i, vp, vp, vp, vp)
    sig("fr_voxel_max_energy", vp, P(f), P(ll))
    sig("fr_sum", vp, ll, P(d), vp)
    sig("fr_sum_noalloc", vp, ll, P(d), vp)
    sig("fr_adjoint_create", P(FrParams), i, i, P(vp))
    sig("fr_adjoint_destroy", vp)
    sig("fr_adjoint_set_images", vp, vp, vp, vp, vp)
    sig("fr_adjoint_forward", vp, vp, d, d, d, P(d))
    sig("fr_adjoint_backward", vp, vp, P(d), P(d))
    sig("fr_adjoint_get_state", vp, i, vp, vp, vp, vp)
    sig("fr_adjoint_info", vp, P(sz), P(f), P(f))
    sig("fr_adjoint_set_gather_chunk", vp, i)
    sig("fr_adjoint_get_gather_chunk", vp, P(i), P(i), P(i))
    sig("fr_step_fwd", vp, i, d, d, vp, vp, vp, vp, vp, vp, vp, vp)
    sig("fr_step_bwd", vp, i, d, d, vp, vp, vp, vp, vp, vp, vp, vp, P(d), P(d))
    sig("fr_warp_fwd", i, i, i, vp, vp, vp)
    sig("fr_warp_bwd", i, i, i, vp, vp, vp, vp, vp)
    sig("fr_gradient_fwd", i, i, i, vp, vp)
    sig("fr_gradient_bwd", i, i, i, vp, vp)
    sig("fr_laplacian_fwd", i, i, i, vp, vp)
    sig("fr_laplacian_bwd", i, i, i, vp, vp)
    sig("fr_dev_malloc", P(vp), sz)
    sig("fr_dev_free", vp)
    sig("fr_memcpy_htod", vp, vp, sz)
    sig("fr_memcpy_dtoh", vp, vp, sz)
    sig("fr_memcpy_dtod", vp, vp, sz)
    sig("fr_dev_memset", vp, i, sz)
    sig("fr_dev_sync")
    sig("fr_device_info", C.c_char_p, sz)
    _lib = L
    return L


def check(rc):
    if rc != FR_OK:
        raise FastRegError(rc, lib().fr_last_error().decode(errors="replace"))
    return rc


# ---- parameters -------------------------------------------------------------------------------

def default_params(**kw) -> FrParams:
    p = FrParams()
    check(lib().fr_params_default(C.byref(p)))
    for k, v in kw.items():
        if k == "vol_dims":
            p.vol_dims[:] = list(v)
        elif k == "vol_size":
            p.vol_size[:] = list(v)
        else:
            setattr(p, k, v)
    return p


def load_ini(path, base: FrParams | None = None) -> FrParams:
    p = base if base is not None else default_params()
    check(lib().fr_params_load_ini(os.fsencode(path), C.byref(p)))
    return p


# ---- device buffers ---------------------------------------------------------------------------

class DeviceArray:
    """A device allocation holding a numpy-typed array (via the library's fr_dev_* plumbing)."""

    def __init__(self, shape, dtype):
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        ptr = C.c_void_p()
        check(lib().fr_dev_malloc(C.byref(ptr), self.nbytes))
        self.ptr = ptr

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.nbytes != self.nbytes:
            raise ValueError(f"upload of {a.nbytes} bytes into a device array of {self.nbytes} bytes")
        check(lib().fr_memcpy_htod(self.ptr, a.ctypes.data_as(C.c_void_p), self.nbytes))

    def get(self):
        out = np.empty(self.shape, self.dtype)
        check(lib().fr_memcpy_dtoh(out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes))
        return out

    def free(self):
        if self.ptr is not None and self.ptr.value:
            lib().fr_dev_free(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_info() -> str:
    buf = C.create_string_buffer(512)
    check(lib().fr_device_info(buf, 512))
    return buf.value.decode()


def warp(img, lab, psi4, dims):
    """apply(): trilinear image and floor-corner labels at psi (float4 per voxel), on the GPU."""
    X, Y, Z = dims
    n = X * Y * Z
    d_img = DeviceArray.from_host(np.asarray(img, np.float32).reshape(n))
    d_lab = DeviceArray.from_host(np.asarray(lab, np.int16).reshape(n))
    d_psi = DeviceArray.from_host(np.asarray(psi4, np.float32).reshape(n, 4))
    o_img, o_lab = DeviceArray(n, np.float32), DeviceArray(n, np.int16)
    check(lib().fr_warp(X, Y, Z, d_img.ptr, d_lab.ptr, d_psi.ptr, o_img.ptr, o_lab.ptr, None))
    return o_img.get(), o_lab.get()


def get_inverse(psi4, psi_inv4, dims):
    """DeformationField::get_inverse on the GPU: 48 sweeps of psi_inv = -trilerp(psi, psi_inv) from
    the given psi_inv (float4 per voxel); returns the result."""
    X, Y, Z = dims
    n = X * Y * Z
    d_psi = DeviceArray.from_host(np.asarray(psi4, np.float32).reshape(n, 4))
    d_inv = DeviceArray.from_host(np.asarray(psi_inv4, np.float32).reshape(n, 4))
    check(lib().fr_get_inverse(X, Y, Z, d_psi.ptr, d_inv.ptr, None))
    return d_inv.get()


def interpolate_gradient(field4, psi4, dims):
    """interpolate_gradient on the GPU: field (float4 per voxel) trilinearly sampled at psi."""
    X, Y, Z = dims
    n = X * Y * Z
    d_f = DeviceArray.from_host(np.asarray(field4, np.float32).reshape(n, 4))
    d_psi = DeviceArray.from_host(np.asarray(psi4, np.float32).reshape(n, 4))
    out = DeviceArray((n, 4), np.float32)
    check(lib().fr_interpolate_gradient(X, Y, Z, d_f.ptr, out.ptr, d_psi.ptr, None))
    return out.get()


def field_sum(a):
    """ScalarField::sum on the GPU (FP64 accumulation of a float32 array)."""
    a = np.ascontiguousarray(a, np.float32).reshape(-1)
    d_a = DeviceArray.from_host(a)
    out = C.c_double()
    check(lib().fr_sum(d_a.ptr, a.size, C.byref(out), None))
    return out.value


# ---- solver context ---------------------------------------------------------------------------

@dataclass
class Stats:
    iterations: int
    stop_reason: int
    last_max_update: float
    last_max_voxel: int
    last_energy: float
    min_det_true: float
    min_det_ref: float
    dice_all: float
    solver_ms: float
    total_ms: float


class Solver:
    """One fr_ctx: the FastReg pipeline for one volume size and parameter set."""

    def __init__(self, params: FrParams):
        self.params = params
        self.dims = tuple(params.vol_dims)
        self.n = self.dims[0] * self.dims[1] * self.dims[2]
        h = C.c_void_p()
        check(lib().fr_create(C.byref(params), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().fr_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # pipeline
    def register(self, tgt, tgt_lab, src, src_lab):
        """FastReg::operator() on host arrays; returns (img1, lab1, img2, lab2) = (T o psi_inv,
        LT o psi_inv, S o psi, LS o psi)."""
        n = self.n
        bufs = [DeviceArray.from_host(np.asarray(tgt, np.float32).reshape(n)),
                DeviceArray.from_host(np.asarray(tgt_lab, np.int16).reshape(n)),
                DeviceArray.from_host(np.asarray(src, np.float32).reshape(n)),
                DeviceArray.from_host(np.asarray(src_lab, np.int16).reshape(n))]
        check(lib().fr_register(self.h, *[b.ptr for b in bufs]))
        return tuple(b.get() for b in bufs)

    # solver pieces
    def init(self, tgt, tgt_lab, src, src_lab):
        n = self.n
        self._inputs = [DeviceArray.from_host(np.asarray(tgt, np.float32).reshape(n)),
                        DeviceArray.from_host(np.asarray(tgt_lab, np.int16).reshape(n)),
                        DeviceArray.from_host(np.asarray(src, np.float32).reshape(n)),
                        DeviceArray.from_host(np.asarray(src_lab, np.int16).reshape(n))]
        check(lib().fr_solver_init(self.h, *[b.ptr for b in self._inputs]))

    def reinit(self):
        """fr_solver_init again from the inputs of the last init (already on the device)."""
        check(lib().fr_solver_init(self.h, *[b.ptr for b in self._inputs]))

    def step(self, t):
        m, am = C.c_float(), C.c_longlong()
        check(lib().fr_solver_step(self.h, t, C.byref(m), C.byref(am)))
        return m.value, am.value

    def run(self):
        check(lib().fr_solver_run(self.h))

    def finalize(self, want_outputs=True):
        n = self.n
        if want_outputs:
            outs = [DeviceArray(n, np.float32), DeviceArray(n, np.int16), DeviceArray(n, np.float32),
                    DeviceArray(n, np.int16)]
            check(lib().fr_solver_finalize(self.h, *[b.ptr for b in outs]))
            return tuple(b.get() for b in outs)
        check(lib().fr_solver_finalize(self.h, None, None, None, None))
        return None

    def set_state(self, v=None, rho=None, psi=None, psi_inv=None):
        n = self.n
        arrs = []
        for a, shape in ((v, (n, 4)), (rho, (n,)), (psi, (n, 4)), (psi_inv, (n, 4))):
            arrs.append(None if a is None else DeviceArray.from_host(np.asarray(a, np.float32).reshape(shape)))
        check(lib().fr_solver_set_state(self.h, *[None if a is None else a.ptr for a in arrs]))

    def get_state(self):
        n = self.n
        v, rho, psi, pinv = DeviceArray((n, 4), np.float32), DeviceArray(n, np.float32), \
            DeviceArray((n, 4), np.float32), DeviceArray((n, 4), np.float32)
        check(lib().fr_solver_get_state(self.h, v.ptr, rho.ptr, psi.ptr, pinv.ptr))
        return {"v": v.get(), "rho": rho.get(), "psi": psi.get(), "psi_inv": pinv.get()}

    def energy(self):
        a, b = C.c_double(), C.c_double()
        check(lib().fr_solver_energy(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def stats(self) -> Stats:
        s = FrStats()
        check(lib().fr_get_stats(self.h, C.byref(s)))
        return Stats(**{k: getattr(s, k) for k, _ in FrStats._fields_})

    def history(self, n=None):
        n = self.params.max_iter if n is None else n
        m = np.zeros(max(n, 1), np.float32)
        e = np.zeros(max(n, 1), np.float64)
        check(lib().fr_get_history(self.h, n, m.ctypes.data_as(C.c_void_p), e.ctypes.data_as(C.c_void_p)))
        return m[:n], e[:n]

    def dice(self, n=None):
        n = self.params.no_labels if n is None else n
        dice = np.zeros(n, np.float64)
        ov = np.zeros(n, np.int64)
        vx = np.zeros(n, np.int64)
        check(lib().fr_get_dice(self.h, n, dice.ctypes.data_as(C.c_void_p), ov.ctypes.data_as(C.c_void_p),
                                vx.ctypes.data_as(C.c_void_p)))
        return dice, ov, vx

    def log(self, size=1 << 20):
        buf = C.create_string_buffer(size)
        check(lib().fr_get_log(self.h, buf, size))
        return buf.value.decode()

    def voxel_max_energy(self):
        """Reductor::voxel_max_energy of the current state: (energy, voxel)."""
        e, v = C.c_float(), C.c_longlong()
        check(lib().fr_voxel_max_energy(self.h, C.byref(e), C.byref(v)))
        return e.value, v.value

    def set_storage(self, storage):
        """STORAGE_F32 / STORAGE_BF16 / STORAGE_F16 storage of v and rho; discards the state
        (init again)."""
        check(lib().fr_set_storage(self.h, int(storage)))

    def storage(self):
        v = C.c_int()
        check(lib().fr_get_storage(self.h, C.byref(v)))
        return v.value

    def set_keep_state(self, keep=True):
        """Keep v, psi and psi_inv from one registration to the next (new images, rho reset), as
        the reference's FastReg does; off by default (every registration starts fresh)."""
        check(lib().fr_set_keep_state(self.h, 1 if keep else 0))

    def keep_state(self):
        v = C.c_int()
        check(lib().fr_get_keep_state(self.h, C.byref(v)))
        return bool(v.value)

    def set_arith(self, arith):
        """ARITH_REFERENCE (the reference's operation order, the default) or ARITH_FMA (its
        multiply-adds contracted; FP32 storage only).  The state is kept."""
        check(lib().fr_set_arith(self.h, int(arith)))

    def arith(self):
        v = C.c_int()
        check(lib().fr_get_arith(self.h, C.byref(v)))
        return v.value

    def set_chunking(self, zc_a=0, zc_b=0):
        """Interior z-slices per block of pass A and pass B (0 = the automatic choice)."""
        check(lib().fr_set_chunking(self.h, zc_a, zc_b))

    def chunking(self):
        """(zc_a, zc_b) in use."""
        v = [C.c_int() for _ in range(2)]
        check(lib().fr_get_chunking(self.h, *[C.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def time_kernels(self, reps=20):
        """Average device ms per launch of (pass A, pass B), timed inside `reps` iterations run
        in sequence."""
        a, b = C.c_float(), C.c_float()
        check(lib().fr_time_kernels(self.h, reps, C.byref(a), C.byref(b)))
        return a.value, b.value


class Batch:
    """fr_batch: the iteration loops of several Solvers (volumes of any sizes, the same iteration
    schedule) run together, one launch per pass per iteration, replayed from a hipGraph."""

    def __init__(self, solvers):
        self.solvers = list(solvers)
        arr = (C.c_void_p * len(self.solvers))(*[s.h.value for s in self.solvers])
        h = C.c_void_p()
        check(lib().fr_batch_create(arr, len(self.solvers), C.byref(h)))
        self.h = h

    def relayout(self):
        check(lib().fr_batch_relayout(self.h))

    def run(self):
        check(lib().fr_batch_run(self.h))

    def time_ms(self):
        ms = C.c_float()
        check(lib().fr_batch_time(self.h, C.byref(ms)))
        return ms.value

    def time_iterations(self, iters):
        """Device ms of `iters` iterations with the stop tests off; the volumes must be
        re-initialised afterwards (Solver.reinit)."""
        ms = C.c_float()
        check(lib().fr_batch_time_iterations(self.h, iters, C.byref(ms)))
        return ms.value

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().fr_batch_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Adjoint:
    """fr_adjoint: K unrolled iterations of the solver in FP64 from (RHO_0, v0, identity,
    identity), the objective E = 0.5 sum (T - S o psi)^2 + label_weight * (label-overlap
    surrogate) of the final state, and its gradients with respect to v0, W_REG and TAU through
    hand-written adjoint kernels, with checkpoints every `ckpt` iterations."""

    def __init__(self, params: FrParams, K: int, ckpt: int):
        self.params = params
        self.dims = tuple(params.vol_dims)
        self.n = self.dims[0] * self.dims[1] * self.dims[2]
        self.K, self.ckpt = K, ckpt
        h = C.c_void_p()
        check(lib().fr_adjoint_create(C.byref(params), K, ckpt, C.byref(h)))
        self.h = h
        self._v0 = None
        self.forward_calls = 0  # counts forward(); backward() differentiates the latest one

    def set_images(self, tgt, tgt_lab, src, src_lab):
        n = self.n
        bufs = [DeviceArray.from_host(np.asarray(tgt, np.float32).reshape(n)),
                DeviceArray.from_host(np.asarray(tgt_lab, np.int16).reshape(n)),
                DeviceArray.from_host(np.asarray(src, np.float32).reshape(n)),
                DeviceArray.from_host(np.asarray(src_lab, np.int16).reshape(n))]
        check(lib().fr_adjoint_set_images(self.h, *[b.ptr for b in bufs]))

    def forward(self, v0=None, w_reg=None, tau=None, label_weight=0.0) -> float:
        """E after K iterations from v0 ((N, 3) float64 host array, a DeviceArray, or None for 0);
        w_reg / tau default to the parameters' W_REG / TAU (as float32 values)."""
        w = float(np.float32(self.params.w_reg)) if w_reg is None else float(w_reg)
        t = float(np.float32(self.params.tau)) if tau is None else float(tau)
        ptr = None
        if isinstance(v0, DeviceArray):
            ptr = v0.ptr
        elif v0 is not None:
            if self._v0 is None:
                self._v0 = DeviceArray((self.n, 3), np.float64)
            self._v0.upload(np.asarray(v0, np.float64).reshape(self.n, 3))
            ptr = self._v0.ptr
        e = C.c_double()
        check(lib().fr_adjoint_forward(self.h, ptr, w, t, float(label_weight), C.byref(e)))
        self.forward_calls += 1
        return e.value

    def backward(self):
        """(dE/dv0 as (N, 3) float64, dE/dW_REG, dE/dTAU) of the last forward pass."""
        g = DeviceArray((self.n, 3), np.float64)
        gw, gt = C.c_double(), C.c_double()
        check(lib().fr_adjoint_backward(self.h, g.ptr, C.byref(gw), C.byref(gt)))
        out = g.get()
        g.free()
        return out, gw.value, gt.value

    def state(self, k):
        """The state after k iterations (k a multiple of ckpt below K, or K): dict of v (N, 3),
        rho (N), psi (N, 3), psi_inv (N, 3), float64."""
        n = self.n
        b = {"v": DeviceArray((n, 3), np.float64), "rho": DeviceArray(n, np.float64),
             "psi": DeviceArray((n, 3), np.float64), "psi_inv": DeviceArray((n, 3), np.float64)}
        check(lib().fr_adjoint_get_state(self.h, k, b["v"].ptr, b["rho"].ptr, b["psi"].ptr, b["psi_inv"].ptr))
        return {k_: a.get() for k_, a in b.items()}

    def _state_bufs(self, st):
        n = self.n
        return [DeviceArray.from_host(np.asarray(st["rho"], np.float64).reshape(n)),
                DeviceArray.from_host(np.asarray(st["v"], np.float64).reshape(n, 3)),
                DeviceArray.from_host(np.asarray(st["psi"], np.float64).reshape(n, 3)),
                DeviceArray.from_host(np.asarray(st["psi_inv"], np.float64).reshape(n, 3))]

    def step_fwd(self, k, state, w_reg, tau):
        """One iteration (number k) from `state` (dict rho, v, psi, psi_inv): the next state."""
        ins = self._state_bufs(state)
        outs = [DeviceArray(b.shape, np.float64) for b in ins]
        check(lib().fr_step_fwd(self.h, k, float(w_reg), float(tau), *[b.ptr for b in ins], *[b.ptr for b in outs]))
        return dict(zip(("rho", "v", "psi", "psi_inv"), (b.get() for b in outs)))

    def step_bwd(self, k, state, grad_next, w_reg, tau):
        """Adjoint of one iteration: (adjoint of `state` as a dict, dE/dW_REG, dE/dTAU) from the
        adjoint of the next state `grad_next` (dict rho, v, psi, psi_inv)."""
        ins = self._state_bufs(state)
        gs = self._state_bufs(grad_next)
        gw, gt = C.c_double(), C.c_double()
        check(lib().fr_step_bwd(self.h, k, float(w_reg), float(tau), *[b.ptr for b in ins], *[b.ptr for b in gs],
                                C.byref(gw), C.byref(gt)))
        self.forward_calls += 1  # the handle's last forward pass is no longer differentiable
        return dict(zip(("rho", "v", "psi", "psi_inv"), (b.get() for b in gs))), gw.value, gt.value

    def set_gather_chunk(self, zc):
        """z-planes per block of the backward pass's record gathers (0 = automatic)."""
        check(lib().fr_adjoint_set_gather_chunk(self.h, int(zc)))

    def gather_chunk(self):
        """(tile width, tile height, z-planes per block) of the backward pass's record gathers."""
        tx, ty, zc = C.c_int(), C.c_int(), C.c_int()
        check(lib().fr_adjoint_get_gather_chunk(self.h, C.byref(tx), C.byref(ty), C.byref(zc)))
        return tx.value, ty.value, zc.value

    def info(self):
        """(device bytes held, device ms of the last forward, of the last backward)."""
        by, f, b = C.c_size_t(), C.c_float(), C.c_float()
        check(lib().fr_adjoint_info(self.h, C.byref(by), C.byref(f), C.byref(b)))
        return by.value, f.value, b.value

    def close(self):
        if getattr(self, "_v0", None) is not None:
            self._v0.free()
            self._v0 = None
        if getattr(self, "h", None) is not None and self.h.value:
            lib().fr_adjoint_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- single operators as FP64 forward/backward pairs ---------------------------------------------

def _dev(a, shape):
    return DeviceArray.from_host(np.ascontiguousarray(np.asarray(a, np.float64).reshape(shape)))


def warp_fwd(f, pos, dims):
    """trilerp(f, pos(p)) for every voxel (f: N, pos: (N, 3) float64)."""
    X, Y, Z = dims
    n = X * Y * Z
    fb, pb, ob = _dev(f, n), _dev(pos, (n, 3)), DeviceArray(n, np.float64)
    check(lib().fr_warp_fwd(X, Y, Z, fb.ptr, pb.ptr, ob.ptr))
    return ob.get()


def warp_bwd(f, pos, grad_out, dims):
    """(dE/df, dE/dpos) of E with dE/d(warp_fwd) = grad_out."""
    X, Y, Z = dims
    n = X * Y * Z
    fb, pb, gb = _dev(f, n), _dev(pos, (n, 3)), _dev(grad_out, n)
    gf, gp = _dev(np.zeros(n), n), _dev(np.zeros((n, 3)), (n, 3))
    check(lib().fr_warp_bwd(X, Y, Z, fb.ptr, pb.ptr, gb.ptr, gf.ptr, gp.ptr))
    return gf.get(), gp.get()


def gradient_fwd(f, dims):
    X, Y, Z = dims
    n = X * Y * Z
    fb, gb = _dev(f, n), DeviceArray((n, 3), np.float64)
    check(lib().fr_gradient_fwd(X, Y, Z, fb.ptr, gb.ptr))
    return gb.get()


def gradient_bwd(grad_grad, dims):
    X, Y, Z = dims
    n = X * Y * Z
    gg, gf = _dev(grad_grad, (n, 3)), _dev(np.zeros(n), n)
    check(lib().fr_gradient_bwd(X, Y, Z, gg.ptr, gf.ptr))
    return gf.get()


def laplacian_fwd(psi, dims):
    X, Y, Z = dims
    n = X * Y * Z
    pb, lb = _dev(psi, (n, 3)), DeviceArray((n, 3), np.float64)
    check(lib().fr_laplacian_fwd(X, Y, Z, pb.ptr, lb.ptr))
    return lb.get()


def laplacian_bwd(grad_L, dims):
    X, Y, Z = dims
    n = X * Y * Z
    gl, gp = _dev(grad_L, (n, 3)), _dev(np.zeros((n, 3)), (n, 3))
    check(lib().fr_laplacian_bwd(X, Y, Z, gl.ptr, gp.ptr))
    return gp.get()
