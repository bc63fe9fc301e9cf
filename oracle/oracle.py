"""ctypes binding of the fastreg oracle (liboracle_fastreg.so) — TEST INFRASTRUCTURE ONLY.

The oracle is the CPU restatement of the reference solver (fr_oracle.c).  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import this module; the product
path never does.

Arrays follow the reference layouts: volumes (N, 2), vector fields (N, 4), Jacobians
(N, 4, 4), scalar fields (N,), all flattened x-fastest (index = x + X*(y + Y*z)).  Precision
"f32" is the FP32 reference-order restatement, "f64" the same algorithm in double.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle_fastreg.so")
_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "all"], cwd=_HERE, check=True)
    return LIB_PATH


class OrParams(C.Structure):
    _fields_ = [("max_iter", C.c_int), ("reset_iter", C.c_int), ("verbosity", C.c_int), ("no_labels", C.c_int),
                ("max_update_norm", C.c_float), ("rho_0", C.c_float), ("tau", C.c_float), ("alpha", C.c_float),
                ("w_reg", C.c_float), ("p", C.c_float), ("C", C.c_float)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library not built: {LIB_PATH} (run `make -C oracle`)")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        for sfx, R in (("_f32", C.c_float), ("_f64", C.c_double)):
            def f(name, res, *args):
                fn = getattr(L, name + sfx)
                fn.restype = res
                fn.argtypes = list(args)
            i = C.c_int
            f("or_create", vp, i, i, i)
            f("or_set_storage", None, vp, i)
            f("or_destroy", None, vp)
            f("or_init", None, vp, vp, vp, vp, vp, R)
            f("or_step", R, vp, R, R, R, R, R, R, C.POINTER(C.c_long))
            f("or_energy", None, vp, C.POINTER(C.c_double), C.POINTER(C.c_double))
            f("or_run_stdout", i, vp, vp, vp, vp, C.POINTER(C.c_int), i)
            f("or_outputs", None, vp, vp, vp, vp, vp)
            f("or_dice", None, vp, i, vp, vp, C.POINTER(C.c_long))
            f("or_min_det", None, vp, C.POINTER(C.c_double), C.POINTER(C.c_double))
            f("or_get_state", None, vp, vp, vp, vp, vp)
            f("or_set_state", None, vp, vp, vp, vp, vp)
            f("or_get_field", i, vp, C.c_char_p, vp)
            f("ork_apply", None, i, i, i, vp, vp, vp)
            f("ork_get_inverse", None, i, i, i, vp, vp, i)
            f("ork_interpolate_gradient", None, i, i, i, vp, vp, vp)
            f("ork_voxel_max_energy", R, i, i, i, vp, vp, vp, R, C.POINTER(C.c_long))
            f("or_voxel_max_energy", R, vp, R, C.POINTER(C.c_long))
            f("ork_volume_gradient", None, i, i, i, vp, vp)
            f("ork_jacobian", None, i, i, i, vp, vp)
            f("ork_deformation_jacobian", None, i, i, i, vp, vp)
            f("ork_laplacian", None, i, i, i, vp, vp)
            f("ork_interpolate_laplacian", None, i, i, i, vp, vp, vp)
            f("ork_laplacian_velocity", None, i, i, i, vp, vp)
            f("ork_dv_v", None, i, i, i, vp, vp)
            f("ork_divergence", None, i, i, i, vp, vp, vp)
            f("ork_update_density", None, i, i, i, vp, vp, R)
            f("ork_potential_gradient", None, i, i, i, vp, vp, vp, vp, vp, vp, R)
            f("ork_update_velocity", None, i, i, i, vp, vp, vp, vp, vp, R, R, R, R, R)
            f("ork_update_psi", None, i, i, i, vp, vp, vp, R)
            f("ork_update_psi_inv", None, i, i, i, vp, vp, vp, R)
            f("ork_reduce_max", R, i, i, i, vp, C.POINTER(C.c_long))
            f("ork_det_ref", R, vp)
            f("ork_det_true", R, vp)
            f("ork_norm", R, vp)
            f("ork_lerp", R, R, R, R)
        L.or_num_threads.restype = C.c_int
        L.or_round16.restype = None
        L.or_round16.argtypes = [C.c_int, C.c_void_p, C.c_long]
        _lib = L
    return _lib


def _dt(prec):
    return np.float32 if prec == "f32" else np.float64


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _chk(a, dtype, shape):
    a = np.ascontiguousarray(a, dtype=dtype)
    if a.shape != shape:
        a = a.reshape(shape)
    return a


def num_threads() -> int:
    return lib().or_num_threads()


STORAGE_F32, STORAGE_BF16, STORAGE_F16 = 0, 1, 2


def round16(storage: int, a) -> np.ndarray:
    """The oracle's storage rounding (bf16 / fp16, nearest even) of a float32 array."""
    out = np.ascontiguousarray(a, dtype=np.float32).copy()
    lib().or_round16(storage, _p(out), out.size)
    return out


@dataclass
class SolverParams:
    """Solver parameters with the reference's params.ini defaults (params.ini:13-25, params.hpp:11-12)."""
    max_iter: int = 2048
    reset_iter: int = -1
    verbosity: int = 0
    no_labels: int = 6
    max_update_norm: float = -1e-10
    rho_0: float = 1.0
    tau: float = 0.2
    alpha: float = 0.004
    w_reg: float = 0.2
    p: float = 2.0
    C: float = 0.25

    def c_struct(self) -> OrParams:
        return OrParams(self.max_iter, self.reset_iter, self.verbosity, self.no_labels, self.max_update_norm,
                        self.rho_0, self.tau, self.alpha, self.w_reg, self.p, self.C)


class Oracle:
    """One registration's state in the oracle (precision 'f32' or 'f64')."""

    def __init__(self, dims, prec: str = "f32", storage: int = 0):
        """storage: 0 FP32 fields; 1 / 2 (FP32 build only): v and rho rounded to bf16 / fp16 when
        computed, as the library's 16-bit storage forms do."""
        self.X, self.Y, self.Z = (int(d) for d in dims)
        self.n = self.X * self.Y * self.Z
        self.prec = prec
        self.sfx = "_" + prec
        self.dtype = _dt(prec)
        self._L = lib()
        self.h = self._fn("or_create")(self.X, self.Y, self.Z)
        if not self.h:
            raise MemoryError("oracle context allocation failed")
        if storage and prec != "f32":
            raise ValueError("the 16-bit storage forms are modelled in the FP32 oracle only")
        self._fn("or_set_storage")(self.h, int(storage))

    def _fn(self, name):
        return getattr(self._L, name + self.sfx)

    def close(self):
        if self.h:
            self._fn("or_destroy")(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init(self, tgt, tgt_lab, src, src_lab, rho_0=1.0):
        n = self.n
        self._keep = [np.ascontiguousarray(tgt, np.float32).reshape(n), np.ascontiguousarray(tgt_lab, np.int16).reshape(n),
                      np.ascontiguousarray(src, np.float32).reshape(n), np.ascontiguousarray(src_lab, np.int16).reshape(n)]
        self._fn("or_init")(self.h, *[_p(a) for a in self._keep], rho_0)
        self._keep = None

    def step(self, t, alpha=0.004, tau=0.2, w_reg=0.2, p=2.0, C_=0.25):
        am = C.c_long(0)
        m = self._fn("or_step")(self.h, alpha, t, tau, w_reg, p, C_, C.byref(am))
        return m, am.value

    def energy(self):
        a, b = C.c_double(), C.c_double()
        self._fn("or_energy")(self.h, C.byref(a), C.byref(b))
        return a.value, b.value

    def voxel_max_energy(self, w_reg):
        """(max per-voxel energy, its voxel) of the current state (Reductor::voxel_max_energy)."""
        arg = C.c_long(0)
        e = self._fn("or_voxel_max_energy")(self.h, w_reg, C.byref(arg))
        return e, arg.value

    def run(self, params: SolverParams, log: bool = False):
        prm = params.c_struct()
        mh = np.zeros(max(params.max_iter, 1), self.dtype)
        eh = np.zeros(max(params.max_iter, 1), np.float64)
        reason = C.c_int(0)
        it = self._fn("or_run_stdout")(self.h, C.byref(prm), _p(mh), _p(eh), C.byref(reason), int(log))
        return {"iterations": it, "max_hist": mh[:it], "energy_hist": eh[:max(it, 1)], "stop_reason": reason.value}

    def outputs(self):
        n = self.n
        img1, img2 = np.zeros(n, np.float32), np.zeros(n, np.float32)
        lab1, lab2 = np.zeros(n, np.int16), np.zeros(n, np.int16)
        self._fn("or_outputs")(self.h, _p(img1), _p(lab1), _p(img2), _p(lab2))
        return img1, lab1, img2, lab2

    def dice_counts(self, no_labels):
        ov = np.zeros(no_labels, np.int64)
        vx = np.zeros(no_labels, np.int64)
        bad = C.c_long(0)
        self._fn("or_dice")(self.h, no_labels, _p(ov), _p(vx), C.byref(bad))
        return ov, vx, bad.value

    def min_det(self):
        a, b = C.c_double(), C.c_double()
        self._fn("or_min_det")(self.h, C.byref(a), C.byref(b))
        return a.value, b.value

    def get_state(self):
        n, dt = self.n, self.dtype
        v, rho, psi, pinv = np.zeros((n, 4), dt), np.zeros(n, dt), np.zeros((n, 4), dt), np.zeros((n, 4), dt)
        self._fn("or_get_state")(self.h, _p(v), _p(rho), _p(psi), _p(pinv))
        return {"v": v, "rho": rho, "psi": psi, "psi_inv": pinv}

    def set_state(self, v=None, rho=None, psi=None, psi_inv=None):
        n, dt = self.n, self.dtype
        arrs = [None if a is None else _chk(a, dt, s) for a, s in
                ((v, (n, 4)), (rho, (n,)), (psi, (n, 4)), (psi_inv, (n, 4)))]
        self._fn("or_set_state")(self.h, *[_p(a) for a in arrs])

    _FIELD_SHAPES = {"nabla_phi_n": 4, "L": 4, "L_o": 4, "dv_v": 4, "delta_v": 4, "nabla_U": 4, "updates": 4,
                     "J_inv": (4, 4), "J": (4, 4), "div": None, "phi_global_psi_inv": 2, "phi_n_psi": 2}

    def field(self, name):
        w = self._FIELD_SHAPES[name]
        shape = (self.n,) if w is None else ((self.n, w) if isinstance(w, int) else (self.n,) + w)
        out = np.zeros(shape, self.dtype)
        if self._fn("or_get_field")(self.h, name.encode(), _p(out)) != 0:
            raise KeyError(name)
        return out


# ---- single-kernel entry points on arrays -------------------------------------------------------

def _k(prec, name):
    return getattr(lib(), name + "_" + prec)


def get_inverse(dims, psi4, psi_inv4, sweeps=48, prec="f32"):
    """DeformationField::get_inverse: `sweeps` sweeps of psi_inv = -trilerp(psi, psi_inv)."""
    X, Y, Z = dims
    n = X * Y * Z
    dt = _dt(prec)
    psi4 = _chk(psi4, dt, (n, 4))
    out = _chk(psi_inv4, dt, (n, 4)).copy()
    _k(prec, "ork_get_inverse")(X, Y, Z, _p(psi4), _p(out), sweeps)
    return out


def interpolate_gradient(dims, grad4, psi4, prec="f32"):
    X, Y, Z = dims
    n = X * Y * Z
    dt = _dt(prec)
    grad4, psi4 = _chk(grad4, dt, (n, 4)), _chk(psi4, dt, (n, 4))
    out = np.zeros((n, 4), dt)
    _k(prec, "ork_interpolate_gradient")(X, Y, Z, _p(grad4), _p(out), _p(psi4))
    return out


def voxel_max_energy(dims, tgt2, src_psi2, J16, w_reg, prec="f32"):
    X, Y, Z = dims
    n = X * Y * Z
    dt = _dt(prec)
    tgt2, src_psi2, J16 = _chk(tgt2, dt, (n, 2)), _chk(src_psi2, dt, (n, 2)), _chk(J16, dt, (n, 4, 4))
    arg = C.c_long(0)
    e = _k(prec, "ork_voxel_max_energy")(X, Y, Z, _p(tgt2), _p(src_psi2), _p(J16), w_reg, C.byref(arg))
    return e, arg.value


def apply(dims, phi2, psi4, prec="f32"):
    X, Y, Z = dims
    n = X * Y * Z
    dt = _dt(prec)
    phi2, psi4 = _chk(phi2, dt, (n, 2)), _chk(psi4, dt, (n, 4))
    out = np.zeros((n, 2), dt)
    _k(prec, "ork_apply")(X, Y, Z, _p(phi2), _p(out), _p(psi4))
    return out


def volume_gradient(dims, vol2, prec="f32"):
    X, Y, Z = dims
    n = X * Y * Z
    dt = _dt(prec)
    vol2 = _chk(vol2, dt, (n, 2))
    out = np.zeros((n, 4), dt)
    _k(prec, "ork_volume_gradient")(X, Y, Z, _p(vol2), _p(out))
    return out


def _unary4(kname, out_shape_tail):
    def fn(dims, f4, prec="f32", out=None):
        X, Y, Z = dims
        n = X * Y * Z
        dt = _dt(prec)
        f4 = _chk(f4, dt, (n, 4))
        o = np.zeros((n,) + out_shape_tail, dt) if out is None else _chk(out, dt, (n,) + out_shape_tail).copy()
        _k(prec, kname)(X, Y, Z, _p(f4), _p(o))
        return o
    return fn


jacobian = _unary4("ork_jacobian", (4, 4))
deformation_jacobian = _unary4("ork_deformation_jacobian", (4, 4))
laplacian = _unary4("ork_laplacian", (4,))
laplacian_velocity = _unary4("ork_laplacian_velocity", (4,))
dv_v = _unary4("ork_dv_v", (4,))


def interpolate_laplacian(dims, L4, psi4, prec="f32"):
    X, Y, Z = dims
    n = X * Y * Z
    dt = _dt(prec)
    L4, psi4 = _chk(L4, dt, (n, 4)), _chk(psi4, dt, (n, 4))
    out = np.zeros((n, 4), dt)
    _k(prec, "ork_interpolate_laplacian")(X, Y, Z, _p(L4), _p(out), _p(psi4))
    return out


def divergence(dims, rho, v4, prec="f32"):
    X, Y, Z = dims
    n = X * Y * Z
    dt = _dt(prec)
    rho, v4 = _chk(rho, dt, (n,)), _chk(v4, dt, (n, 4))
    out = np.zeros(n, dt)
    _k(prec, "ork_divergence")(X, Y, Z, _p(rho), _p(v4), _p(out))
    return out


def update_density(dims, rho, div, alpha, prec="f32"):
    X, Y, Z = dims
    n = X * Y * Z
    dt = _dt(prec)
    r = _chk(rho, dt, (n,)).copy()
    div = _chk(div, dt, (n,))
    _k(prec, "ork_update_density")(X, Y, Z, _p(r), _p(div), alpha)
    return r


def potential_gradient(dims, phi_n2, phi_g_psi_inv2, nabla4, J16, Lo4, w_reg, prec="f32"):
    X, Y, Z = dims
    n = X * Y * Z
    dt = _dt(prec)
    a = [_chk(phi_n2, dt, (n, 2)), _chk(phi_g_psi_inv2, dt, (n, 2)), _chk(nabla4, dt, (n, 4)),
         _chk(J16, dt, (n, 4, 4)), _chk(Lo4, dt, (n, 4))]
    out = np.zeros((n, 4), dt)
    _k(prec, "ork_potential_gradient")(X, Y, Z, *[_p(x) for x in a], _p(out), w_reg)
    return out


def update_velocity(dims, v4, dvv4, dlt4, rho, nU4, alpha, t, tau, p, C_, prec="f32"):
    X, Y, Z = dims
    n = X * Y * Z
    dt = _dt(prec)
    v = _chk(v4, dt, (n, 4)).copy()
    a = [_chk(dvv4, dt, (n, 4)), _chk(dlt4, dt, (n, 4)), _chk(rho, dt, (n,)), _chk(nU4, dt, (n, 4))]
    _k(prec, "ork_update_velocity")(X, Y, Z, _p(v), *[_p(x) for x in a], alpha, t, tau, p, C_)
    return v


def update_psi(dims, psi4, v4, alpha, prec="f32"):
    X, Y, Z = dims
    n = X * Y * Z
    dt = _dt(prec)
    psi = _chk(psi4, dt, (n, 4)).copy()
    v = _chk(v4, dt, (n, 4))
    upd = np.zeros((n, 4), dt)
    _k(prec, "ork_update_psi")(X, Y, Z, _p(psi), _p(v), _p(upd), alpha)
    return psi, upd


def update_psi_inv(dims, psi_inv4, J16, v4, alpha, prec="f32"):
    X, Y, Z = dims
    n = X * Y * Z
    dt = _dt(prec)
    pinv = _chk(psi_inv4, dt, (n, 4)).copy()
    a = [_chk(J16, dt, (n, 4, 4)), _chk(v4, dt, (n, 4))]
    _k(prec, "ork_update_psi_inv")(X, Y, Z, _p(pinv), *[_p(x) for x in a], alpha)
    return pinv


def reduce_max(dims, upd4, prec="f32"):
    X, Y, Z = dims
    n = X * Y * Z
    dt = _dt(prec)
    u = _chk(upd4, dt, (n, 4))
    am = C.c_long(0)
    m = _k(prec, "ork_reduce_max")(X, Y, Z, _p(u), C.byref(am))
    return m, am.value


def det_ref(M, prec="f32"):
    M = _chk(M, _dt(prec), (4, 4))
    return _k(prec, "ork_det_ref")(_p(M))


def det_true(M, prec="f32"):
    M = _chk(M, _dt(prec), (4, 4))
    return _k(prec, "ork_det_true")(_p(M))


def norm(v, prec="f32"):
    v = _chk(v, _dt(prec), (4,))
    return _k(prec, "ork_norm")(_p(v))


def lerp(v0, v1, t, prec="f32"):
    return _k(prec, "ork_lerp")(v0, v1, t)
