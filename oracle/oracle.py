"""TEST INFRASTRUCTURE ONLY: ctypes front end of the CPU oracle.

Loads ``oracle/liboracle_td2d.so`` (the from-scratch C restatement in
``td2d_oracle.c``) and, when present, ``oracle/_ref/libtd2d_ref.so`` (the
reference's own CPU path compiled from its sources by ``oracle/Makefile``).
Only tests, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg
import this module; the product package never does.

Every array argument follows the reference's layouts
(torchwi/propagator/td2d/wrapper_cpu.cpp:44-112):
  vpad  (dimy*dimx,) x-fastest padded model
  w     (nt,) wavelet
  frd   (nt, nx) shot gather, nx = dimx - order
  virt  (nt, dimy*dimx) virtual source (full wavefield storage)
  resid (nt, nx) residual (the gather's adjoint input)
  grad  (dimy*dimx,) gradient on the padded grid
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_td2d.so")
_REF_PATH = os.path.join(_HERE, "_ref", "libtd2d_ref.so")

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i16p = np.ctypeslib.ndpointer(dtype=np.int16, flags="C_CONTIGUOUS")
_ci, _cf = ctypes.c_int, ctypes.c_float

_lib = None
_ref = None


def build():
    """Compile the restatement (and the reference, when its sources exist)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    if os.path.isdir("/root/reference/torchwi/propagator/td2d"):
        subprocess.run(["make", "-s", "-C", _HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        geo = [_ci, _ci, _ci, _ci, _cf, _cf]
        L.orc_forward_f32.argtypes = [_f32p, ctypes.c_void_p, _f32p, _f32p, _f32p, _f32p, _f32p] + geo + [_cf, _cf, _cf]
        L.orc_backward_f32.argtypes = [_f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p] + geo + [_cf]
        L.orc_exa_forward_f32.argtypes = [_f32p, _f32p, _i16p, _f32p, _f32p, _f32p, _f32p, _f32p] + geo + [_cf, _cf, _cf]
        L.orc_exa_backward_f32.argtypes = [_f32p, _f32p, _i16p, _f32p, _f32p, _f32p, _f32p, _f32p] + geo + [_cf]
        L.orc_forward_f64.argtypes = [_f64p, ctypes.c_void_p, _f64p, _f64p, _f64p, _f64p, _f64p] + geo + [_cf, _cf, _cf]
        L.orc_backward_f64.argtypes = [_f64p, _f64p, _f64p, _f64p, _f64p, _f64p, _f64p] + geo + [_cf]
        L.orc_step_f32.argtypes = [_f32p, _f32p, _f32p, _f32p, _ci, _ci, _ci, _cf, _cf]
        L.orc_step_f64.argtypes = [_f64p, _f64p, _f64p, _f64p, _ci, _ci, _ci, ctypes.c_double, ctypes.c_double]
        for f in (L.orc_forward_f32, L.orc_backward_f32, L.orc_exa_forward_f32,
                  L.orc_exa_backward_f32, L.orc_forward_f64, L.orc_backward_f64,
                  L.orc_step_f32, L.orc_step_f64):
            f.restype = _ci
        _lib = L
    return _lib


def ref_available():
    return os.path.exists(_REF_PATH)


def ref():
    """The reference's compiled CPU path, or None when it was not built."""
    global _ref
    if _ref is None and ref_available():
        R = ctypes.CDLL(_REF_PATH)
        geo = [_ci, _ci, _ci, _ci, _cf, _cf]
        R.ref_forward.argtypes = [_f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p] + geo + [_cf, _cf, _cf]
        R.ref_backward.argtypes = [_f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p] + geo + [_cf]
        R.ref_forward_only.argtypes = [_f32p, _f32p, _f32p, _f32p, _f32p, _f32p] + geo + [_cf, _cf, _cf]
        for f in (R.ref_forward, R.ref_backward, R.ref_forward_only):
            f.restype = None
        _ref = R
    return _ref


def _check(rc):
    if rc != 0:
        raise ValueError("oracle: unsupported order (use 2, 4 or 8)")


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32).ravel()


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64).ravel()


def forward(vpad, w, order, dimx, dimy, h, dt, sx, sy, ry, store_virt=True, impl="oracle"):
    """FP32 forward modelling.  Returns (frd (nt,nx), virt (nt,dimxy) or None).

    impl="oracle" runs the restatement; impl="ref" runs the compiled reference."""
    vpad, w = _f32(vpad), _f32(w)
    nt, nx, n = len(w), dimx - order, dimx * dimy
    frd = np.zeros(nt * nx, np.float32)
    virt = np.zeros(nt * n, np.float32) if store_virt else None
    u = [np.zeros(n, np.float32) for _ in range(3)]
    if impl == "ref":
        R = ref()
        if store_virt:
            R.ref_forward(frd, virt, *u, vpad, w, order, dimx, dimy, nt, h, dt, sx, sy, ry)
        else:
            R.ref_forward_only(frd, *u, vpad, w, order, dimx, dimy, nt, h, dt, sx, sy, ry)
    else:
        vp = virt.ctypes.data_as(ctypes.c_void_p) if store_virt else None
        _check(lib().orc_forward_f32(frd, vp, *u, vpad, w, order, dimx, dimy, nt, h, dt, sx, sy, ry))
    return frd.reshape(nt, nx), (virt.reshape(nt, n) if store_virt else None)


def backward(virt, vpad, resid, order, dimx, dimy, h, dt, ry, impl="oracle"):
    """FP32 adjoint + imaging condition.  Returns grad (dimy*dimx,) padded grid."""
    vpad, virt, resid = _f32(vpad), _f32(virt), _f32(resid)
    n = dimx * dimy
    nt = resid.size // (dimx - order)
    grad = np.zeros(n, np.float32)
    u = [np.zeros(n, np.float32) for _ in range(3)]
    if impl == "ref":
        ref().ref_backward(grad, virt, *u, vpad, resid, order, dimx, dimy, nt, h, dt, ry)
    else:
        _check(lib().orc_backward_f32(grad, virt, *u, vpad, resid, order, dimx, dimy, nt, h, dt, ry))
    return grad


def exa_forward(vpad, w, order, dimx, dimy, h, dt, sx, sy, ry):
    vpad, w = _f32(vpad), _f32(w)
    nt, nx, n = len(w), dimx - order, dimx * dimy
    frd = np.zeros(nt * nx, np.float32)
    exa = np.zeros(n, np.float32)
    iexa = np.zeros(n, np.int16)
    u = [np.zeros(n, np.float32) for _ in range(3)]
    _check(lib().orc_exa_forward_f32(frd, exa, iexa, *u, vpad, w, order, dimx, dimy, nt, h, dt, sx, sy, ry))
    return frd.reshape(nt, nx), exa, iexa


def exa_backward(exa, iexa, vpad, resid, order, dimx, dimy, h, dt, ry):
    vpad, exa, resid = _f32(vpad), _f32(exa), _f32(resid)
    iexa = np.ascontiguousarray(iexa, dtype=np.int16).ravel()
    n = dimx * dimy
    nt = resid.size // (dimx - order)
    grad = np.zeros(n, np.float32)
    u = [np.zeros(n, np.float32) for _ in range(3)]
    _check(lib().orc_exa_backward_f32(grad, exa, iexa, *u, vpad, resid, order, dimx, dimy, nt, h, dt, ry))
    return grad


def forward64(vpad, w, order, dimx, dimy, h, dt, sx, sy, ry, store_virt=True):
    """FP64 forward of the same recursion (geometry truncated as in FP32)."""
    vpad, w = _f64(vpad), _f64(w)
    nt, nx, n = len(w), dimx - order, dimx * dimy
    frd = np.zeros(nt * nx, np.float64)
    virt = np.zeros(nt * n, np.float64) if store_virt else None
    u = [np.zeros(n, np.float64) for _ in range(3)]
    vp = virt.ctypes.data_as(ctypes.c_void_p) if store_virt else None
    _check(lib().orc_forward_f64(frd, vp, *u, vpad, w, order, dimx, dimy, nt, h, dt, sx, sy, ry))
    return frd.reshape(nt, nx), (virt.reshape(nt, n) if store_virt else None)


def backward64(virt, vpad, resid, order, dimx, dimy, h, dt, ry):
    vpad, virt, resid = _f64(vpad), _f64(virt), _f64(resid)
    n = dimx * dimy
    nt = resid.size // (dimx - order)
    grad = np.zeros(n, np.float64)
    u = [np.zeros(n, np.float64) for _ in range(3)]
    _check(lib().orc_backward_f64(grad, virt, *u, vpad, resid, order, dimx, dimy, nt, h, dt, ry))
    return grad


def step32(u1, u2, vpad, order, dimx, dimy, h, dt, u3=None):
    """One interior + Keys update from arbitrary (u1, u2) (FP32 mirror)."""
    u3 = np.zeros(dimx * dimy, np.float32) if u3 is None else _f32(u3).copy()
    _check(lib().orc_step_f32(u3, _f32(u1), _f32(u2), _f32(vpad), order, dimx, dimy, h, dt))
    return u3


def step64(u1, u2, vpad, order, dimx, dimy, h, dt, u3=None):
    u3 = np.zeros(dimx * dimy, np.float64) if u3 is None else _f64(u3).copy()
    _check(lib().orc_step_f64(u3, _f64(u1), _f64(u2), _f64(vpad), order, dimx, dimy, h, dt))
    return u3


def pad_vel(vel, order):
    """Edge-replicate vel (nx, ny), y fastest, into vpad (dimy*dimx,) x-fastest,
    as TimeOperator.pad_vel does with ReplicationPad2d (TimeOperator.py:55-58)."""
    ne = order // 2
    v = np.asarray(vel, dtype=np.float32).T  # (ny, nx)
    return np.ascontiguousarray(np.pad(v, ne, mode="edge")).ravel()


def crop_grad(grad, order, nx, ny):
    """Crop a padded-grid gradient to the model and transpose to (nx, ny)
    (TimeFunction.py:143-145)."""
    ne = order // 2
    g = np.asarray(grad).reshape(ny + order, nx + order)[ne:ne + ny, ne:ne + nx]
    return np.ascontiguousarray(g.T)
