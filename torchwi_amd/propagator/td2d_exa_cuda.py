"""Excitation-amplitude (EXA) forward / gradient on the GPU.

Same functions and argument lists as the reference's ``td2d_exa_cuda``
module (wrapper_exa_cuda.cpp:9-89).  EXA keeps, per cell, the virtual source
of largest magnitude and its time step (int16), so memory is O(dimx*dimy)
instead of O(nt*dimx*dimy); the gradient is approximate by construction."""
import torch

from .. import _lib
from ._td2d_common import check_common, device_tensor, wavelet_on


def forward(frd, exa, iexa, u1, u2, u3, vpad, w, order, dimx, dimy, nt, h, dt, sx, sy, ry):
    """Reference: fdm_forward, wrapper_exa_cuda.cpp:9-45 -> cuda_fdm_forward,
    propagator_exa_cuda.cu:5-84."""
    check_common(order, dimx, dimy, nt)
    n = dimx * dimy
    nx = dimx - order
    device_tensor("forward", frd, numel=nt * nx)
    device_tensor("exa", exa, numel=n)
    device_tensor("iexa", iexa, dtype=torch.int16, numel=n)
    for name, t in (("u1", u1), ("u2", u2), ("u3", u3), ("vel", vpad)):
        device_tensor(name, t, numel=n)
    wd = wavelet_on(w, frd.device, nt)
    rc = _lib.lib().tw_td2d_exa_forward(
        _lib.ptr(frd), _lib.ptr(exa), _lib.ptr(iexa), _lib.ptr(u1), _lib.ptr(u2), _lib.ptr(u3),
        _lib.ptr(vpad), _lib.ptr(wd), int(order), int(dimx), int(dimy), int(nt),
        float(h), float(dt), float(sx), float(sy), float(ry), _lib.stream_of(frd))
    _lib.check(rc, "td2d_exa_cuda.forward")


def backward(exa, iexa, u1, u2, u3, vpad, resid, order, dimx, dimy, nt, h, dt, ry):
    """Reference: fdm_backward, wrapper_exa_cuda.cpp:47-83 -> cuda_fdm_backward,
    propagator_exa_cuda.cu:87-159 (residual indexed with the true nx)."""
    check_common(order, dimx, dimy, nt)
    n = dimx * dimy
    nx = dimx - order
    device_tensor("exa", exa, numel=n)
    device_tensor("iexa", iexa, dtype=torch.int16, numel=n)
    for name, t in (("u1", u1), ("u2", u2), ("u3", u3), ("vel", vpad)):
        device_tensor(name, t, numel=n)
    resid = device_tensor("resid", resid.contiguous() if isinstance(resid, torch.Tensor) else resid,
                          numel=nt * nx)
    grad = torch.zeros_like(vpad)
    rc = _lib.lib().tw_td2d_exa_backward(
        _lib.ptr(grad), _lib.ptr(exa), _lib.ptr(iexa), _lib.ptr(u1), _lib.ptr(u2), _lib.ptr(u3),
        _lib.ptr(vpad), _lib.ptr(resid), int(order), int(dimx), int(dimy), int(nt),
        float(h), float(dt), float(ry), _lib.stream_of(grad))
    _lib.check(rc, "td2d_exa_cuda.backward")
    return grad
