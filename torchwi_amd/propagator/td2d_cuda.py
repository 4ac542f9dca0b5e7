"""Full-storage forward / adjoint gradient on the GPU.

Same functions and argument lists as the reference's ``td2d_cuda`` /
``td2d_cpu`` modules (torchwi/propagator/td2d/wrapper_cuda.cpp,
wrapper_cpu.cpp:44-118); results follow the reference CPU path.
"""
import torch

from .. import _lib
from ._td2d_common import check_common, device_tensor, wavelet_on


def forward(frd, virt, u1, u2, u3, vpad, w, order, dimx, dimy, nt, h, dt, sx, sy, ry):
    """Shot gather frd (nt*nx) and virtual sources virt (nt*dimx*dimy).

    Reference: fdm_forward, wrapper_cpu.cpp:44-78 -> c_fdm_forward,
    propagator_cpu.cpp:5-79."""
    check_common(order, dimx, dimy, nt)
    n = dimx * dimy
    nx = dimx - order
    device_tensor("forward", frd, numel=nt * nx)
    device_tensor("source_wavefield", virt, numel=nt * n)
    for name, t in (("u1", u1), ("u2", u2), ("u3", u3), ("vel", vpad)):
        device_tensor(name, t, numel=n)
    wd = wavelet_on(w, frd.device, nt)
    rc = _lib.lib().tw_td2d_forward(
        _lib.ptr(frd), _lib.ptr(virt), _lib.ptr(u1), _lib.ptr(u2), _lib.ptr(u3),
        _lib.ptr(vpad), _lib.ptr(wd), int(order), int(dimx), int(dimy), int(nt),
        float(h), float(dt), float(sx), float(sy), float(ry), _lib.stream_of(frd))
    _lib.check(rc, "td2d_cuda.forward")


def backward(virt, u1, u2, u3, vpad, resid, order, dimx, dimy, nt, h, dt, ry):
    """Gradient on the padded grid (dimx*dimy), freshly zeroed then
    accumulated over all steps.

    Reference: fdm_backward, wrapper_cpu.cpp:80-112 -> c_fdm_backward,
    propagator_cpu.cpp:82-152."""
    check_common(order, dimx, dimy, nt)
    n = dimx * dimy
    nx = dimx - order
    device_tensor("source_wavefield", virt, numel=nt * n)
    for name, t in (("u1", u1), ("u2", u2), ("u3", u3), ("vel", vpad)):
        device_tensor(name, t, numel=n)
    resid = device_tensor("resid", resid.contiguous() if isinstance(resid, torch.Tensor) else resid,
                          numel=nt * nx)
    grad = torch.zeros_like(vpad)
    rc = _lib.lib().tw_td2d_backward(
        _lib.ptr(grad), _lib.ptr(virt), _lib.ptr(u1), _lib.ptr(u2), _lib.ptr(u3),
        _lib.ptr(vpad), _lib.ptr(resid), int(order), int(dimx), int(dimy), int(nt),
        float(h), float(dt), float(ry), _lib.stream_of(grad))
    _lib.check(rc, "td2d_cuda.backward")
    return grad
