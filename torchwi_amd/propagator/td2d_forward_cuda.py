"""Forward modelling without storage on the GPU.

Same functions and argument lists as the reference's ``td2d_forward_cuda`` /
``td2d_forward_cpu`` modules (wrapper_forward_cpu.cpp:27-70)."""
import torch

from .. import _lib
from ._td2d_common import check_common, device_tensor, wavelet_on


def forward(frd, u1, u2, u3, vpad, w, order, dimx, dimy, nt, h, dt, sx, sy, ry):
    """Reference: fdm_forward, wrapper_forward_cpu.cpp:27-58 ->
    c_fdm_forward_only, propagator_forward_cpu.cpp:5-71."""
    check_common(order, dimx, dimy, nt)
    n = dimx * dimy
    nx = dimx - order
    device_tensor("forward", frd, numel=nt * nx)
    for name, t in (("u1", u1), ("u2", u2), ("u3", u3), ("vel", vpad)):
        device_tensor(name, t, numel=n)
    wd = wavelet_on(w, frd.device, nt)
    rc = _lib.lib().tw_td2d_forward_only(
        _lib.ptr(frd), _lib.ptr(u1), _lib.ptr(u2), _lib.ptr(u3), _lib.ptr(vpad), _lib.ptr(wd),
        int(order), int(dimx), int(dimy), int(nt), float(h), float(dt),
        float(sx), float(sy), float(ry), _lib.stream_of(frd))
    _lib.check(rc, "td2d_forward_cuda.forward")


def backward():
    """The reference's stub returns zeros(1) (wrapper_forward_cpu.cpp:60-64)."""
    return torch.zeros(1)
