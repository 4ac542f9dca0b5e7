"""Argument checks shared by the td2d extension modules.

The reference's pybind wrappers check contiguity (and CUDA residency) of
every tensor and let ``data_ptr<float>()`` reject other dtypes, raising
RuntimeError (wrapper_cuda.cpp:40-42, wrapper_cpu.cpp:40-41).  These checks
keep that behaviour and add the size checks the reference lacks, since an
undersized buffer would otherwise be written out of bounds on the device.
"""
import torch


def device_tensor(name, t, dtype=torch.float32, numel=None):
    if not isinstance(t, torch.Tensor):
        raise RuntimeError("%s must be a tensor" % name)
    if t.device.type != "cuda":
        raise RuntimeError("%s must be a CUDA (HIP) tensor" % name)
    if not t.is_contiguous():
        raise RuntimeError("%smust be contiguous" % name)
    if t.dtype != dtype:
        raise RuntimeError("%s: expected scalar type %s but found %s" % (name, dtype, t.dtype))
    if numel is not None and t.numel() < numel:
        raise RuntimeError("%s has %d elements, needs %d" % (name, t.numel(), numel))
    return t


def wavelet_on(w, device, nt):
    """The reference keeps the wavelet on the host (TimeOperator.py:28) and
    reads w[it] there; here it is staged to the device once per call."""
    if not isinstance(w, torch.Tensor):
        raise RuntimeError("w must be a tensor")
    if w.dtype != torch.float32:
        raise RuntimeError("w: expected scalar type Float but found %s" % w.dtype)
    if not w.is_contiguous():
        raise RuntimeError("wmust be contiguous")
    if w.numel() < nt:
        raise RuntimeError("w has %d samples, needs nt=%d" % (w.numel(), nt))
    return w.to(device, non_blocking=False) if w.device != device else w


def check_common(order, dimx, dimy, nt):
    if order not in (2, 4, 8):
        raise RuntimeError("Wrong order: use 2, 4, and 8 (got %r)" % (order,))
    if dimx <= order or dimy <= order or nt <= 0:
        raise RuntimeError("invalid grid dimx=%d dimy=%d nt=%d for order %d" % (dimx, dimy, nt, order))
