"""ctypes binding of the C ABI in ``include/torchwi_hip.h``.

The library is ``torchwi_amd/lib/libtorchwi_hip.so`` (built by
``python -m torchwi_amd.build``).  There is no fallback: if the library is
missing or cannot be loaded, every entry point raises.
"""
import ctypes
import os
import threading

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libtorchwi_hip.so")

_lock = threading.Lock()
_lib = None

_vp = ctypes.c_void_p
_ci = ctypes.c_int
_cf = ctypes.c_float

# Exported symbol -> (restype, argtypes).  Kept in the order of the header.
SIGNATURES = {
    "tw_last_error": (ctypes.c_char_p, []),
    "tw_version": (ctypes.c_char_p, []),
    "tw_td2d_forward": (_ci, [_vp] * 6 + [_vp, _ci, _ci, _ci, _ci, _cf, _cf, _cf, _cf, _cf, _vp]),
    "tw_td2d_forward_only": (_ci, [_vp] * 5 + [_vp, _ci, _ci, _ci, _ci, _cf, _cf, _cf, _cf, _cf, _vp]),
    "tw_td2d_backward": (_ci, [_vp] * 7 + [_ci, _ci, _ci, _ci, _cf, _cf, _cf, _vp]),
    "tw_td2d_exa_forward": (_ci, [_vp] * 7 + [_vp, _ci, _ci, _ci, _ci, _cf, _cf, _cf, _cf, _cf, _vp]),
    "tw_td2d_exa_backward": (_ci, [_vp] * 8 + [_ci, _ci, _ci, _ci, _cf, _cf, _cf, _vp]),
    "tw_plan_workspace_bytes": (ctypes.c_size_t, [_vp, _vp, _ci]),
    "tw_plan_create": (_ci, [_vp, _vp, _ci, _vp, ctypes.c_size_t, _vp, ctypes.POINTER(_vp)]),
    "tw_plan_forward": (_ci, [_vp, _vp]),
    "tw_plan_backward": (_ci, [_vp, _ci, _vp]),
    "tw_plan_destroy": (None, [_vp]),
    "tw_plan_shot_groups": (_ci, [_vp]),
    "tw_plan_create_static": (_ci, [_vp, _vp, _ci, _vp, ctypes.POINTER(_vp)]),
    "tw_plan_set_timing": (_ci, [_vp, _ci]),
    "tw_plan_timing": (_ci, [_vp, _ci, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_longlong)]),
    "tw_stream_copy": (_ci, [_vp, _vp, ctypes.c_size_t, _vp]),
}

# Step-kernel kinds of tw_plan_timing (TW_KIND_* in the header).
KINDS = ("fwd", "fwd_virt", "recomp", "adj")


class TwPlanSpec(ctypes.Structure):
    _fields_ = [("order", _ci), ("nt", _ci), ("h", _cf), ("dt", _cf), ("ckpt_every", _ci),
                ("tile_rows", _ci), ("graph_steps", _ci), ("time_block", _ci), ("forward_only", _ci),
                ("tile_cols", _ci), ("grad_group", _ci), ("streams", _ci)]


class TwShotSpec(ctypes.Structure):
    _fields_ = [
        ("dimx", _ci), ("dimy", _ci), ("ld", _ci),
        ("sx", _cf), ("sy", _cf), ("ry", _cf),
        ("rx0", _ci), ("nrec", _ci),
        ("vel", _vp), ("vm", _vp), ("vc3", _vp), ("w", _vp), ("u", _vp), ("lam", _vp), ("frd", _vp),
        ("resid", _vp), ("virt", _vp), ("ckpt", _vp), ("grad", _vp),
    ]


def lib():
    """Load the library (once).  Raises RuntimeError when it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    "torchwi_amd: HIP library %s is not built; run "
                    "`python -m torchwi_amd.build` (no CPU fallback exists)" % LIB_PATH)
            L = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                f = getattr(L, name)
                f.restype = res
                f.argtypes = args
            _lib = L
    return _lib


def last_error():
    return lib().tw_last_error().decode(errors="replace")


def check(rc, what):
    if rc != 0:
        raise RuntimeError("torchwi_amd %s failed (status %d): %s" % (what, rc, last_error()))


def stream_of(t):
    """The raw hipStream_t of torch's current stream on t's device."""
    return ctypes.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None
