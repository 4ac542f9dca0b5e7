"""CPU-side checks of the C ABI and the host-side surface (no GPU calls)."""
import ctypes
import os
import re

import pytest
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    names = []
    inc = os.path.join(ROOT, "include")
    for f in sorted(os.listdir(inc)):
        if f.endswith(".h"):
            text = open(os.path.join(inc, f)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names += re.findall(r"\b(tw_\w+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    from torchwi_amd import _lib
    L = _lib.lib()
    names = header_functions()
    assert "tw_td2d_forward" in names and "tw_plan_backward" in names
    for n in names:
        assert hasattr(L, n), n
        assert n in _lib.SIGNATURES, "no ctypes signature for %s" % n
    assert L.tw_version().startswith(b"torchwi_amd")


def test_library_is_built_for_gfx950():
    from torchwi_amd import _lib
    data = open(_lib.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_cpu_device_is_refused_not_emulated():
    import torchwi_amd as torchwi
    w = torch.zeros(10)
    with pytest.raises(NotImplementedError):
        torchwi.Time2d(20, 10, 10.0, w, 1e-3, 8, "cpu")
    with pytest.raises(NotImplementedError):
        torchwi.Time2dForward(20, 10, 10.0, w, 1e-3, 8, "cpu")


def test_get_operator_surface():
    from torchwi_amd.operator import TimeFunction
    with pytest.raises(NotImplementedError):
        TimeFunction.get_operator("cuda", prop="bogus")
    f = TimeFunction.get_operator("cuda", forward_only=True)
    assert callable(f)
    assert callable(TimeFunction.get_operator("cuda", exa=True))


def test_plan_validation_errors_without_gpu():
    """Plan validation happens on the host before any device call."""
    from torchwi_amd import _lib
    L = _lib.lib()
    spec = _lib.TwPlanSpec(order=6, nt=10, h=10.0, dt=1e-3, ckpt_every=0)
    shots = (_lib.TwShotSpec * 1)()
    assert L.tw_plan_workspace_bytes(ctypes.byref(spec), shots, 1) == 0
    assert "order" in _lib.last_error()
    spec.order = 8
    q = shots[0]
    q.dimx, q.dimy, q.ld = 40, 30, 40
    q.sx, q.sy, q.ry = 5000.0, 25.0, 25.0
    assert L.tw_plan_workspace_bytes(ctypes.byref(spec), shots, 1) == 0
    assert "source" in _lib.last_error()
