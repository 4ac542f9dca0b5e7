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
    for cls in (torchwi.Time2dBatch, torchwi.Time2dForwardBatch):
        with pytest.raises(NotImplementedError):
            cls(20, 10, 10.0, w, 1e-3, 8, "cpu")


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


def header_struct_fields(name):
    """Field names of `typedef struct name {...} name;` in include/*.h, in order."""
    inc = os.path.join(ROOT, "include")
    for f in sorted(os.listdir(inc)):
        text = re.sub(r"/\*.*?\*/", "", open(os.path.join(inc, f)).read(), flags=re.S)
        m = re.search(r"typedef struct %s \{(.*?)\} %s;" % (name, name), text, flags=re.S)
        if m:
            fields = []
            for decl in m.group(1).split(";"):
                decl = decl.strip()
                if not decl:
                    continue
                parts = decl.split(None, 1) if not decl.startswith("const ") else ["const"] + decl[6:].split(None, 1)
                for v in parts[-1].split(","):
                    fields.append(v.strip().lstrip("*").strip())
            return fields
    raise AssertionError("struct %s not found" % name)


@pytest.mark.parametrize("name", ["TwPlanSpec", "TwShotSpec"])
def test_ctypes_structs_follow_the_header(name):
    """The ctypes mirrors of the C-ABI structs have the header's fields in the
    header's order (a binding that drifts would pass garbage silently)."""
    from torchwi_amd import _lib
    want = header_struct_fields(name)
    got = [f for f, _ in getattr(_lib, name)._fields_]
    assert got == want


def test_grad_group_is_refused_on_the_host():
    """grad_group > 1 is refused by plan validation (no device call)."""
    from torchwi_amd import _lib
    L = _lib.lib()
    spec = _lib.TwPlanSpec(order=8, nt=10, h=10.0, dt=1e-3, ckpt_every=0, grad_group=2)
    shots = (_lib.TwShotSpec * 2)()
    for q in shots:
        q.dimx, q.dimy, q.ld = 40, 30, 64
        q.sx, q.sy, q.ry = 105.0, 25.0, 25.0
        q.rx0, q.nrec = 0, 32
        for p in ("vel", "vm", "vc3", "w", "u", "lam", "frd", "resid", "virt", "grad"):
            setattr(q, p, 256)
    assert L.tw_plan_workspace_bytes(ctypes.byref(spec), shots, 2) == 0
    assert "grad_group" in _lib.last_error()
    spec.grad_group = 1
    assert L.tw_plan_workspace_bytes(ctypes.byref(spec), shots, 2) > 0


def test_static_plan_limits_are_refused_before_any_device_call():
    """tw_plan_create_static refuses, with a clear error and without touching
    a device, a plan larger than the static workspace (tw_plan_workspace_bytes
    tells the size) and graph capture, which may allocate."""
    from torchwi_amd import _lib
    L = _lib.lib()
    h = 10.0
    dimx, dimy = 3008, 1008
    ld = 3008

    def shots(n):
        arr = (_lib.TwShotSpec * n)()
        for q in arr:
            q.dimx, q.dimy, q.ld = dimx, dimy, ld
            q.sx, q.sy, q.ry = 100.5 * h, 2.5 * h, 2.5 * h
            q.rx0, q.nrec = 0, dimx - 8
            for p in ("vel", "vm", "vc3", "w", "u", "lam", "frd", "resid", "virt", "grad"):
                setattr(q, p, 4096)
        return arr

    spec = _lib.TwPlanSpec(order=8, nt=50, h=h, dt=1e-3, ckpt_every=0, graph_steps=0, time_block=1)
    big = shots(256)   # 64 x 16 tiles: 256 x 47 x 63 of them, 6 MB of tile list
    need = L.tw_plan_workspace_bytes(ctypes.byref(spec), big, 256)
    assert need > 4194304
    plan = ctypes.c_void_p()
    rc = L.tw_plan_create_static(ctypes.byref(spec), big, 256, None, ctypes.byref(plan))
    assert rc == -1 and not plan.value
    msg = _lib.last_error()
    assert "static workspace" in msg and str(need) in msg and "4194304" in msg, msg
    spec.graph_steps = 100
    assert L.tw_plan_create_static(ctypes.byref(spec), shots(1), 1, None, ctypes.byref(plan)) == -1
    assert "graph_steps = 0" in _lib.last_error()
