"""Property-based tests (hypothesis) on the CPU: random geometries and inputs.

* the oracle against the reference's own compiled CPU path (oracle/_ref,
  built from the reference's sources where they exist): gathers, virtual
  sources and gradients bit for bit on random grids, orders, positions and
  step counts;
* the batched plan's host-side validation (no device call): every geometry
  inside the rules is accepted with a positive workspace, sources, receivers
  and windows outside them are refused with a message naming the problem;
* RSF files: write/read round trips of random arrays and axis headers.
"""
import ctypes
import os

import numpy as np
import pytest
from hypothesis import HealthCheck, given, settings
from hypothesis import strategies as st

from oracle import oracle as O

SETTINGS = settings(max_examples=40, deadline=None, suppress_health_check=[HealthCheck.too_slow])


@st.composite
def cases(draw):
    order = draw(st.sampled_from([2, 4, 8]))
    nx = draw(st.integers(6, 28))
    ny = draw(st.integers(6, 18))
    nt = draw(st.integers(1, 24))
    isx = draw(st.integers(0, nx - 1))
    isy = draw(st.integers(0, ny - 1))
    iry = draw(st.integers(0, ny - 1))
    seed = draw(st.integers(0, 2**31 - 1))
    return order, nx, ny, nt, isx, isy, iry, seed


@pytest.mark.skipif(not O.ref_available(), reason="the reference's CPU path is not built here")
@SETTINGS
@given(cases())
def test_oracle_equals_reference_on_random_cases(c):
    order, nx, ny, nt, isx, isy, iry, seed = c
    rng = np.random.default_rng(seed)
    h, dt = 10.0, 1e-3
    vel = (1500.0 + 2500.0 * rng.random((nx, ny))).astype(np.float32)
    w = rng.standard_normal(nt).astype(np.float32)
    vpad = O.pad_vel(vel, order)
    geo = (order, nx + order, ny + order, h, dt)
    pos = ((isx + 0.5) * h, (isy + 0.5) * h, (iry + 0.5) * h)
    frd_o, virt_o = O.forward(vpad, w, *geo, *pos)
    frd_r, virt_r = O.forward(vpad, w, *geo, *pos, impl="ref")
    assert np.array_equal(frd_o, frd_r) and np.array_equal(virt_o, virt_r)
    resid = rng.standard_normal((nt, nx)).astype(np.float32)
    g_o = O.backward(virt_o, vpad, resid, *geo, pos[2])
    g_r = O.backward(virt_r, vpad, resid, *geo, pos[2], impl="ref")
    assert np.array_equal(g_o, g_r)


def _shot(L, dimx, dimy, ld, sx, sy, ry, rx0, nrec):
    q = L.TwShotSpec()
    q.dimx, q.dimy, q.ld = dimx, dimy, ld
    q.sx, q.sy, q.ry = sx, sy, ry
    q.rx0, q.nrec = rx0, nrec
    for p in ("vel", "vm", "vc3", "w", "u", "lam", "frd", "resid", "virt", "grad"):
        setattr(q, p, 4096)
    return q


@SETTINGS
@given(order=st.sampled_from([2, 4, 8]), nx=st.integers(8, 3000), ny=st.integers(8, 1500),
       fx=st.floats(0, 1, exclude_max=True), fy=st.floats(0, 1, exclude_max=True),
       fr=st.floats(0, 1, exclude_max=True), a=st.floats(0, 1), b=st.floats(0, 1),
       tile=st.sampled_from([(0, 0), (64, 32), (64, 64), (128, 8), (128, 16)]))
def test_plan_accepts_every_valid_geometry(order, nx, ny, fx, fy, fr, a, b, tile):
    from torchwi_amd import _lib
    L = _lib.lib()
    h = 10.0
    isx, isy, iry = int(fx * nx), int(fy * ny), int(fr * ny)
    rx0 = int(a * (nx - 1))
    nrec = 1 + int(b * (nx - rx0 - 1))
    dimx, dimy = nx + order, ny + order
    ld = (dimx + 31) // 32 * 32
    spec = _lib.TwPlanSpec(order=order, nt=50, h=h, dt=1e-3, ckpt_every=0, tile_cols=tile[0], tile_rows=tile[1],
                           graph_steps=10)
    shots = (_lib.TwShotSpec * 2)(_shot(_lib, dimx, dimy, ld, (isx + 0.5) * h, (isy + 0.5) * h, (iry + 0.5) * h,
                                        rx0, nrec),
                                  _shot(_lib, dimx, dimy, ld, 0.5 * h, 0.5 * h, 0.5 * h, 0, nx))
    assert L.tw_plan_workspace_bytes(ctypes.byref(spec), shots, 2) > 0, _lib.last_error()


@SETTINGS
@given(order=st.sampled_from([2, 4, 8]), nx=st.integers(8, 400), ny=st.integers(8, 200),
       what=st.sampled_from(["src_x", "src_y", "rec_y", "window", "ld"]), over=st.integers(0, 50))
def test_plan_refuses_geometry_outside_the_rules(order, nx, ny, what, over):
    from torchwi_amd import _lib
    L = _lib.lib()
    h = 10.0
    dimx, dimy = nx + order, ny + order
    ld = (dimx + 31) // 32 * 32
    sx, sy, ry, rx0, nrec = 0.5 * h, 0.5 * h, 0.5 * h, 0, nx
    if what == "src_x":
        sx = (nx + over + 0.5) * h
    elif what == "src_y":
        sy = (ny + over + 0.5) * h
    elif what == "rec_y":
        ry = (ny + over + 0.5) * h
    elif what == "window":
        rx0, nrec = nx - 1 - over % nx, over % nx + 2
    else:
        ld = dimx - 1 - over % 3 if over % 2 else dimx + 1 + 4 * (over % 5)
    spec = _lib.TwPlanSpec(order=order, nt=50, h=h, dt=1e-3, ckpt_every=0)
    shots = (_lib.TwShotSpec * 1)(_shot(_lib, dimx, dimy, ld, sx, sy, ry, rx0, nrec))
    if what == "ld" and ld >= dimx and ld % 4 == 0:
        return   # that stride is valid
    assert L.tw_plan_workspace_bytes(ctypes.byref(spec), shots, 1) == 0
    msg = _lib.last_error()
    key = {"src_x": "source", "src_y": "source", "rec_y": "receiver", "window": "receiver window",
           "ld": "ld"}[what]
    assert key in msg, msg


@SETTINGS
@given(shape=st.lists(st.integers(1, 7), min_size=1, max_size=3),
       typ=st.sampled_from(["float", "int", "complex"]), form=st.sampled_from(["native", "ascii"]),
       o1=st.floats(-1e3, 1e3, allow_nan=False), d1=st.floats(1e-3, 1e3, allow_nan=False),
       seed=st.integers(0, 2**31 - 1))
def test_rsf_round_trip(tmp_path_factory, shape, typ, form, o1, d1, seed):
    from torchwi_amd.io import rsf
    if form == "ascii" and (typ == "complex" or len(shape) > 2):
        return   # the ASCII form holds real 1-D / 2-D tables (numpy savetxt/loadtxt)
    rng = np.random.default_rng(seed)
    dt = {"float": np.float32, "int": np.int32, "complex": np.complex64}[typ]
    a = (rng.standard_normal(shape) * 100).astype(dt)
    if typ == "complex":
        a = (a + 1j * rng.standard_normal(shape)).astype(dt)
    f = str(tmp_path_factory.mktemp("rsf") / "a.rsf")
    keys = {}
    for k, n in enumerate(reversed(shape)):   # n1 is the fastest axis
        keys["n%d" % (k + 1)] = n
    keys["o1"], keys["d1"] = o1, d1
    rsf.tofile(f, a, data_format="%s_%s" % (form, typ), **keys)
    got = rsf.fromfile(f, "data")
    assert got.dtype == dt
    if form == "ascii":
        assert np.allclose(got.reshape(a.shape), a, rtol=0, atol=0) or np.array_equal(got.reshape(a.shape), a)
    else:
        assert got.shape == a.shape and np.array_equal(got, a)
    n1, o1r, d1r = rsf.fromfile(f, "n1 o1 d1")
    assert (n1, o1r, d1r) == (shape[-1], float(o1), float(d1))
