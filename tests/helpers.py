"""Shared helpers of the test suite (seeded small cases, error measures)."""
import numpy as np


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    d = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (d if d > 0 else 1.0))


def small_case(order, nx=37, ny=23, nt=200, seed=0, h=10.0, dt=1e-3, isx=None, isy=2, iry=2):
    """A seeded random-velocity case on an (nx, ny) model.  Positions are
    (i + 0.5)*h so the float truncation lands on cell i."""
    rng = np.random.default_rng(seed)
    vel = (1500.0 + 2500.0 * rng.random((nx, ny))).astype(np.float32)
    t = np.arange(nt) * dt
    f0 = 15.0
    w = ((1 - 2 * (np.pi * f0 * (t - 0.06)) ** 2) * np.exp(-(np.pi * f0 * (t - 0.06)) ** 2)).astype(np.float32)
    isx = nx // 2 if isx is None else isx
    return dict(vel=vel, w=w, order=order, nx=nx, ny=ny, dimx=nx + order, dimy=ny + order,
                nt=nt, h=h, dt=dt, sx=(isx + 0.5) * h, sy=(isy + 0.5) * h, ry=(iry + 0.5) * h)


def record_drift(case, **errors):
    """Record the relative L2 errors of one result (e.g. frd=..., grad=...
    against the FP64 run of the same recursion).  When TW_DRIFT_JSON names a
    file, the values are merged into it under `case`, so a test run leaves a
    per-result record; they are printed either way."""
    import json
    import os
    print("fp64 drift %s: %s" % (case, ", ".join("%s %.3e" % kv for kv in errors.items())), flush=True)
    path = os.environ.get("TW_DRIFT_JSON")
    if not path:
        return
    try:
        with open(path) as f:
            data = json.load(f)
    except (OSError, ValueError):
        data = {}
    data[case] = {k: float(v) for k, v in errors.items()}
    with open(path, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
