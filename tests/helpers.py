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
