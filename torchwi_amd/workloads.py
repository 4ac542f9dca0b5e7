"""Synthetic inputs of the benchmark workloads (seeded, FP32, SI units).

Models are returned in the reference's user layout ``vel[ix, iz]`` of shape
(nx, nz), depth fastest (TimeOperator.py:20-22); z is the reference's y.
Source/receiver positions are returned as physical coordinates (i + 0.5)*h,
which the propagator truncates to cell i exactly.
"""
import numpy as np
from scipy import ndimage

from .source.wavelets import ricker


def cell_pos(i, h):
    return (np.asarray(i, dtype=np.float64) + 0.5) * h


def _smooth_noise(rng, shape, sigma):
    n = ndimage.gaussian_filter(rng.standard_normal(shape), sigma, mode="wrap")
    return n / np.abs(n).max()


def three_layer(nz, nx, seed=0, sigma=3.0):
    """3 layers 1500/2500/3500 m/s, interfaces at depths z1~U[20,45],
    z2~U[55,85] cells (scaled to nz/100), Gaussian-smoothed (nz, nx)."""
    rng = np.random.default_rng(seed)
    z1 = rng.uniform(20, 45) * nz / 100.0
    z2 = rng.uniform(55, 85) * nz / 100.0
    z = np.arange(nz)[:, None] * np.ones((1, nx))
    v = np.where(z < z1, 1500.0, np.where(z < z2, 2500.0, 3500.0))
    return ndimage.gaussian_filter(v, sigma, mode="nearest")


def dipping_layers(nz, nx, nlayers=8, vmin=1500.0, vmax=4500.0, dip_deg=5.0, seed=1,
                   perturb=0.05, corr=20.0):
    """nlayers dipping layers, velocities linear vmin..vmax, a shared dip with
    random offsets, times (1 + p) with p Gaussian-correlated noise scaled to
    max|p| = perturb.  Returns (nz, nx)."""
    rng = np.random.default_rng(seed)
    vels = np.linspace(vmin, vmax, nlayers)
    base = np.linspace(0, nz, nlayers + 1)[1:-1]
    offs = rng.uniform(-0.3, 0.3, nlayers - 1) * nz / nlayers
    slope = np.tan(np.deg2rad(dip_deg))
    x = np.arange(nx)[None, :]
    z = np.arange(nz)[:, None]
    v = np.full((nz, nx), vels[0])
    for k in range(nlayers - 1):
        iface = base[k] + offs[k] + slope * (x - nx / 2)
        v = np.where(z >= iface, vels[k + 1], v)
    if perturb:
        v = v * (1.0 + perturb * _smooth_noise(rng, (nz, nx), corr))
    return v


def smooth_layered(nz, nx, vmin=1500.0, vmax=4000.0, seed=3, nlayers=10, sigma=8.0):
    """Smooth random layered model: random layer thicknesses, velocities
    increasing with depth plus jitter, Gaussian-smoothed."""
    rng = np.random.default_rng(seed)
    cuts = np.sort(rng.uniform(0, nz, nlayers - 1))
    vels = np.sort(rng.uniform(vmin, vmax, nlayers))
    z = np.arange(nz)[:, None] * np.ones((1, nx))
    idx = np.searchsorted(cuts, z)
    v = vels[idx]
    return np.clip(ndimage.gaussian_filter(v, sigma, mode="nearest"), vmin, vmax)


def to_user(v_zx):
    """(nz, nx) -> the reference's vel layout (nx, nz), float32."""
    return np.ascontiguousarray(np.asarray(v_zx, dtype=np.float32).T)


def three_layer_100x200(order=8, pad=20, nt=1000):
    """Plain reference-path case: 100x200 model, h=10 m, dt=1 ms, nt=1000,
    Keys ABC behind a pad-cell edge-replicated border (left, right, bottom;
    the top is the free surface), source (z=2, x=100), receivers at z=2 on
    every cell, ricker(10); observed data from a +5% Gaussian blob."""
    nz, nx, h, dt = 100, 200, 10.0, 1e-3
    v = three_layer(nz, nx, seed=0)
    zz, xx = np.mgrid[0:nz, 0:nx]
    blob = np.exp(-((zz - 50) ** 2 + (xx - 100) ** 2) / (2 * 10.0 ** 2))
    v_true = v * (1.0 + 0.05 * blob)
    padw = ((0, pad), (pad, pad))
    vp = np.pad(v, padw, mode="edge")
    vp_true = np.pad(v_true, padw, mode="edge")
    return dict(
        nx=nx + 2 * pad, ny=nz + pad, h=h, dt=dt, order=order, nt=nt, pad=pad,
        vel=to_user(vp), vel_true=to_user(vp_true),
        w=ricker(10.0, dt, nt).astype(np.float32),
        sx=[float(cell_pos(100 + pad, h))], sy=[float(cell_pos(2, h))], ry=[float(cell_pos(2, h))],
        crop=(pad, pad + nx))


def dipping_384x1024(order=8, nshots=32, nt=4000):
    """Plain HIP form: 384x1024 model, h=12.5 m, dt=1 ms, nt=4000, 8 dipping
    layers 1500-4500 m/s with +-5% correlated perturbation (corr. 20 cells),
    32 ricker(8) shots evenly spaced at z=2, 1024 receivers at z=2."""
    nz, nx, h, dt = 384, 1024, 12.5, 1e-3
    v = dipping_layers(nz, nx, seed=1)
    xs = (np.arange(nshots) + 0.5) * nx / nshots
    return dict(
        nx=nx, ny=nz, h=h, dt=dt, order=order, nt=nt,
        vel=to_user(v), w=ricker(8.0, dt, nt).astype(np.float32),
        sx=[float(cell_pos(int(x), h)) for x in xs], sy=[float(cell_pos(2, h))] * nshots,
        ry=[float(cell_pos(2, h))] * nshots)


def dipping_1000x3000(order=8, nshots=64, nt=8000):
    """Differentiation at scale: 1000x3000, h=10 m, dt=0.8 ms, nt=8000,
    64 ricker(6) shots, 3000 receivers, dipping-layer family."""
    nz, nx, h, dt = 1000, 3000, 10.0, 8e-4
    v = dipping_layers(nz, nx, seed=2)
    xs = (np.arange(nshots) + 0.5) * nx / nshots
    return dict(
        nx=nx, ny=nz, h=h, dt=dt, order=order, nt=nt,
        vel=to_user(v), w=ricker(6.0, dt, nt).astype(np.float32),
        sx=[float(cell_pos(int(x), h)) for x in xs], sy=[float(cell_pos(2, h))] * nshots,
        ry=[float(cell_pos(2, h))] * nshots)


def uneven_600x4000(order=8, nshots=256, nt=6000, seed=4, margin=40):
    """Batched uneven shots: 600x4000 model, h=10 m, dt=1 ms, nt=6000, smooth
    random layered 1500-4000 m/s; 256 shots with receiver windows of
    U{400..2000} cells (seeded), each shot's sub-grid its window +-40 cells
    (clipped to the model), the shot at the window centre, receivers on the
    window's columns at z=2."""
    nz, nx, h, dt = 600, 4000, 10.0, 1e-3
    v = smooth_layered(nz, nx, seed=3)
    rng = np.random.default_rng(seed)
    widths = rng.integers(400, 2001, nshots)
    centres = (np.arange(nshots) + 0.5) * nx / nshots
    wins = []
    for k in range(nshots):
        wdt = int(widths[k])
        c = int(centres[k])
        r0 = max(0, min(nx - wdt, c - wdt // 2))
        x0 = max(0, r0 - margin)
        x1 = min(nx, r0 + wdt + margin)
        src = min(max(c, r0), r0 + wdt - 1)
        wins.append(dict(x0=x0, x1=x1, sx=float(cell_pos(src - x0, h)), sy=float(cell_pos(2, h)),
                         ry=float(cell_pos(2, h)), rx0=r0 - x0, nrec=wdt))
    return dict(nx=nx, ny=nz, h=h, dt=dt, order=order, nt=nt, vel=to_user(v),
                w=ricker(10.0, dt, nt).astype(np.float32), windows=wins)
