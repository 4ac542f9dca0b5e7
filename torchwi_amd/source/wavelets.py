"""Source wavelets.  Same functions and results as torchwi/source/wavelets.py.

``ricker(freq, dt, nt)`` is the time-shifted Ricker used by the reference's
examples: with vm = freq/2/1.08 its peak frequency is freq/2.16, and the
shift t0 = icut*dt puts the first sample where |a| < 1e-3*a[0] at t = 0.
"""
import numpy as np


def _ricker_value(vm2, t):
    # scalar np.exp on a Python float, as the reference evaluates it; the
    # vectorised ufunc can differ in the last bit for tiny tail values
    arg = np.pi ** 2 * t ** 2
    return (1.0 - 2.0 * vm2 * arg) * np.exp(-vm2 * arg)


def ricker(freq, dt, nt, dtype=np.float64):
    """Reference: torchwi/source/wavelets.py:3-23 (FP64, cast by the caller)."""
    vm = freq / 2.0 / 1.08
    vm2 = vm ** 2
    first = [_ricker_value(vm2, i * dt) for i in range(nt)]
    icut = next((i for i in range(nt) if np.abs(first[i]) < first[0] * 1.0e-3), 0)
    t0 = icut * dt
    return np.array([_ricker_value(vm2, t0 - i * dt) for i in range(nt)], dtype=dtype)


def fdgaus(cutoff, dt, nt, dtype=np.float64):
    """First derivative of a Gaussian, normalised to max 1.
    Reference: torchwi/source/wavelets.py:26-52."""
    a = np.pi * (5.0 * cutoff / 8.0) ** 2
    amp = np.sqrt(a / np.pi)
    t = np.arange(nt) * dt
    w = amp * np.exp(np.maximum(-a * t ** 2, -32.0))
    t0 = 0.0
    below = np.nonzero(w < 0.001 * w[0])[0]
    if len(below):
        t0 = (int(below[0]) - 1) * dt
    tt = t - t0
    w = -2.0 * np.sqrt(a) * a * tt * np.exp(np.maximum(-a * tt ** 2, -32.0)) / np.sqrt(np.pi)
    w = w / w.max()
    return w.astype(dtype)
