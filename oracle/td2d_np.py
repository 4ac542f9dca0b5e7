"""TEST INFRASTRUCTURE ONLY: brute-force FP64 numpy restatement of one time
step of the propagator, for tiny grids.

Written independently of td2d_oracle.c, on (dimy, dimx) arrays with explicit
shifted slices, so that a wrong index, sign or coefficient in either one
shows up as a disagreement.  Follows the reference's step definition
(torchwi/propagator/td2d/td2d_cpu.cpp:6-146; cross-checked against the
readable torchwi/propagator/td2d_base.py:5-89, which indexes (x, y)).
"""
from fractions import Fraction

import numpy as np


def lap_weights(order):
    """Central second-difference weights derived from the Taylor conditions
    sum_k w_k k^(2j) = [j == 1] (j = 0..ne), solved exactly in rationals --
    independent of the decimal literals the reference writes."""
    ne = order // 2
    # unknowns c_1..c_ne: sum_k c_k * 2 * k^(2j) / (2j)! = delta_{j,1}, j=1..ne
    A = [[Fraction(2 * k ** (2 * j)) for k in range(1, ne + 1)] for j in range(1, ne + 1)]
    b = [Fraction(0) for _ in range(ne)]
    b[0] = Fraction(2)  # sum_k 2 c_k k^2 = 2 (second derivative * 2!/2)
    # Gaussian elimination over rationals
    n = ne
    M = [row[:] + [b[i]] for i, row in enumerate(A)]
    for i in range(n):
        piv = next(r for r in range(i, n) if M[r][i] != 0)
        M[i], M[piv] = M[piv], M[i]
        for r in range(n):
            if r != i and M[r][i] != 0:
                f = M[r][i] / M[i][i]
                M[r] = [x - f * y for x, y in zip(M[r], M[i])]
    c = [M[i][n] / M[i][i] for i in range(n)]
    c0 = -2 * sum(c)
    return [float(c0)] + [float(x) for x in c]


CM = np.cos(np.pi / 6) * np.cos(np.pi / 12)
CP = np.cos(np.pi / 6) + np.cos(np.pi / 12)


def step(u1, u2, v, order, h, dt):
    """u3 from (u1, u2) on (dimy, dimx) FP64 arrays; top ne rows left zero."""
    ne = order // 2
    W = lap_weights(order)
    dimy, dimx = u2.shape
    u3 = np.zeros_like(u2)
    I = (slice(ne, dimy - ne), slice(ne, dimx - ne))
    lap = 2 * W[0] * u2[I]
    for k in range(1, ne + 1):
        lap = lap + W[k] * (u2[ne - k:dimy - ne - k, ne:dimx - ne] + u2[ne + k:dimy - ne + k, ne:dimx - ne]
                            + u2[ne:dimy - ne, ne - k:dimx - ne - k] + u2[ne:dimy - ne, ne + k:dimx - ne + k])
    u3[I] = 2 * u2[I] - u1[I] + v[I] ** 2 * dt ** 2 / h ** 2 * lap

    def keys(iy, ix, dy, dx):
        um0, uo0 = u1[iy, ix], u2[iy, ix]
        um1, uo1 = u1[iy + dy, ix + dx], u2[iy + dy, ix + dx]
        uo2 = u2[iy + 2 * dy, ix + 2 * dx]
        vv = v[iy, ix]
        uxx = (uo0 - 2 * uo1 + uo2) / h ** 2
        uxt = ((uo0 - um0) - (uo1 - um1)) / (h * dt)
        return -vv ** 2 * dt ** 2 / CM * (uxx + CP / vv * uxt) + 2 * uo0 - um0

    for iy in range(ne, dimy - ne):
        for ix in range(dimx - ne, dimx):
            u3[iy, ix] = keys(iy, ix, 0, -1)
        for ix in range(0, ne):
            u3[iy, ix] = keys(iy, ix, 0, +1)
    for iy in range(dimy - ne, dimy):
        for ix in range(dimx):
            u3[iy, ix] = keys(iy, ix, -1, 0)
    return u3
