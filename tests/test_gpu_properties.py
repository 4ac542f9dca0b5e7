"""Property-based GPU parity (hypothesis, a fixed derandomized set of cases):
random grids, orders, step counts and positions through the drop-in
single-shot module and through batched plans, one step per launch and
temporally blocked, against the oracle bit for bit."""
import numpy as np
import pytest
import torch
from hypothesis import HealthCheck, given, settings
from hypothesis import strategies as st

from oracle import oracle as O

pytestmark = pytest.mark.gpu
DEV = "cuda"
SETTINGS = settings(max_examples=15, deadline=None, derandomize=True, database=None,
                    suppress_health_check=[HealthCheck.too_slow, HealthCheck.function_scoped_fixture])


def _vel_w(nx, ny, nt, seed):
    rng = np.random.default_rng(seed)
    vel = (1500.0 + 2500.0 * rng.random((nx, ny))).astype(np.float32)
    w = rng.standard_normal(nt).astype(np.float32)
    return rng, vel, w


@SETTINGS
@given(order=st.sampled_from([2, 4, 8]), nx=st.integers(8, 160), ny=st.integers(8, 100), nt=st.integers(1, 40),
       fx=st.floats(0, 1, exclude_max=True), fy=st.floats(0, 1, exclude_max=True),
       fr=st.floats(0, 1, exclude_max=True), seed=st.integers(0, 2**31 - 1))
def test_single_shot_random_cases_bitexact(order, nx, ny, nt, fx, fy, fr, seed):
    from torchwi_amd.propagator import td2d_cuda
    h, dt = 10.0, 1e-3
    sx, sy, ry = (int(fx * nx) + 0.5) * h, (int(fy * ny) + 0.5) * h, (int(fr * ny) + 0.5) * h
    print(f"case order={order} nx={nx} ny={ny} nt={nt} sx={sx} sy={sy} ry={ry}", flush=True)
    rng, vel, w = _vel_w(nx, ny, nt, seed)
    dimx, dimy = nx + order, ny + order
    n = dimx * dimy
    vpad = O.pad_vel(vel, order)
    frd_o, virt_o = O.forward(vpad, w, order, dimx, dimy, h, dt, sx, sy, ry)
    resid = rng.standard_normal(frd_o.shape).astype(np.float32)
    grad_o = O.backward(virt_o, vpad, resid, order, dimx, dimy, h, dt, ry)
    vp = torch.from_numpy(vpad).to(DEV)
    frd = torch.zeros(nt * nx, device=DEV)
    virt = torch.zeros(nt * n, device=DEV)
    u = [torch.zeros(n, device=DEV) for _ in range(3)]
    td2d_cuda.forward(frd, virt, *u, vp, torch.from_numpy(w), order, dimx, dimy, nt, h, dt, sx, sy, ry)
    g = td2d_cuda.backward(virt, *u, vp, torch.from_numpy(resid).to(DEV), order, dimx, dimy, nt, h, dt, ry)
    assert np.array_equal(frd.cpu().numpy().reshape(nt, nx), frd_o)
    assert np.array_equal(virt.cpu().numpy().reshape(nt, n), virt_o)
    assert np.array_equal(g.cpu().numpy(), grad_o)


@SETTINGS
@given(order=st.sampled_from([2, 4, 8]), nx=st.integers(20, 300), ny=st.integers(12, 150), nt=st.integers(5, 40),
       S=st.integers(1, 4), T=st.sampled_from([2, 3, 4]), m=st.integers(0, 4), seed=st.integers(0, 2**31 - 1))
def test_batched_and_blocked_random_cases_bitexact(order, nx, ny, nt, S, T, m, seed):
    """One step per launch and T steps per launch give the same gathers and
    per-shot gradients (checkpointed when K = T*m > 0), and the first shot's
    gather equals the oracle's."""
    from torchwi_amd import Time2dBatch
    h, dt = 10.0, 1e-3
    K = T * m
    rng, vel, w = _vel_w(nx, ny, nt, seed)
    sx = [(int(rng.integers(0, nx)) + 0.5) * h for _ in range(S)]
    sy = [(int(rng.integers(0, ny)) + 0.5) * h for _ in range(S)]
    ry = [(int(rng.integers(0, ny)) + 0.5) * h for _ in range(S)]
    print(f"case order={order} nx={nx} ny={ny} nt={nt} S={S} T={T} K={K}", flush=True)
    v = torch.from_numpy(vel).to(DEV)
    wt = torch.from_numpy(w)
    res = []
    for tb in (1, T):
        mdl = Time2dBatch(nx, ny, h, wt, dt, order, DEV, ckpt_every=K, time_block=tb)
        frd = mdl.modeling(v, sx, sy, ry).clone()
        loss, g = mdl.misfit_gradient(v, sx, sy, ry, 0.5 * frd)
        res.append((frd, [mdl.plan.grad_of(s).clone() for s in range(S)], g.clone()))
    assert torch.equal(res[0][0], res[1][0]), "gathers"
    for s in range(S):
        assert torch.equal(res[0][1][s], res[1][1][s]), f"shot {s} gradient"
    assert torch.equal(res[0][2], res[1][2])
    vpad = O.pad_vel(vel, order)
    frd_o, _ = O.forward(vpad, w, order, nx + order, ny + order, h, dt, sx[0], sy[0], ry[0], store_virt=False)
    assert np.array_equal(res[0][0][0].cpu().numpy(), frd_o)
