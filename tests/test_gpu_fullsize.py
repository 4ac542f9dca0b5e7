"""GPU checks at the full sizes of the measured workloads, through properties
that do not depend on size (the oracle is checked against at small sizes,
where it finishes in seconds):

* 32 shots on the 384 x 1024 model, nt = 4000 (the benchmark's gradient):
  full virtual-source storage, checkpoints every 64 steps (one step per
  launch) and checkpoints every 66 steps with the temporally blocked kernel
  give bit-identical gathers, per-shot gradients, summed gradient and loss;
  shot 0 through the single-shot drop-in module (one direct launch per
  step, its own kernel) gives the same gather and gradient bit for bit;
* 4 shots on the 1000 x 3000 model, nt = 8000, checkpoints every 90 steps:
  one step per launch and 3 steps per launch agree bit for bit;
* 256 shots on uneven windows of the 600 x 4000 model, nt = 6000, in one
  plan of the temporally blocked kernel: two of the shots run alone with
  one step per launch give the same gathers and gradients bit for bit.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _gradients(c, vel, d_obs, S, **kw):
    from torchwi_amd import Time2dBatch
    m = Time2dBatch(c["nx"], c["ny"], c["h"], torch.from_numpy(c["w"]), c["dt"], c["order"], DEV, **kw)
    loss, g = m.misfit_gradient(vel, c["sx"][:S], c["sy"][:S], c["ry"][:S], d_obs)
    out = (float(loss), g.clone(), m.plan.frd_all().clone(), [m.plan.grad_of(s).clone() for s in range(S)],
           m.plan.time_block, m.plan.K)
    del m
    torch.cuda.empty_cache()
    return out


def _observed(c, vel, S, K):
    from torchwi_amd import Time2dBatch
    m = Time2dBatch(c["nx"], c["ny"], c["h"], torch.from_numpy(c["w"]), c["dt"], c["order"], DEV,
                    ckpt_every=K, time_block=1)
    # data of a slightly faster model, so the residuals are not zero
    d = m.modeling(vel * 1.01, c["sx"][:S], c["sy"][:S], c["ry"][:S]).clone()
    del m
    torch.cuda.empty_cache()
    return d


def _same(a, b, what):
    assert a[4] != b[4] or a[5] != b[5], "the two runs must differ in kernel or storage"
    assert a[0] == b[0], f"{what}: loss {a[0]} vs {b[0]}"
    assert torch.equal(a[2], b[2]), f"{what}: gathers"
    assert torch.equal(a[1], b[1]), f"{what}: summed gradient"
    for s, (x, y) in enumerate(zip(a[3], b[3])):
        assert torch.equal(x, y), f"{what}: gradient of shot {s}"


def test_benchmark_gradient_full_size_storage_and_kernels_agree():
    from torchwi_amd import Time2d
    from torchwi_amd import workloads as W
    c = W.dipping_384x1024()
    S = len(c["sx"])
    vel = torch.from_numpy(c["vel"]).to(DEV)
    print("case 32 shots 384x1024 nt=4000: observed data", flush=True)
    d_obs = _observed(c, vel, S, 64)
    print("case 32 shots 384x1024 nt=4000: full storage", flush=True)
    full = _gradients(c, vel, d_obs, S, ckpt_every=0, time_block=1)
    assert (full[4], full[5]) == (1, 0)
    assert np.isfinite(full[0]) and full[0] > 0
    assert bool(torch.isfinite(full[1]).all()) and float(full[1].abs().max()) > 0
    print("case 32 shots 384x1024 nt=4000: checkpoints every 64, one step per launch", flush=True)
    _same(full, _gradients(c, vel, d_obs, S, ckpt_every=64, time_block=1), "K=64")
    print("case 32 shots 384x1024 nt=4000: checkpoints every 66, 3 steps per launch", flush=True)
    _same(full, _gradients(c, vel, d_obs, S, ckpt_every=66, time_block=3), "K=66 blocked")

    print("case 384x1024 nt=4000: shot 0 through the single-shot module", flush=True)
    m = Time2d(c["nx"], c["ny"], c["h"], torch.from_numpy(c["w"]), c["dt"], c["order"], DEV)
    v = vel.clone().requires_grad_(True)
    frd = m(v, torch.tensor([c["sx"][0]]), torch.tensor([c["sy"][0]]), torch.tensor([c["ry"][0]]))
    assert torch.equal(frd, full[2][0])
    frd.backward(frd - d_obs[0])
    ne = c["order"] // 2
    g0 = full[3][0][ne:ne + c["ny"], ne:ne + c["nx"]].transpose(0, 1)
    assert torch.equal(v.grad, g0)
    del m, frd, v
    torch.cuda.empty_cache()


def test_large_model_full_size_blocked_equals_one_step():
    from torchwi_amd import workloads as W
    c = W.dipping_1000x3000()
    S = 4
    vel = torch.from_numpy(c["vel"]).to(DEV)
    print("case 4 shots 1000x3000 nt=8000: observed data", flush=True)
    d_obs = _observed(c, vel, S, 90)
    print("case 4 shots 1000x3000 nt=8000: checkpoints every 90, one step per launch", flush=True)
    one = _gradients(c, vel, d_obs, S, ckpt_every=90, time_block=1)
    assert np.isfinite(one[0]) and one[0] > 0
    print("case 4 shots 1000x3000 nt=8000: checkpoints every 90, 3 steps per launch", flush=True)
    _same(one, _gradients(c, vel, d_obs, S, ckpt_every=90, time_block=3), "blocked")


def test_uneven_windows_full_size_batched_equals_alone():
    from torchwi_amd import Time2dWindows
    from torchwi_amd import workloads as W
    c = W.uneven_600x4000()
    order, nx, ny, nt, h, dt = c["order"], c["nx"], c["ny"], c["nt"], c["h"], c["dt"]
    wins = c["windows"]
    w = torch.from_numpy(c["w"])
    vel = torch.from_numpy(c["vel"]).to(DEV)
    print("case 256 uneven windows of 600x4000 nt=6000: batched, 3 steps per launch", flush=True)
    m = Time2dWindows(nx, ny, h, w, dt, order, wins, DEV, time_block=3)
    assert m.plan.time_block == 3 and m.plan.K > 0
    d_obs = [f.clone() for f in m.modeling(vel * 1.01)]
    loss, g = m.misfit_gradient(vel, d_obs)
    assert np.isfinite(float(loss)) and float(loss) > 0
    K = m.plan.K
    for s in (0, len(wins) // 2):
        print(f"case uneven window {s} alone, one step per launch", flush=True)
        frd_b = m.plan.frd_of(s).clone()
        g_b = m.plan.grad_of(s).clone()
        one = Time2dWindows(nx, ny, h, w, dt, order, [wins[s]], DEV, ckpt_every=K, time_block=1)
        f1 = one.modeling(vel * 1.01)[0]
        assert torch.equal(f1, d_obs[s]), f"window {s}: gathers"
        one.misfit_gradient(vel, [d_obs[s]])
        assert torch.equal(one.plan.grad_of(0), g_b), f"window {s}: gradient"
        assert torch.equal(one.plan.frd_of(0), frd_b), f"window {s}: gathers at the gradient's model"
        del one
        torch.cuda.empty_cache()
    del m
    torch.cuda.empty_cache()
