"""GPU checks at the benchmark workloads' own full sizes, against the oracle
directly where the oracle finishes in seconds to a minute on the host cores,
and through exact properties elsewhere.

* 32 shots on the 384 x 1024 model, nt = 4000, full storage: shot 0's gather
  and gradient equal the oracle's forward and adjoint (full virtual-source
  storage, 6.5 GB of host memory) bit for bit; the FP32-vs-FP64 error of
  both is recorded.
* 64 shots on the 1000 x 3000 model, nt = 8000, checkpoints every 90 steps
  (the differentiation-at-scale plan, 215 GB): the temporally blocked kernel
  and the one-step kernel give bit-identical gathers, loss and per-shot
  gradients for all 64 shots; shot 0's gather equals the oracle's
  forward-only run bit for bit; the gradient agrees with central finite
  differences along 5 seeded smooth directions within the FP32 noise bound
  derived below.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests.helpers import record_drift, rel_l2

pytestmark = pytest.mark.gpu
DEV = "cuda"


def test_384x1024_32_shots_shot0_gather_and_gradient_equal_oracle():
    from torchwi_amd import Time2dBatch
    from torchwi_amd import workloads as W
    c = W.dipping_384x1024()
    order, nx, ny, nt, h, dt = c["order"], c["nx"], c["ny"], c["nt"], c["h"], c["dt"]
    dimx, dimy = nx + order, ny + order
    print("case 32 shots 384x1024 nt=4000 full storage: batched gradient", flush=True)
    m = Time2dBatch(nx, ny, h, torch.from_numpy(c["w"]), dt, order, DEV, ckpt_every=0, time_block=1)
    vel = torch.from_numpy(c["vel"]).to(DEV)
    d_obs = m.modeling(vel * 1.01, c["sx"], c["sy"], c["ry"]).clone()
    m.misfit_gradient(vel, c["sx"], c["sy"], c["ry"], d_obs)
    p = m.plan
    frd0 = p.frd_of(0).cpu().numpy()
    resid0 = p.resid_of(0).cpu().numpy()
    grad0 = p.grad_of(0)[:, :dimx].cpu().numpy().ravel()
    assert np.array_equal(resid0, frd0 - d_obs[0].cpu().numpy())
    del m, p, d_obs
    torch.cuda.empty_cache()

    print("case 384x1024 shot 0: oracle forward + adjoint (FP32, full storage)", flush=True)
    vpad = O.pad_vel(c["vel"], order)
    a = (order, dimx, dimy, h, dt)
    frd_o, virt_o = O.forward(vpad, c["w"], *a, c["sx"][0], c["sy"][0], c["ry"][0])
    assert np.array_equal(frd0, frd_o), "shot 0 gather"
    grad_o = O.backward(virt_o, vpad, resid0, *a, c["ry"][0])
    del virt_o
    assert np.array_equal(grad0, grad_o), "shot 0 gradient"
    print("case 384x1024 shot 0: FP64 forward + adjoint", flush=True)
    frd64, virt64 = O.forward64(vpad, c["w"], *a, c["sx"][0], c["sy"][0], c["ry"][0])
    grad64 = O.backward64(virt64, vpad, resid0, *a, c["ry"][0])
    del virt64
    record_drift("384x1024, order 8, nt=4000, shot 0 of 32", frd=rel_l2(frd0, frd64),
                 grad=rel_l2(grad0, grad64))
    assert rel_l2(frd0, frd64) < 1e-4 and rel_l2(grad0, grad64) < 1e-3


def _misfit64(frd, d_obs):
    J = 0.0
    for s in range(frd.shape[0]):
        r = (frd[s] - d_obs[s]).double().view(-1)
        J += 0.5 * float(torch.dot(r, r))
    return J


def test_1000x3000_64_shots_blocked_equals_one_step_oracle_gather_and_finite_differences():
    """Central differences along 5 seeded smooth directions d_k (seeds
    10-14: white noise, Gaussian sigma = 20 cells, unit max-norm, zero
    within 10 cells of the left, right and bottom edges), eps = 1e-3:
    FD(e) = (J(v(1 + e d)) - J(v(1 - e d))) / (2e), J = 0.5 ||d(v) - d_obs||^2
    summed in FP64 over the FP32 gathers, against <g, v d>.

    Error budget.  FD(e) = <g, v d> + T(e) + N(e).
    * T(e) = (e^2 / 6) J''' + O(e^4) is the central difference's truncation.
      It is removed to O(e^4) by Richardson extrapolation from e and 2e:
      FD_R = (4 FD(e) - FD(2e)) / 3.
    * N(e) is the FP32 noise.  With the computed gathers d~ = d + err,
      ||err|| <= gamma ||d|| (gamma = the gathers' relative L2 error against
      an FP64 run of the same recursion, measured here on shot 0 and doubled
      for the other shots), J~ - J = <r, err> + O(err^2), so
      |N(e)| <= 2 gamma ||r|| ||d|| / (2e) = gamma ||r|| ||d|| / e, and for FD_R
      |N_R| <= (4/3 + 1/6) gamma ||r|| ||d|| / e = 1.5 gamma ||r|| ||d|| / e.
      (The observed data's own error is common to J(+e) and J(-e); it
      enters only through r(+e) - r(-e) = O(e), a small fraction of this.)
    The check is |FD_R - <g, v d>| <= 3 gamma_0 ||r|| ||d|| / e (gamma_0 of shot
    0, the factor 2 for the other shots), which must also be well below the
    derivative itself (a bound larger than |<g, v d>| / 10 fails the test as
    vacuous)."""
    from torchwi_amd import Time2dBatch
    from torchwi_amd import workloads as W
    from torchwi_amd.diffcheck import direction
    c = W.dipping_1000x3000(nshots=64, nt=8000)
    order, nx, ny, nt, h, dt = c["order"], c["nx"], c["ny"], c["nt"], c["h"], c["dt"]
    S, K, eps = 64, 90, 1e-3
    w = torch.from_numpy(c["w"])
    vel = torch.from_numpy(c["vel"]).to(DEV)
    vtrue = vel * (1.0 + 0.02 * torch.from_numpy(direction(nx, ny, seed=99, sigma=30.0)).to(DEV))
    pos = (c["sx"], c["sy"], c["ry"])

    print("case 1000x3000 64 shots: 64 shots 1000x3000 nt=8000 K=90, blocked kernel: observed data", flush=True)
    m = Time2dBatch(nx, ny, h, w, dt, order, DEV, ckpt_every=K, time_block=3)
    d_obs = m.modeling(vtrue, *pos).clone()
    assert m.plan.time_block == 3 and m.plan.K == K
    print("case 1000x3000 64 shots: blocked checkpointed gradient", flush=True)
    loss_b, g = m.misfit_gradient(vel, *pos, d_obs)
    loss_b, g = float(loss_b), g.clone()
    frd_b = m.plan.frd_all().clone()
    grads_b = torch.stack([m.plan.grad_of(s) for s in range(S)])
    assert bool(torch.isfinite(g).all()) and float(g.abs().max()) > 0
    dn = float(torch.linalg.vector_norm(frd_b.double()))
    rn = float(torch.linalg.vector_norm((frd_b - d_obs).double()))

    fds = []
    for k in range(5):
        dvec = torch.from_numpy(direction(nx, ny, seed=10 + k)).to(DEV)
        J = {}
        for e in (eps, -eps, 2 * eps, -2 * eps):
            print(f"case 1000x3000 64 shots: direction seed {10 + k}, eps {e:+g}", flush=True)
            J[e] = _misfit64(m.modeling(vel * (1.0 + e * dvec), *pos), d_obs)
        fd1 = (J[eps] - J[-eps]) / (2 * eps)
        fd2 = (J[2 * eps] - J[-2 * eps]) / (4 * eps)
        an = float(torch.sum((g * vel * dvec).double()))
        fds.append((10 + k, fd1, (4 * fd1 - fd2) / 3, an))
    del m
    torch.cuda.empty_cache()

    print("case 1000x3000 64 shots: one-step checkpointed gradient", flush=True)
    m1 = Time2dBatch(nx, ny, h, w, dt, order, DEV, ckpt_every=K, time_block=1)
    loss_1, g1 = m1.misfit_gradient(vel, *pos, d_obs)
    assert m1.plan.time_block == 1
    assert float(loss_1) == loss_b, "loss"
    assert torch.equal(m1.plan.frd_all(), frd_b), "gathers"
    for s in range(S):
        assert torch.equal(m1.plan.grad_of(s), grads_b[s]), f"gradient of shot {s}"
    assert torch.equal(g1, g), "summed gradient"
    del m1, grads_b, d_obs, g1
    torch.cuda.empty_cache()

    print("case 1000x3000 64 shots: shot 0 gather, oracle forward-only FP32 and FP64", flush=True)
    vpad = O.pad_vel(c["vel"], order)
    a = (order, nx + order, ny + order, h, dt, c["sx"][0], c["sy"][0], c["ry"][0])
    f0 = frd_b[0].cpu().numpy()
    frd_o, _ = O.forward(vpad, c["w"], *a, store_virt=False)
    assert np.array_equal(f0, frd_o), "shot 0 gather vs oracle"
    frd64, _ = O.forward64(vpad, c["w"], *a, store_virt=False)
    gamma0 = rel_l2(f0, frd64)
    record_drift("1000x3000, order 8, nt=8000, shot 0 of 64", frd=gamma0)

    bound = 3.0 * max(gamma0, 1e-7) * rn * dn / eps
    for seed, fd1, fdr, an in fds:
        print(f"1000x3000 FD seed {seed}: adjoint {an:.6e} FD {fd1:.6e} (rel {abs(fd1 - an) / abs(an):.2e}) "
              f"Richardson {fdr:.6e} (|diff| {abs(fdr - an):.3e}, bound {bound:.3e})", flush=True)
        record_drift(f"1000x3000, 64 shots: central differences, direction seed {seed}, eps 1e-3", adjoint=an, fd=fd1,
                     fd_richardson=fdr, fd_rel_err=abs(fd1 - an) / abs(an),
                     richardson_rel_err=abs(fdr - an) / abs(an), noise_bound_abs=bound)
    scale = max(abs(an) for _, _, _, an in fds)
    assert bound < 0.1 * scale, "noise bound %.3e is not below a tenth of the largest derivative %.3e" % (bound, scale)
    for seed, fd1, fdr, an in fds:
        assert abs(fdr - an) <= bound, f"direction {seed}: |FD_R - <g, v d>| = {abs(fdr - an):.3e} > {bound:.3e}"
