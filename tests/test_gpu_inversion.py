"""Device-resident inversion epochs (torchwi_amd.inversion.TimeInversion)
against the reference's own epoch structure (examples/time/inv_torch.py:
123-143) run through the drop-in ``Time2d`` module one shot at a time.  The
loss agrees to rounding.  With the velocity parameterisation the summed
gradient adds the per-shot gradients in shot order exactly as autograd
accumulates them, so the parameter is the same after every epoch, bit for
bit; with sloth or slowness the per-shot loop applies the change of
variables to each shot's gradient before adding them, the batched epoch to
their sum, so the parameters agree to rounding."""
import numpy as np
import pytest
import torch

from tests.helpers import small_case

pytestmark = pytest.mark.gpu
DEV = "cuda"


@pytest.mark.parametrize("parameter,optimizer,K", [("velocity", "adam", 0), ("sloth", "nag", 9),
                                                    ("slowness", "gd", 0)])
def test_inversion_epochs_match_the_per_shot_loop(parameter, optimizer, K):
    import torchwi_amd as torchwi
    from torchwi_amd.inversion import TimeInversion
    from torchwi_amd.parameter import SlothParameter, SlownessParameter, VelocityParameter
    order = 8
    c = small_case(order, nx=60, ny=30, nt=80, seed=23)
    h, dt = c["h"], c["dt"]
    S = 3
    sx = [(i * 25 + 5.5) * h for i in range(S)]
    sy = [2.5 * h] * S
    ry = [2.5 * h] * S
    print(f"case parameter={parameter} optimizer={optimizer} K={K}", flush=True)
    w = torch.from_numpy(c["w"])
    vtrue = torch.from_numpy(c["vel"]).to(DEV)
    vinit = torch.from_numpy(c["vel"] * 0.97).to(DEV)
    v0 = vinit.clone()
    obs = torchwi.Time2dBatch(c["nx"], c["ny"], h, w, dt, order, DEV, ckpt_every=0).modeling(vtrue, sx, sy, ry).clone()

    inv = TimeInversion(vinit, sx, sy, ry, h, w, dt, order, DEV, vmin=1400.0, vmax=4200.0, parameter=parameter,
                        optimizer=optimizer, lr=1e-3 if parameter == "velocity" else 1e-9, ckpt_every=K)
    # the reference's epoch, one drop-in Time2d call per shot
    cls = {"velocity": VelocityParameter, "sloth": SlothParameter, "slowness": SlownessParameter}[parameter]
    velgen = cls(vinit.clone(), 1400.0, 4200.0)
    lr = 1e-3 if parameter == "velocity" else 1e-9
    opt = {"adam": lambda p: torch.optim.Adam(p, lr=lr),
           "nag": lambda p: torch.optim.SGD(p, lr=lr, momentum=0.9, nesterov=True),
           "gd": lambda p: torch.optim.SGD(p, lr=lr)}[optimizer](velgen.parameters())
    mod = torchwi.Time2d(c["nx"], c["ny"], h, w, dt, order, DEV)
    gnorm0 = None
    for epoch in range(3):
        loss_b = inv.epoch(obs)
        opt.zero_grad()
        total = 0.0
        for s in range(S):
            frd = mod(velgen(), torch.tensor([sx[s]]), torch.tensor([sy[s]]), torch.tensor([ry[s]]))
            loss = 0.5 * torch.sum((frd - obs[s].view_as(frd)) ** 2)
            loss.backward()
            total += loss.item()
        if gnorm0 is None:
            gnorm0 = velgen.gradient().norm(float("inf"))
        velgen.normalize_gradient(gnorm0)
        opt.step()
        assert abs(float(loss_b) - total) <= 1e-5 * abs(total), (epoch, float(loss_b), total)
        a, b = inv.velgen.par.data, velgen.par.data
        if parameter == "velocity":
            assert torch.equal(a, b), f"epoch {epoch}: parameters differ"
        else:
            assert float(torch.linalg.norm(a - b) / torch.linalg.norm(b)) < 1e-6, f"epoch {epoch}"
    assert torch.equal(vinit, v0), "the caller's initial model was modified"
    moved = inv.velgen.par.data - cls(v0.clone()).par.data
    assert float(moved.abs().max()) > 0


@pytest.mark.parametrize("parameter,optimizer", [("velocity", "gd"), ("velocity", "adam"), ("slowness", "gd")])
def test_inversion_epochs_match_an_oracle_epoch(parameter, optimizer):
    """The same epochs computed on the CPU from the oracle: per shot the
    oracle's forward and adjoint (residual frd - d_obs, the adjoint input of
    0.5 * ||frd - d_obs||^2), the per-shot gradients cropped and added in shot
    order, then the parameterisation's chain rule, the first-epoch
    normalisation and the optimiser step with torch on the CPU
    (inv_torch.py:57-64, 123-143; Param.py:34-41).  Velocity: the normalised
    gradient is bit-identical every epoch, and so is the parameter after a
    gradient-descent step; Adam's step (square roots and divisions, computed
    by different kernels on the two devices) and slowness's chain rule agree
    to rounding.  The loss agrees to the FP32 summation order."""
    from oracle import oracle as O
    from torchwi_amd.inversion import TimeInversion
    from torchwi_amd.parameter import SlownessParameter, VelocityParameter
    order = 8
    c = small_case(order, nx=60, ny=30, nt=80, seed=31)
    nx, ny, nt, h, dt = c["nx"], c["ny"], c["nt"], c["h"], c["dt"]
    dimx, dimy = nx + order, ny + order
    S = 3
    sx = [(i * 25 + 5.5) * h for i in range(S)]
    sy = [2.5 * h] * S
    ry = [2.5 * h] * S
    print(f"case oracle epoch parameter={parameter} optimizer={optimizer}", flush=True)
    vpad_t = O.pad_vel(c["vel"], order)
    obs = np.stack([O.forward(vpad_t, c["w"], order, dimx, dimy, h, dt, sx[s], sy[s], ry[s], store_virt=False)[0]
                    for s in range(S)])
    vinit = (c["vel"] * 0.97).astype(np.float32)
    lr = {"velocity": 1e-3, "slowness": 1e-9}[parameter] if optimizer == "gd" else 2e-3
    inv = TimeInversion(torch.from_numpy(vinit).to(DEV), sx, sy, ry, h, torch.from_numpy(c["w"]), dt, order, DEV,
                        vmin=1400.0, vmax=4200.0, parameter=parameter, optimizer=optimizer, lr=lr, ckpt_every=0)
    cls = {"velocity": VelocityParameter, "slowness": SlownessParameter}[parameter]
    velgen = cls(torch.from_numpy(vinit.copy()), 1400.0, 4200.0)
    opt = (torch.optim.Adam(velgen.parameters(), lr=lr) if optimizer == "adam"
           else torch.optim.SGD(velgen.parameters(), lr=lr))
    gnorm0 = None
    obs_d = torch.from_numpy(obs).to(DEV)
    for epoch in range(3):
        loss_gpu = float(inv.epoch(obs_d))
        # the oracle epoch
        opt.zero_grad()
        v = velgen()
        vpad = O.pad_vel(v.detach().numpy(), order)
        gsum, loss = None, 0.0
        for s in range(S):
            frd, virt = O.forward(vpad, c["w"], order, dimx, dimy, h, dt, sx[s], sy[s], ry[s])
            r = frd - obs[s]
            loss += 0.5 * float(np.sum(r.astype(np.float64) ** 2))
            g = O.crop_grad(O.backward(virt, vpad, r, order, dimx, dimy, h, dt, ry[s]), order, nx, ny)
            gsum = g if gsum is None else gsum + g
        v.backward(torch.from_numpy(gsum))
        if gnorm0 is None:
            gnorm0 = velgen.gradient().norm(float("inf"))
        velgen.normalize_gradient(gnorm0)
        opt.step()
        assert abs(loss_gpu - loss) <= 1e-5 * loss, (epoch, loss_gpu, loss)
        ga, gb = inv.velgen.par.grad.cpu(), velgen.par.grad
        pa, pb = inv.velgen.par.data.cpu(), velgen.par.data
        if parameter == "velocity":
            assert torch.equal(ga, gb), f"epoch {epoch}: normalised gradient"
        else:
            assert float(torch.linalg.norm(ga - gb) / torch.linalg.norm(gb)) < 1e-6, f"epoch {epoch}: gradient"
        if parameter == "velocity" and optimizer == "gd":
            assert torch.equal(pa, pb), f"epoch {epoch}: parameter after the step"
        else:
            assert float(torch.linalg.norm(pa - pb) / torch.linalg.norm(pb)) < 1e-6, f"epoch {epoch}: parameter"
    assert float((inv.velgen.par.data.cpu() - cls(torch.from_numpy(vinit)).par.data).abs().max()) > 0
