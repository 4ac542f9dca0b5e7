"""Device-resident inversion epochs (torchwi_amd.inversion.TimeInversion)
against the reference's own epoch structure (examples/time/inv_torch.py:
123-143) run through the drop-in ``Time2d`` module one shot at a time.  The
loss agrees to rounding.  With the velocity parameterisation the summed
gradient adds the per-shot gradients in shot order exactly as autograd
accumulates them, so the parameter is the same after every epoch, bit for
bit; with sloth or slowness the per-shot loop applies the change of
variables to each shot's gradient before adding them, the batched epoch to
their sum, so the parameters agree to rounding."""
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
