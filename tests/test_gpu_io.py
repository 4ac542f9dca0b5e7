"""RSF inputs through the batched engine and back to gather files (GPU).

The flow of the reference's examples (examples/time/forward_modeling.py and
inv_torch.py): wavelet, velocity and shot table from RSF files, one gather
file per shot, then a misfit gradient against those files.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests.helpers import small_case
from torchwi_amd.io import read_gathers, rsf, write_gather

pytestmark = pytest.mark.gpu
DEV = "cuda"


def test_rsf_inputs_gather_files_and_gradient(tmp_path):
    from torchwi_amd import Time2dBatch
    order = 8
    c = small_case(order, nx=60, ny=30, nt=90, seed=17)
    nx, ny, nt, h, dt = c["nx"], c["ny"], c["nt"], c["h"], c["dt"]
    S = 3
    sxy = np.array([[(i + 0.5) * h, 2.5 * h] for i in (5, 30, 55)], np.float32)
    fw, fv, fs = (str(tmp_path / n) for n in ("wavelet.rsf", "vel.rsf", "shot.rsf"))
    rsf.tofile(fw, c["w"], n1=nt, d1=dt)
    rsf.tofile(fv, c["vel"], n1=ny, d1=h, n2=nx, d2=h)
    rsf.tofile(fs, sxy, n1=2, n2=S)
    print("case: rsf round trip through Time2dBatch, nx=%d ny=%d nt=%d" % (nx, ny, nt), flush=True)

    nt_, dt_, w = rsf.fromfile(fw, "n1 d1 data")
    ny_, nx_, h_, vel = rsf.fromfile(fv, "n1 n2 d1 data")
    nshot, shots = rsf.fromfile(fs, "n2 data")
    assert (nt_, ny_, nx_, nshot) == (nt, ny, nx, S) and vel.shape == (nx, ny)
    m = Time2dBatch(nx_, ny_, h_, torch.from_numpy(w), dt_, order, DEV, ckpt_every=0)
    v = torch.from_numpy(vel).to(DEV)
    frd = m.modeling(v, shots[:, 0], shots[:, 1], shots[:, 1]).clone()
    pat = str(tmp_path / "seismo.%04d")
    for s in range(S):
        write_gather(pat, s, frd[s])
    d_obs = read_gathers(pat, range(S), nt, nx, device=DEV)
    assert torch.equal(d_obs, frd)

    # gather of shot 1 equals the oracle on the same inputs
    vpad = O.pad_vel(c["vel"], order)
    args = (order, nx + order, ny + order, h, dt)
    frd_o, virt_o = O.forward(vpad, c["w"], *args, float(shots[1, 0]), float(shots[1, 1]), float(shots[1, 1]))
    assert np.array_equal(d_obs[1].cpu().numpy(), frd_o)

    # zero residual: zero loss and gradient; a scaled copy: the oracle's gradient
    loss, g = m.misfit_gradient(v, shots[:, 0], shots[:, 1], shots[:, 1], d_obs)
    assert float(loss) == 0.0 and not torch.any(g != 0)
    loss, g = m.misfit_gradient(v, shots[:, 0], shots[:, 1], shots[:, 1], 0.5 * d_obs)
    resid = (frd[1] - 0.5 * d_obs[1]).cpu().numpy()
    g1 = O.crop_grad(O.backward(virt_o, vpad, resid, *args, float(shots[1, 1])), order, nx, ny)
    got = m.plan.grad_of(1)[order // 2:order // 2 + ny, order // 2:order // 2 + nx].cpu().numpy().T
    assert np.array_equal(got, g1)
