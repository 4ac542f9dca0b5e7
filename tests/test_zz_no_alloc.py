"""The allocation-free plan entry point (tw_plan_create_static) against the
workspace entry point (tw_plan_create) on the same inputs: gathers, loss and
per-shot gradients bit for bit, on two plans in a row, the second smaller,
each through the C ABI (torchwi_amd.shots.ShotPlan binds both)."""
import numpy as np
import pytest
import torch

from tests.helpers import small_case

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _plan(c, S, K, tb, static):
    from torchwi_amd.shots import ShotPlan
    nx, ny, h = c["nx"], c["ny"], c["h"]
    w = torch.from_numpy(c["w"])
    shots = [dict(model=0, sx=(3 + 7 * s + 0.5) * h, sy=2.5 * h, ry=2.5 * h, rx0=0, nrec=nx, w=w)
             for s in range(S)]
    p = ShotPlan(c["order"], c["nt"], h, c["dt"], [(nx, ny)], shots, K, DEV, graph_steps=0, time_block=tb,
                 static_workspace=static)
    p.models[0].load(torch.from_numpy(c["vel"]).to(DEV))
    return p


def _gradient(c, S, K, tb, static):
    p = _plan(c, S, K, tb, static)
    p.forward()
    r = p.resid_all()
    rng = np.random.default_rng(5)
    r.copy_(p.frd_all() - torch.from_numpy(rng.standard_normal(tuple(r.shape)).astype(np.float32)).to(DEV))
    p.backward()
    torch.cuda.synchronize()
    out = (p.frd_all().clone(), [p.grad_of(s).clone() for s in range(S)])
    return p, out


@pytest.mark.parametrize("case", [dict(order=8, nx=150, ny=70, nt=150, S=5, K=12, tb=3),
                                  dict(order=4, nx=61, ny=33, nt=90, S=2, K=0, tb=1)])
def test_static_workspace_plan_equals_workspace_plan(case):
    """Run as two calls in a row: the second case is the smaller one."""
    print("case no-alloc %s" % case, flush=True)
    c = small_case(case["order"], nx=case["nx"], ny=case["ny"], nt=case["nt"], seed=case["nx"])
    ref_plan, ref = _gradient(c, case["S"], case["K"], case["tb"], False)
    del ref_plan
    st_plan, got = _gradient(c, case["S"], case["K"], case["tb"], True)
    # the workspace is held: a second allocation-free plan on this device is refused
    with pytest.raises(RuntimeError, match="held by another plan"):
        _plan(c, 1, case["K"], case["tb"], True)
    del st_plan
    assert torch.equal(got[0], ref[0]), "gathers"
    for s, (a, b) in enumerate(zip(got[1], ref[1])):
        assert torch.equal(a, b), f"gradient of shot {s}"
    # released on destruction: the next call may take it
    again = _plan(c, 1, case["K"], case["tb"], True)
    del again
