"""GPU tests of the continuous-batching request loop."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DEV = "cuda"
BUDGET = 6 << 30
CAP = 512 << 20


def _stream(n=6, seed=3):
    from torchwi_amd.serving import make_stream
    return make_stream(n, 50.0, seed=seed, grids=((40, 96), (64, 130), (50, 200)), nt_range=(150, 260),
                       shots_range=(1, 9), device=DEV)


def test_results_identical_alone_and_in_the_mix():
    """Each request's result is bit-identical when served alone and when
    served while the others occupy the other slots."""
    from torchwi_amd.serving import Server
    mixed = Server(DEV, budget_bytes=BUDGET, slots=4, cap_bytes=CAP).run(_stream(), realtime=False)
    assert all(r.error is None for r in mixed)
    for k, req in enumerate(_stream()):
        print(f"case request={k} kind={req.kind} grid={req.nx}x{req.ny} nt={req.nt} shots={req.nshots}",
              flush=True)
        alone = Server(DEV, budget_bytes=BUDGET, slots=4, cap_bytes=CAP).run([req], realtime=False)[0]
        m = mixed[k]
        assert alone.error is None
        if req.kind == "forward":
            assert torch.equal(alone.frd, m.frd)
        else:
            assert alone.loss == m.loss
            assert torch.equal(alone.grad, m.grad)


def test_forward_request_matches_batched_engine():
    from torchwi_amd import Time2dBatch
    from torchwi_amd.serving import Server
    req = [r for r in _stream(8, seed=4) if r.kind == "forward"][0]
    res = Server(DEV, budget_bytes=BUDGET, slots=2, cap_bytes=CAP).run([req], realtime=False)[0]
    m = Time2dBatch(req.nx, req.ny, req.h, req.w, req.dt, req.order, DEV, ckpt_every=0)
    frd = m.modeling(req.vel.to(DEV), req.sx, req.sy, req.ry)
    assert torch.equal(res.frd, frd)


def test_chunked_gradient_equals_unchunked_sum():
    """A gradient request split into shot chunks (small cap) gives the same
    gradient as one chunk, up to the order of the shot sum."""
    from torchwi_amd.serving import Server
    req = [r for r in _stream(10, seed=6) if r.kind == "gradient" and r.nshots >= 3][0]
    one = Server(DEV, budget_bytes=BUDGET, slots=1, cap_bytes=CAP).run([req], realtime=False)[0]
    req2 = [r for r in _stream(10, seed=6) if r.kind == "gradient" and r.nshots >= 3][0]
    from torchwi_amd.serving import chunk_plan
    per_shot = chunk_plan(req2, CAP)[2] // chunk_plan(req2, CAP)[0]
    many = Server(DEV, budget_bytes=BUDGET, slots=1, cap_bytes=int(1.5 * per_shot)).run([req2], realtime=False)[0]
    assert many.chunks > 1
    g1, g2 = one.grad.cpu().numpy(), many.grad.cpu().numpy()
    assert np.linalg.norm(g1 - g2) <= 1e-6 * np.linalg.norm(g1)


def test_oversize_request_rejected_others_served():
    from torchwi_amd.serving import Server
    reqs = _stream(4, seed=8)
    big = _stream(1, seed=9)[0]
    big.rid = 99
    big.vel = torch.full((3000, 2000), 2000.0)
    srv = Server(DEV, budget_bytes=64 << 20, slots=2, cap_bytes=32 << 20)
    res = srv.run(reqs[:2] + [big], realtime=False)
    assert res[-1].error is not None and "HBM" in res[-1].error
    assert all(r.error is None for r in res[:2]) or any("HBM" in (r.error or "") for r in res[:2])
