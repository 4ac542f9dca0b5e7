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


def _check_same(alone, m):
    assert alone.error is None and m.error is None, (alone.error, m.error)
    if alone.kind == "forward":
        assert torch.equal(alone.frd, m.frd)
    else:
        assert alone.loss == m.loss
        assert torch.equal(alone.grad, m.grad)


def test_requests_of_one_key_share_a_plan_and_match_alone():
    """Due chunks of requests with the same kind, grid, order and time axis
    run in one plan (each on its own model and wavelet); every result is
    bit-identical to the request served alone."""
    from torchwi_amd.serving import Server, make_request
    spec = [("gradient", 3, 11), ("gradient", 2, 12), ("forward", 3, 13), ("forward", 1, 14),
            ("gradient", 4, 15)]
    mk = lambda k: [make_request(kind, 48, 120, 180, S, 10.0, seed, device=DEV, rid=i)  # noqa: E731
                    for i, (kind, S, seed) in enumerate(spec)][k]
    srv = Server(DEV, budget_bytes=BUDGET, slots=8, cap_bytes=CAP)
    mixed = srv.run([mk(i) for i in range(len(spec))], realtime=False)
    assert srv.max_batch_requests >= 3, "same-key requests were not batched together"
    assert all(r.shared_batches > 0 for r in mixed)
    for i in range(len(spec)):
        print(f"case merged request={i} {spec[i]}", flush=True)
        alone = Server(DEV, budget_bytes=BUDGET, slots=8, cap_bytes=CAP, merge=False).run([mk(i)],
                                                                                          realtime=False)[0]
        assert alone.shared_batches == 0
        _check_same(alone, mixed[i])


def test_failing_request_is_isolated():
    """A request whose second chunk cannot be built (a source below the
    grid) ends with an error after its first chunk ran; its slot and budget
    are released and the other requests, including one batched with it,
    finish with the same results as alone."""
    from torchwi_amd.serving import Server, chunk_plan, make_request
    good = [make_request("gradient", 40, 96, 150, 4, 12.0, 21, device=DEV, rid=0),
            make_request("forward", 40, 96, 150, 3, 12.0, 22, device=DEV, rid=1),
            make_request("gradient", 40, 96, 150, 4, 12.0, 23, device=DEV, rid=2)]
    bad = make_request("gradient", 40, 96, 150, 4, 12.0, 24, device=DEV, rid=3)
    bad.sy[3] = (40 + 30) * bad.h          # below the model: refused by the library
    from torchwi_amd.shots import default_ckpt_interval, plan_bytes
    K = default_ckpt_interval(150, 3)
    cap = plan_bytes([(96, 40, 8)], [dict(model=0, nrec=96)] * 2, 150, K) + 4096
    assert chunk_plan(bad, cap)[:2] == (2, K)   # two shots per chunk: the bad shot is in the second
    srv = Server(DEV, budget_bytes=BUDGET, slots=4, cap_bytes=cap)
    res = srv.run(good + [bad], realtime=False)
    assert res[3].error is not None and "source" in res[3].error, res[3].error
    assert res[3].chunks == 1
    assert srv.in_use == 0 and sorted(srv.free_slots) == list(range(4)) and not srv.active
    assert all(r.error is None for r in res[:3])
    for i, req in enumerate(good):
        fresh = make_request(req.kind, 40, 96, 150, req.nshots, 12.0, 21 + i, device=DEV, rid=i)
        alone = Server(DEV, budget_bytes=BUDGET, slots=4, cap_bytes=cap, merge=False).run([fresh],
                                                                                          realtime=False)[0]
        _check_same(alone, res[i])


def test_serving_grid_classes_alone_equal_mix():
    """One request of each serving grid class (256x512, 512x1024, 800x2400,
    1200x3600 at h = 10 m, dt = 1 ms) with nt >= 2000, plus a second request
    sharing a key, served as one stream: each result is bit-identical to the
    request served alone."""
    from torchwi_amd.serving import GRIDS, Server, make_request
    spec = [("gradient", 256, 512, 2400, 6, 8.0), ("forward", 512, 1024, 3000, 5, 12.0),
            ("gradient", 800, 2400, 2000, 4, 6.0), ("forward", 1200, 3600, 2200, 3, 10.0),
            ("gradient", 256, 512, 2400, 3, 14.0)]
    assert {(s[1], s[2]) for s in spec} == set(GRIDS)
    mk = lambda i: make_request(spec[i][0], *spec[i][1:], seed=100 + i, device=DEV, rid=i)  # noqa: E731
    srv = Server(DEV, slots=4)
    mixed = srv.run([mk(i) for i in range(len(spec))], realtime=False)
    print(f"serving grid classes mix: max concurrent {srv.max_concurrent}, largest batch {srv.max_batch_requests} requests",
          flush=True)
    for i in range(len(spec)):
        print(f"case serving grid classes request={i} {spec[i]} alone", flush=True)
        alone = Server(DEV, slots=4, merge=False).run([mk(i)], realtime=False)[0]
        _check_same(alone, mixed[i])


def test_cancel_waiting_and_in_flight_requests():
    """Cancelling a waiting request removes it at once; cancelling one whose
    chunk is on the device (in a plan it shares with another request) lets
    that chunk run, launches none of its later chunks and releases its slot
    and budget; the other requests finish with their results alone, bit for
    bit."""
    from torchwi_amd.serving import Server, chunk_plan, make_request
    from torchwi_amd.shots import default_ckpt_interval, plan_bytes
    mk = lambda: [make_request("gradient", 40, 96, 150, 4, 12.0, 31, device=DEV, rid=0),  # noqa: E731
                  make_request("gradient", 40, 96, 150, 4, 12.0, 32, device=DEV, rid=1),
                  make_request("forward", 40, 96, 150, 3, 12.0, 33, device=DEV, rid=2),
                  make_request("gradient", 40, 96, 150, 2, 12.0, 34, device=DEV, rid=3)]
    K = default_ckpt_interval(150, 3)
    cap = plan_bytes([(96, 40, 8)], [dict(model=0, nrec=96)] * 2, 150, K) + 4096
    reqs = mk()
    assert chunk_plan(reqs[1], cap)[0] == 2          # request 1 runs as two chunks of two shots
    srv = Server(DEV, budget_bytes=BUDGET, slots=2, cap_bytes=cap)
    res = {r.rid: srv.submit(r) for r in reqs}
    srv.poll()                                        # admits 0 and 1 (two slots), launches their first chunks
    assert [w[0].rid for w in srv.waiting] == [2, 3]
    assert any(m[0].req.rid == 1 for b in srv.batches for m in b.members), "request 1 is not in flight"
    print("case cancel: waiting request 3, in-flight request 1", flush=True)
    assert srv.cancel(3) and res[3].cancelled and res[3].error == "cancelled"
    assert srv.cancel(1) and not srv.cancel(42)
    while not srv.idle():
        srv.poll()
    torch.cuda.synchronize()
    assert res[1].cancelled and res[1].error == "cancelled" and res[1].chunks == 1
    assert srv.in_use == 0 and sorted(srv.free_slots) == [0, 1] and not srv.active and not srv.batches
    for rid in (0, 2):
        fresh = mk()[rid]
        alone = Server(DEV, budget_bytes=BUDGET, slots=2, cap_bytes=cap, merge=False).run([fresh],
                                                                                          realtime=False)[0]
        _check_same(alone, res[rid])
