"""GPU parity of the gfx950 propagator against the CPU oracle.

The HIP kernels reproduce the reference CPU path's mixed-precision
arithmetic, so the bar here is bit-exact equality with the oracle (which is
itself bit-exact with the compiled reference, tests/test_oracle.py).  The
parity tolerance for floating point, relative L2 <= 1e-5, is asserted as well, and the FP32
error against the FP64 run of the same recursion is recorded per result (tests.helpers.record_drift).
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests.helpers import record_drift, rel_l2, small_case

pytestmark = pytest.mark.gpu

TOL = 1e-5  # relative L2 parity tolerance for floating-point results

DEV = "cuda"


def _run_module_forward_backward(c, resid=None):
    from torchwi_amd.propagator import td2d_cuda
    order, dimx, dimy, nt = c["order"], c["dimx"], c["dimy"], c["nt"]
    n = dimx * dimy
    vpad = torch.from_numpy(O.pad_vel(c["vel"], order)).to(DEV)
    frd = torch.zeros(nt * (dimx - order), device=DEV)
    virt = torch.zeros(nt * n, device=DEV)
    u = [torch.zeros(n, device=DEV) for _ in range(3)]
    w = torch.from_numpy(c["w"])
    td2d_cuda.forward(frd, virt, *u, vpad, w, order, dimx, dimy, nt, c["h"], c["dt"],
                      c["sx"], c["sy"], c["ry"])
    out = dict(frd=frd.cpu().numpy().reshape(nt, -1), virt=virt.cpu().numpy().reshape(nt, n))
    if resid is not None:
        g = td2d_cuda.backward(virt, *u, vpad, torch.from_numpy(resid).to(DEV), order, dimx, dimy,
                               nt, c["h"], c["dt"], c["ry"])
        out["grad"] = g.cpu().numpy()
    return out


@pytest.mark.parametrize("order", [2, 4, 8])
@pytest.mark.parametrize("shape", [(37, 23), (150, 70)])
def test_forward_backward_bitexact(order, shape):
    nx, ny = shape
    c = small_case(order, nx=nx, ny=ny, nt=160, seed=order + nx)
    print(f"case order={order} nx={nx} ny={ny}", flush=True)
    vpad = O.pad_vel(c["vel"], order)
    args = (order, c["dimx"], c["dimy"], c["h"], c["dt"])
    frd_o, virt_o = O.forward(vpad, c["w"], *args, c["sx"], c["sy"], c["ry"])
    rng = np.random.default_rng(1)
    resid = rng.standard_normal(frd_o.shape).astype(np.float32)
    grad_o = O.backward(virt_o, vpad, resid, *args, c["ry"])
    got = _run_module_forward_backward(c, resid)
    frd64, virt64 = O.forward64(vpad, c["w"], *args, c["sx"], c["sy"], c["ry"])
    grad64 = O.backward64(virt64, vpad, resid, *args, c["ry"])
    print(f"  rel-L2 vs oracle f32: frd {rel_l2(got['frd'], frd_o):.3e} grad {rel_l2(got['grad'], grad_o):.3e}")
    record_drift(f"small order={order} nx={nx} ny={ny} nt=160", frd=rel_l2(got["frd"], frd64),
                 grad=rel_l2(got["grad"], grad64))
    assert rel_l2(got["frd"], frd_o) <= TOL and rel_l2(got["grad"], grad_o) <= TOL
    assert np.array_equal(got["frd"], frd_o)
    assert np.array_equal(got["virt"], virt_o)
    assert np.array_equal(got["grad"], grad_o)


@pytest.mark.parametrize("order", [2, 4, 8])
def test_forward_only_bitexact(order):
    from torchwi_amd.propagator import td2d_forward_cuda
    c = small_case(order, nx=90, ny=41, nt=150, seed=7)
    vpad = O.pad_vel(c["vel"], order)
    frd_o, _ = O.forward(vpad, c["w"], order, c["dimx"], c["dimy"], c["h"], c["dt"],
                         c["sx"], c["sy"], c["ry"], store_virt=False)
    n = c["dimx"] * c["dimy"]
    frd = torch.zeros(c["nt"] * c["nx"], device=DEV)
    u = [torch.zeros(n, device=DEV) for _ in range(3)]
    td2d_forward_cuda.forward(frd, *u, torch.from_numpy(vpad).to(DEV), torch.from_numpy(c["w"]),
                              order, c["dimx"], c["dimy"], c["nt"], c["h"], c["dt"],
                              c["sx"], c["sy"], c["ry"])
    assert np.array_equal(frd.cpu().numpy().reshape(c["nt"], -1), frd_o)


@pytest.mark.parametrize("order", [2, 4, 8])
def test_exa_bitexact(order):
    from torchwi_amd.propagator import td2d_exa_cuda
    c = small_case(order, nx=70, ny=33, nt=140, seed=11)
    vpad = O.pad_vel(c["vel"], order)
    args = (order, c["dimx"], c["dimy"], c["h"], c["dt"])
    frd_o, exa_o, iexa_o = O.exa_forward(vpad, c["w"], *args, c["sx"], c["sy"], c["ry"])
    resid = np.random.default_rng(2).standard_normal(frd_o.shape).astype(np.float32)
    grad_o = O.exa_backward(exa_o, iexa_o, vpad, resid, *args, c["ry"])
    n = c["dimx"] * c["dimy"]
    frd = torch.zeros(c["nt"] * c["nx"], device=DEV)
    exa = torch.zeros(n, device=DEV)
    iexa = torch.zeros(n, dtype=torch.int16, device=DEV)
    u = [torch.zeros(n, device=DEV) for _ in range(3)]
    vp = torch.from_numpy(vpad).to(DEV)
    td2d_exa_cuda.forward(frd, exa, iexa, *u, vp, torch.from_numpy(c["w"]), order, c["dimx"],
                          c["dimy"], c["nt"], c["h"], c["dt"], c["sx"], c["sy"], c["ry"])
    g = td2d_exa_cuda.backward(exa, iexa, *u, vp, torch.from_numpy(resid).to(DEV), order,
                               c["dimx"], c["dimy"], c["nt"], c["h"], c["dt"], c["ry"])
    assert np.array_equal(frd.cpu().numpy().reshape(c["nt"], -1), frd_o)
    assert np.array_equal(exa.cpu().numpy(), exa_o)
    assert np.array_equal(iexa.cpu().numpy(), iexa_o)
    assert np.array_equal(g.cpu().numpy(), grad_o)


def test_time2d_autograd_matches_oracle():
    """The drop-in module surface: Time2d forward + loss.backward()."""
    import torchwi_amd as torchwi
    order = 8
    c = small_case(order, nx=64, ny=40, nt=180, seed=3)
    vpad = O.pad_vel(c["vel"], order)
    args = (order, c["dimx"], c["dimy"], c["h"], c["dt"])
    frd_o, virt_o = O.forward(vpad, c["w"], *args, c["sx"], c["sy"], c["ry"])
    true = (0.9 * frd_o).astype(np.float32)
    grad_o = O.crop_grad(O.backward(virt_o, vpad, frd_o - true, *args, c["ry"]), order, c["nx"], c["ny"])

    model = torchwi.Time2d(c["nx"], c["ny"], c["h"], torch.from_numpy(c["w"]), c["dt"], order, DEV)
    vel = torch.nn.Parameter(torch.from_numpy(c["vel"]).to(DEV))
    sx, sy, ry = (torch.tensor([c[k]]) for k in ("sx", "sy", "ry"))
    frd = model(vel, sx, sy, ry)
    assert tuple(frd.shape) == (c["nt"], c["nx"])
    loss = 0.5 * torch.sum((frd - torch.from_numpy(true).to(DEV)) ** 2)
    loss.backward()
    assert tuple(vel.grad.shape) == (c["nx"], c["ny"])
    assert np.array_equal(frd.detach().cpu().numpy(), frd_o)
    assert np.array_equal(vel.grad.cpu().numpy(), grad_o)


def test_time2dforward_module():
    import torchwi_amd as torchwi
    order = 4
    c = small_case(order, nx=50, ny=30, nt=100, seed=4)
    vpad = O.pad_vel(c["vel"], order)
    frd_o, _ = O.forward(vpad, c["w"], order, c["dimx"], c["dimy"], c["h"], c["dt"],
                         c["sx"], c["sy"], c["ry"], store_virt=False)
    m = torchwi.Time2dForward(c["nx"], c["ny"], c["h"], torch.from_numpy(c["w"]), c["dt"], order, DEV)
    with torch.no_grad():
        frd = m(torch.from_numpy(c["vel"]).to(DEV), torch.tensor([c["sx"]]), torch.tensor([c["sy"]]),
                torch.tensor([c["ry"]]))
    assert np.array_equal(frd.cpu().numpy(), frd_o)


@pytest.mark.parametrize("order", [2, 8])
def test_batch_matches_single_and_checkpoint_matches_full(order):
    """Batched lockstep shots equal one-shot calls bit for bit, and the
    checkpointed gradient equals the full-storage gradient bit for bit."""
    from torchwi_amd import Time2dBatch
    c = small_case(order, nx=100, ny=45, nt=230, seed=5)
    sx = [(i + 0.5) * c["h"] for i in (5, 33, 61, 94)]
    sy = [(2 + 0.5) * c["h"]] * 4
    ry = [(3 + 0.5) * c["h"]] * 4
    vel = torch.from_numpy(c["vel"]).to(DEV)
    w = torch.from_numpy(c["w"])
    full = Time2dBatch(c["nx"], c["ny"], c["h"], w, c["dt"], order, DEV, ckpt_every=0)
    d_obs = torch.zeros(4, c["nt"], c["nx"], device=DEV)
    frd_full = full.modeling(vel, sx, sy, ry).clone()
    d_obs.copy_(0.8 * frd_full)
    loss_f, g_full = full.misfit_gradient(vel, sx, sy, ry, d_obs)
    g_full = g_full.clone()
    per_shot_full = [full.plan.grad_of(s).clone() for s in range(4)]
    for K in (7, 64, 1000):
        ck = Time2dBatch(c["nx"], c["ny"], c["h"], w, c["dt"], order, DEV, ckpt_every=K)
        loss_c, g_ck = ck.misfit_gradient(vel, sx, sy, ry, d_obs)
        assert torch.equal(ck.plan.frd_all(), frd_full)
        assert torch.equal(g_ck, g_full), f"K={K}"
        assert float(loss_c) == float(loss_f)
    # each shot of the batch equals the one-shot oracle
    vpad = O.pad_vel(c["vel"], order)
    args = (order, c["dimx"], c["dimy"], c["h"], c["dt"])
    dimx = c["dimx"]
    for s in range(4):
        frd_o, virt_o = O.forward(vpad, c["w"], *args, sx[s], sy[s], ry[s])
        assert np.array_equal(frd_full[s].cpu().numpy(), frd_o)
        resid = frd_o - 0.8 * frd_o
        resid = (frd_full[s] - d_obs[s]).cpu().numpy()
        grad_o = O.backward(virt_o, vpad, resid, *args, ry[s])
        got = per_shot_full[s][:, :dimx].cpu().numpy().ravel()
        assert np.array_equal(got, grad_o)


def test_three_layer_100x200_forward_gradient_bitexact():
    """The reference's CPU-runnable case: 100x200 model behind a 20-cell pad,
    nt=1000, order 8, gradient of the L2 misfit against the blob model."""
    from torchwi_amd import workloads as W
    import torchwi_amd as torchwi
    c = W.three_layer_100x200()
    order, nx, ny, nt = c["order"], c["nx"], c["ny"], c["nt"]
    dimx, dimy = nx + order, ny + order
    vpad = O.pad_vel(c["vel"], order)
    vpad_t = O.pad_vel(c["vel_true"], order)
    a = (order, dimx, dimy, c["h"], c["dt"], c["sx"][0], c["sy"][0], c["ry"][0])
    d_obs, _ = O.forward(vpad_t, c["w"], *a, store_virt=False)
    frd_o, virt_o = O.forward(vpad, c["w"], *a)
    grad_o = O.crop_grad(O.backward(virt_o, vpad, frd_o - d_obs, order, dimx, dimy, c["h"], c["dt"],
                                    c["ry"][0]), order, nx, ny)
    frd64, virt64 = O.forward64(vpad, c["w"], *a)
    d64, _ = O.forward64(vpad_t, c["w"], *a, store_virt=False)
    grad64 = O.crop_grad(O.backward64(virt64, vpad, frd64 - d64, order, dimx, dimy, c["h"], c["dt"],
                                      c["ry"][0]), order, nx, ny)

    model = torchwi.Time2d(nx, ny, c["h"], torch.from_numpy(c["w"]), c["dt"], order, DEV)
    vel = torch.nn.Parameter(torch.from_numpy(c["vel"]).to(DEV))
    frd = model(vel, torch.tensor([c["sx"][0]]), torch.tensor([c["sy"][0]]), torch.tensor([c["ry"][0]]))
    loss = 0.5 * torch.sum((frd - torch.from_numpy(d_obs).to(DEV)) ** 2)
    loss.backward()
    f, g = frd.detach().cpu().numpy(), vel.grad.cpu().numpy()
    print(f"three-layer 100x200: rel-L2 vs oracle frd {rel_l2(f, frd_o):.2e} grad {rel_l2(g, grad_o):.2e}")
    record_drift("three-layer 100x200 + 20-cell pad, order 8, nt=1000", frd=rel_l2(f, frd64),
                 grad=rel_l2(g, grad64))
    assert np.array_equal(f, frd_o)
    assert np.array_equal(g, grad_o)


def test_invalid_arguments_raise():
    from torchwi_amd.propagator import td2d_cuda
    n = 20 * 20
    z = lambda k: torch.zeros(k, device=DEV)  # noqa: E731
    with pytest.raises(RuntimeError, match="order"):
        td2d_cuda.forward(z(100), z(100 * n), z(n), z(n), z(n), z(n) + 2000, torch.zeros(10),
                          6, 20, 20, 10, 10.0, 1e-3, 50.0, 20.0, 20.0)
    with pytest.raises(RuntimeError, match="source"):
        td2d_cuda.forward(z(10 * 12), z(10 * n), z(n), z(n), z(n), z(n) + 2000, torch.zeros(10),
                          8, 20, 20, 10, 10.0, 1e-3, 5000.0, 20.0, 20.0)
    with pytest.raises(RuntimeError, match="elements"):
        td2d_cuda.forward(z(5), z(10 * n), z(n), z(n), z(n), z(n) + 2000, torch.zeros(10),
                          8, 20, 20, 10, 10.0, 1e-3, 50.0, 20.0, 20.0)


@pytest.mark.parametrize("order,nx,ny,tile_rows,tile_cols", [
    (8, 150, 70, 32, 0), (8, 150, 70, 64, 0), (4, 131, 57, 32, 0), (2, 67, 29, 64, 0), (8, 61, 32, 0, 0),
    (8, 301, 70, 32, 128), (8, 150, 45, 16, 128), (4, 131, 57, 0, 128), (2, 67, 29, 16, 128),
    (8, 150, 70, 16, 64), (4, 77, 41, 16, 0), (8, 61, 32, 8, 128)])
def test_batch_kernel_shapes_bitexact(order, nx, ny, tile_rows, tile_cols):
    """The vectorised batched kernel over odd widths (dimx % 4 != 0), several
    tiles in x and y, every tile shape (64 x 16/32/64, 128 x 8/16/32):
    gathers and per-shot gradients equal the oracle bit for bit."""
    from torchwi_amd import Time2dBatch
    c = small_case(order, nx=nx, ny=ny, nt=140, seed=nx + ny)
    sx = [(i + 0.5) * c["h"] for i in (1, nx // 2, nx - 2)]
    sy = [(i + 0.5) * c["h"] for i in (0, 5, ny - 1)]
    ry = [(i + 0.5) * c["h"] for i in (0, 3, ny - 1)]
    print(f"case order={order} nx={nx} ny={ny} tile_rows={tile_rows} tile_cols={tile_cols}", flush=True)
    w = torch.from_numpy(c["w"])
    m = Time2dBatch(nx, ny, c["h"], w, c["dt"], order, DEV, ckpt_every=0, tile_rows=tile_rows,
                    tile_cols=tile_cols)
    vel = torch.from_numpy(c["vel"]).to(DEV)
    frd = m.modeling(vel, sx, sy, ry).clone()
    d_obs = torch.zeros_like(frd)
    m.misfit_gradient(vel, sx, sy, ry, d_obs)
    vpad = O.pad_vel(c["vel"], order)
    args = (order, c["dimx"], c["dimy"], c["h"], c["dt"])
    dimx = c["dimx"]
    for s in range(3):
        frd_o, virt_o = O.forward(vpad, c["w"], *args, sx[s], sy[s], ry[s])
        assert np.array_equal(frd[s].cpu().numpy(), frd_o), f"shot {s} gather"
        grad_o = O.backward(virt_o, vpad, frd_o, *args, ry[s])
        got = m.plan.grad_of(s)[:, :dimx].cpu().numpy().ravel()
        assert np.array_equal(got, grad_o), f"shot {s} gradient"
        pad = m.plan.grad_of(s)[:, dimx:]
        assert not torch.any(pad != 0), "row padding must stay zero"


@pytest.mark.parametrize("K", [0, 13])
def test_graph_replay_matches_direct_launches(K):
    """hipGraph-captured step blocks (partial last block / segment included)
    give the same bits as one direct launch per step."""
    from torchwi_amd import Time2dBatch
    order = 8
    c = small_case(order, nx=120, ny=50, nt=230, seed=21)
    sx = [(i + 0.5) * c["h"] for i in (10, 60, 110)]
    sy = [2.5 * c["h"]] * 3
    ry = [3.5 * c["h"]] * 3
    vel = torch.from_numpy(c["vel"]).to(DEV)
    w = torch.from_numpy(c["w"])
    out = []
    for gs in (0, 7, 100):
        print(f"case K={K} graph_steps={gs}", flush=True)
        m = Time2dBatch(c["nx"], c["ny"], c["h"], w, c["dt"], order, DEV, ckpt_every=K, graph_steps=gs)
        frd = m.modeling(vel, sx, sy, ry).clone()
        loss, g = m.misfit_gradient(vel, sx, sy, ry, 0.5 * frd)
        out.append((frd, g.clone(), float(loss)))
    for frd, g, loss in out[1:]:
        assert torch.equal(frd, out[0][0])
        assert torch.equal(g, out[0][1])
        assert loss == out[0][2]


@pytest.mark.parametrize("time_block,K", [(1, 0), (3, 12)])
def test_windows_match_oracle_on_cropped_models(time_block, K):
    """Shots on uneven aperture windows of one model, in one plan: each gather
    and gradient equals the oracle on the cropped sub-model (the reference's
    Time2d on that sub-model), and the full gradient is their windowed sum;
    with one step per launch and with the blocked kernel (3 steps) over
    checkpoints."""
    from torchwi_amd import Time2dWindows
    order, nx, ny, nt, h, dt = 8, 260, 40, 150, 10.0, 1e-3
    rng = np.random.default_rng(31)
    vel = (1500 + 2000 * rng.random((nx, ny))).astype(np.float32)
    t = np.arange(nt) * dt
    w = ((1 - 2 * (np.pi * 15 * (t - 0.05)) ** 2) * np.exp(-(np.pi * 15 * (t - 0.05)) ** 2)).astype(np.float32)
    wins = []
    for (x0, x1) in ((0, 97), (40, 201), (150, 260), (61, 62 + 70)):
        nxw = x1 - x0
        wins.append(dict(x0=x0, x1=x1, sx=(nxw // 2 + 0.5) * h, sy=2.5 * h, ry=2.5 * h,
                         rx0=5, nrec=nxw - 10))
    print(f"case windows time_block={time_block} K={K}", flush=True)
    m = Time2dWindows(nx, ny, h, torch.from_numpy(w), dt, order, wins, DEV, ckpt_every=K, time_block=time_block)
    velt = torch.from_numpy(vel).to(DEV)
    frds = [f.clone() for f in m.modeling(velt)]
    d_obs = [0.7 * f for f in frds]
    loss, g = m.misfit_gradient(velt, d_obs)
    want = np.zeros((nx, ny), np.float64)
    for s, q in enumerate(wins):
        sub = vel[q["x0"]:q["x1"], :]
        nxw = q["x1"] - q["x0"]
        vpad = O.pad_vel(sub, order)
        a = (order, nxw + order, ny + order, h, dt)
        frd_o, virt_o = O.forward(vpad, w, *a, q["sx"], q["sy"], q["ry"])
        win = frd_o[:, q["rx0"]:q["rx0"] + q["nrec"]]
        assert np.array_equal(frds[s].cpu().numpy(), win), f"window {s}"
        resid = np.zeros_like(frd_o)
        resid[:, q["rx0"]:q["rx0"] + q["nrec"]] = win - 0.7 * win
        # the residual the GPU formed is frd - 0.7 frd in float32
        resid[:, q["rx0"]:q["rx0"] + q["nrec"]] = (frds[s] - d_obs[s]).cpu().numpy()
        g_o = O.crop_grad(O.backward(virt_o, vpad, resid, *a, q["ry"]), order, nxw, ny)
        got = m.plan.grad_of(s)[order // 2:order // 2 + ny, order // 2:order // 2 + nxw].cpu().numpy().T
        assert np.array_equal(got, g_o), f"window {s} gradient"
        want[q["x0"]:q["x1"], :] += g_o
    assert rel_l2(g.cpu().numpy(), want) < 1e-6


@pytest.mark.parametrize("order,T,nx,ny,nt,K,rows", [
    (8, 3, 150, 70, 100, 0, 0), (8, 3, 150, 70, 101, 12, 0), (8, 2, 61, 130, 77, 8, 0),
    (8, 4, 200, 64, 90, 16, 0), (4, 3, 131, 57, 95, 9, 0), (2, 3, 67, 29, 80, 0, 0),
    (2, 4, 70, 150, 60, 8, 0), (8, 3, 20, 12, 40, 6, 0), (8, 4, 300, 200, 63, 0, 1000),
    (8, 4, 141, 203, 70, 8, 37), (8, 2, 250, 90, 51, 0, 16), (4, 4, 77, 120, 45, 12, 50),
    (8, 4, 601, 90, 41, 0, 0), (8, 3, 450, 100, 60, 12, 24), (4, 2, 513, 77, 50, 10, 20),
    (8, 4, 230, 70, 33, 8, 0)])
def test_temporal_blocking_bitexact(order, T, nx, ny, nt, K, rows):
    """The temporally blocked kernel (T steps per launch over column strips
    and row chunks with redundant halos) gives the same gathers and per-shot
    gradients as one step per launch, which equal the oracle; with
    checkpoints (K a multiple of T) too; rows: chunk height (0 automatic)."""
    from torchwi_amd import Time2dBatch
    c = small_case(order, nx=nx, ny=ny, nt=nt, seed=nx * 7 + T)
    sx = [(i + 0.5) * c["h"] for i in (1, nx // 2, nx - 2)]
    sy = [(i + 0.5) * c["h"] for i in (0, ny // 3, ny - 1)]
    ry = [(i + 0.5) * c["h"] for i in (1, 0, ny // 2)]
    print(f"case order={order} T={T} nx={nx} ny={ny} nt={nt} K={K} rows={rows}", flush=True)
    vel = torch.from_numpy(c["vel"]).to(DEV)
    w = torch.from_numpy(c["w"])
    res = []
    for tb in (1, T):
        m = Time2dBatch(nx, ny, c["h"], w, c["dt"], order, DEV, ckpt_every=K, time_block=tb,
                        tile_rows=rows if tb > 1 else 0)
        frd = m.modeling(vel, sx, sy, ry).clone()
        loss, g = m.misfit_gradient(vel, sx, sy, ry, 0.6 * frd)
        res.append((frd, g.clone(), [m.plan.grad_of(s).clone() for s in range(3)], float(loss)))
    assert torch.equal(res[0][0], res[1][0]), "gathers"
    for s in range(3):
        assert torch.equal(res[0][2][s], res[1][2][s]), f"shot {s} gradient"
    assert torch.equal(res[0][1], res[1][1])
    vpad = O.pad_vel(c["vel"], order)
    frd_o, _ = O.forward(vpad, c["w"], order, c["dimx"], c["dimy"], c["h"], c["dt"], sx[1], sy[1], ry[1],
                         store_virt=False)
    assert np.array_equal(res[1][0][1].cpu().numpy(), frd_o)


@pytest.mark.parametrize("streams,nt,K", [(1, 50, 6), (2, 50, 6), (2, 37, 9), (1, 20, 30)])
def test_blocked_backward_twice_after_one_forward(streams, nt, K):
    """A checkpointed blocked plan's forward pass stores the last segment's
    virtual sources in the segment buffer, and its backward adjoints that
    segment without recomputing it.  A second backward after the same
    forward finds the buffer holding another segment's and recomputes every
    segment: both give the per-shot gradients of one step per launch, bit for
    bit (K = 30 > nt: one segment, stored whole by the forward)."""
    from torchwi_amd import Time2dBatch
    order, nx, ny, S = 8, 120, 60, 3
    c = small_case(order, nx=nx, ny=ny, nt=nt, seed=nt + K)
    sx = [(i + 0.5) * c["h"] for i in (2, nx // 2, nx - 3)]
    sy = [(i + 0.5) * c["h"] for i in (0, ny // 3, ny - 2)]
    ry = [(i + 0.5) * c["h"] for i in (1, 0, ny // 2)]
    print(f"case streams={streams} nt={nt} K={K}", flush=True)
    vel = torch.from_numpy(c["vel"]).to(DEV)
    w = torch.from_numpy(c["w"])
    ref = Time2dBatch(nx, ny, c["h"], w, c["dt"], order, DEV, ckpt_every=0, time_block=1)
    d_obs = 0.6 * ref.modeling(vel, sx, sy, ry).clone()
    ref.misfit_gradient(vel, sx, sy, ry, d_obs)
    want = [ref.plan.grad_of(s).clone() for s in range(S)]
    m = Time2dBatch(nx, ny, c["h"], w, c["dt"], order, DEV, ckpt_every=K, time_block=3, streams=streams)
    m.misfit_gradient(vel, sx, sy, ry, d_obs)
    first = [m.plan.grad_of(s).clone() for s in range(S)]
    m.plan.backward()
    second = [m.plan.grad_of(s).clone() for s in range(S)]
    for s in range(S):
        assert torch.equal(first[s], want[s]), f"shot {s}: first backward"
        assert torch.equal(second[s], want[s]), f"shot {s}: second backward"


@pytest.mark.skip(reason="grad_group > 1 is switched off: its first MI355X run (case 8-5-2-0-100) ended in a "
                         "GPU memory fault whose cause has not been found; the plan refuses grad_group > 1")
@pytest.mark.parametrize("order,S,G,K,graph", [
    (8, 5, 2, 0, 100), (8, 8, 4, 12, 100), (4, 3, 4, 0, 0), (2, 6, 3, 7, 0), (8, 7, 7, 9, 0)])
def test_grad_group_matches_per_shot(order, S, G, K, graph):
    """With grad_group = G the adjoint steps G shots of the model through each
    tile and accumulates their imaging condition into one gradient (the
    group's first shot's buffer; the others stay 0).  Gathers and loss are
    the per-shot plan's bit for bit; every group gradient equals the sum of
    its shots' per-shot gradients, and the batch gradient the per-shot batch
    gradient, to rounding (relative L2 <= 1e-6): only the order of the sum
    over shots and time differs."""
    from torchwi_amd import Time2dBatch
    c = small_case(order, nx=90, ny=50, nt=120, seed=S * 11 + G)
    h = c["h"]
    sx = [(5 + 80 * i / max(S - 1, 1) + 0.5) * h for i in range(S)]
    sy = [(i % 3 + 0.5) * h for i in range(S)]
    ry = [(1 + i % 2 + 0.5) * h for i in range(S)]
    print(f"case order={order} S={S} G={G} K={K} graph_steps={graph}", flush=True)
    vel = torch.from_numpy(c["vel"]).to(DEV)
    w = torch.from_numpy(c["w"])
    res = []
    for gg in (1, G):
        m = Time2dBatch(c["nx"], c["ny"], h, w, c["dt"], order, DEV, ckpt_every=K, time_block=1,
                        graph_steps=graph, grad_group=gg)
        frd = m.modeling(vel, sx, sy, ry).clone()
        loss, g = m.misfit_gradient(vel, sx, sy, ry, 0.6 * frd)
        res.append((frd, float(loss), g.clone(), [m.plan.grad_of(s).clone() for s in range(S)]))
    assert torch.equal(res[0][0], res[1][0]), "gathers"
    assert res[0][1] == res[1][1], "loss"
    assert rel_l2(res[1][2].cpu().numpy(), res[0][2].cpu().numpy()) < 1e-6
    for s in range(S):
        if s % G:
            assert not torch.any(res[1][3][s] != 0), f"shot {s} is not a group leader"
    for lead in range(0, S, G):
        want = res[0][3][lead].clone()
        for s in range(lead + 1, min(lead + G, S)):
            want.add_(res[0][3][s])
        assert rel_l2(res[1][3][lead].cpu().numpy(), want.cpu().numpy()) < 1e-6, f"group {lead}"


def test_grad_group_refused():
    """grad_group > 1 is refused with an error naming it (blocked kernel,
    multi-row tiles, and while the group adjoint is switched off, every plan);
    no kernel runs."""
    from torchwi_amd.shots import ShotPlan
    c = small_case(8, nx=40, ny=30, nt=30, seed=3)
    w = torch.from_numpy(c["w"])
    shots = [dict(model=0, sx=10.5 * c["h"], sy=2.5 * c["h"], ry=2.5 * c["h"], rx0=0, nrec=40, w=w)] * 2
    with pytest.raises(RuntimeError, match="grad_group"):
        ShotPlan(8, 30, c["h"], c["dt"], [(40, 30)], shots, 0, DEV, time_block=2, grad_group=2)
    with pytest.raises(RuntimeError, match="grad_group"):
        ShotPlan(8, 30, c["h"], c["dt"], [(40, 30)], shots, 0, DEV, tile_rows=32, grad_group=2)
    with pytest.raises(RuntimeError, match="grad_group"):
        ShotPlan(8, 30, c["h"], c["dt"], [(40, 30)], shots, 0, DEV, grad_group=2)


def test_forward_only_plan_and_autotune():
    """A forward-only plan (no virtual-source storage) models the same gathers
    as a gradient plan and refuses the adjoint; the autotuner returns one of
    its candidates with a timing for each."""
    from torchwi_amd import Time2dBatch
    from torchwi_amd.autotune import CANDIDATES, autotune
    from torchwi_amd.shots import ShotPlan
    order = 8
    c = small_case(order, nx=90, ny=40, nt=120, seed=41)
    w = torch.from_numpy(c["w"])
    vel = torch.from_numpy(c["vel"]).to(DEV)
    shots = [dict(model=0, sx=(i + 0.5) * c["h"], sy=2.5 * c["h"], ry=2.5 * c["h"], rx0=0, nrec=c["nx"], w=w)
             for i in (10, 50)]
    p = ShotPlan(order, c["nt"], c["h"], c["dt"], [(c["nx"], c["ny"])], shots, 0, DEV, forward_only=True)
    p.models[0].load(vel)
    p.forward()
    m = Time2dBatch(c["nx"], c["ny"], c["h"], w, c["dt"], order, DEV, ckpt_every=0)
    frd = m.modeling(vel, [q["sx"] for q in shots], [q["sy"] for q in shots], [q["ry"] for q in shots])
    assert torch.equal(p.frd_all(), frd)
    with pytest.raises(RuntimeError, match="forward_only"):
        p.backward()
    best, times = autotune(order, c["h"], c["dt"], [(c["nx"], c["ny"])],
                           [dict(q, w=None) for q in shots], DEV, nt_probe=30)
    assert best in [dict(x) for x in CANDIDATES]
    assert len(times) == len(CANDIDATES)


@pytest.mark.parametrize("order,S,nt,streams", [(8, 5, 40, 0), (4, 3, 31, 1), (2, 4, 29, 2)])
def test_forward_only_blocked_plan_bitexact(order, S, nt, streams):
    """A forward-only modelling plan with the blocked kernel (3 steps per
    launch, no virtual sources, no checkpoints; one or two shot groups) gives
    the gathers of the one-step forward-only plan bit for bit, and those equal
    the oracle's."""
    from torchwi_amd.shots import ShotPlan
    c = small_case(order, nx=150, ny=60, nt=nt, seed=S * 5 + order)
    h = c["h"]
    w = torch.from_numpy(c["w"])
    vel = torch.from_numpy(c["vel"]).to(DEV)
    shots = [dict(model=0, sx=(9 + 31 * s + 0.5) * h, sy=(s % 3 + 0.5) * h, ry=(1 + s % 2 + 0.5) * h, rx0=0,
                  nrec=c["nx"], w=w) for s in range(S)]
    print(f"case forward-only blocked order={order} S={S} nt={nt} streams={streams}", flush=True)
    res = []
    for tb in (3, 1):
        p = ShotPlan(order, nt, h, c["dt"], [(c["nx"], c["ny"])], shots, 0, DEV, forward_only=True, time_block=tb,
                     streams=streams)
        p.models[0].load(vel)
        p.forward()
        res.append(p.frd_all().clone())
        del p
    assert torch.equal(res[0], res[1]), "blocked vs one-step gathers"
    vpad = O.pad_vel(c["vel"], order)
    frd_o, _ = O.forward(vpad, c["w"], order, c["dimx"], c["dimy"], h, c["dt"], shots[1]["sx"], shots[1]["sy"],
                         shots[1]["ry"], store_virt=False)
    assert np.array_equal(res[0][1].cpu().numpy(), frd_o)


def test_auto_time_block_large_plan_bitexact():
    """A plan above shots.AUTO_BLOCK_CELLS cells per step runs the blocked
    kernel by default (3 steps per launch, many strips and chunks per shot)
    and gives the gathers and per-shot gradients of one step per launch."""
    from torchwi_amd import Time2dBatch
    from torchwi_amd.shots import AUTO_BLOCK_CELLS
    order, nx, ny, nt, h, dt, S = 8, 1024, 912, 24, 10.0, 1e-3, 80
    assert S * (nx + order) * (ny + order) >= AUTO_BLOCK_CELLS
    rng = np.random.default_rng(5)
    vel = torch.from_numpy((1500 + 2500 * rng.random((nx, ny))).astype(np.float32)).to(DEV)
    t = np.arange(nt) * dt
    w = torch.from_numpy(((1 - 2 * (np.pi * 25 * (t - 0.01)) ** 2) *
                          np.exp(-(np.pi * 25 * (t - 0.01)) ** 2)).astype(np.float32))
    sx = [(13 * s % nx + 0.5) * h for s in range(S)]
    sy = [(s % 7 + 0.5) * h for s in range(S)]
    ry = [(s % 5 + 0.5) * h for s in range(S)]
    print(f"case auto time block S={S} nx={nx} ny={ny} nt={nt}", flush=True)
    res = []
    for tb in ("auto", 1):
        m = Time2dBatch(nx, ny, h, w, dt, order, DEV, ckpt_every=0, time_block=tb)
        frd = m.modeling(vel, sx, sy, ry).clone()
        m.misfit_gradient(vel, sx, sy, ry, 0.5 * frd)
        res.append((m.plan.time_block, frd, [m.plan.grad_of(s).clone() for s in (0, S // 2, S - 1)]))
        del m
        torch.cuda.empty_cache()
    assert res[0][0] == 3 and res[1][0] == 1
    assert torch.equal(res[0][1], res[1][1])
    for a, b in zip(res[0][2], res[1][2]):
        assert torch.equal(a, b)


@pytest.mark.parametrize("order,S,K,nt,windows,tb,graph", [
    (8, 5, 12, 40, False, 3, 100), (4, 4, 0, 30, False, 3, 100), (8, 6, 9, 28, True, 3, 100),
    (8, 5, 12, 40, False, 1, 100), (4, 3, 0, 130, False, 1, 100), (2, 4, 7, 30, False, 1, 0),
    (8, 6, 9, 28, True, 1, 10)])
def test_two_shot_groups_on_two_streams_bitexact(order, S, K, nt, windows, tb, graph):
    """Two shot groups on two streams (streams = 2; each group marching through
    its steps and checkpoint segments on its own stream), with the blocked
    kernel and with one step per launch (hipGraph blocks or direct launches),
    give the gathers and per-shot gradients of one stream bit for bit, and
    those equal the oracle's (shot 1)."""
    from torchwi_amd import Time2dBatch, Time2dWindows
    c = small_case(order, nx=150, ny=60, nt=nt, seed=S * 7 + K)
    h = c["h"]
    vel = torch.from_numpy(c["vel"]).to(DEV)
    w = torch.from_numpy(c["w"])
    sx = [(9 + 27 * s + 0.5) * h for s in range(S)]
    sy = [(s % 3 + 0.5) * h for s in range(S)]
    ry = [(1 + s % 2 + 0.5) * h for s in range(S)]
    print(f"case two shot groups order={order} S={S} K={K} nt={nt} windows={windows} tb={tb} graph={graph}",
          flush=True)
    res = []
    for streams in (2, 1):
        if windows:
            wins = [dict(x0=10 * s, x1=10 * s + 90, sx=(20 + s + 0.5) * h, sy=sy[s], ry=ry[s], rx0=3, nrec=80)
                    for s in range(S)]
            m = Time2dWindows(c["nx"], c["ny"], h, w, c["dt"], order, wins, DEV, ckpt_every=K, time_block=tb,
                              graph_steps=graph, streams=streams)
            frd = [f.clone() for f in m.modeling(vel)]
            m.misfit_gradient(vel, [0.5 * f for f in frd])
            frd = torch.stack(frd)
        else:
            m = Time2dBatch(c["nx"], c["ny"], h, w, c["dt"], order, DEV, ckpt_every=K, time_block=tb,
                            graph_steps=graph, streams=streams)
            frd = m.modeling(vel, sx, sy, ry).clone()
            m.misfit_gradient(vel, sx, sy, ry, 0.5 * frd)
        assert m.plan.ngroups == streams
        res.append((frd, [m.plan.grad_of(s).clone() for s in range(S)]))
        del m
    assert torch.equal(res[0][0], res[1][0]), "gathers"
    for s in range(S):
        assert torch.equal(res[0][1][s], res[1][1][s]), f"gradient of shot {s}"
    if not windows:
        vpad = O.pad_vel(c["vel"], order)
        a = (order, c["dimx"], c["dimy"], h, c["dt"])
        frd_o, virt_o = O.forward(vpad, c["w"], *a, sx[1], sy[1], ry[1])
        assert np.array_equal(res[0][0][1].cpu().numpy(), frd_o)
        g_o = O.backward(virt_o, vpad, 0.5 * frd_o, *a, ry[1]).reshape(c["dimy"], c["dimx"])
        assert np.array_equal(res[0][1][1][:, :c["dimx"]].cpu().numpy(), g_o)


def test_stream_copy():
    """The measurement copy (tw_stream_copy) copies exactly, including a
    size that is not a multiple of the workgroup, and refuses sizes and
    pointers that are not 16-byte multiples and overlapping buffers."""
    from torchwi_amd import _lib
    L = _lib.lib()
    s = torch.cuda.current_stream().cuda_stream
    for n in (1 << 20, 4004):
        a = torch.rand(n, device=DEV)
        b = torch.zeros_like(a)
        assert L.tw_stream_copy(b.data_ptr(), a.data_ptr(), 4 * n, s) == 0
        torch.cuda.synchronize()
        assert torch.equal(a, b)
    a = torch.rand(64, device=DEV)
    b = torch.zeros_like(a)
    assert L.tw_stream_copy(b.data_ptr(), a.data_ptr(), 12, s) == -1
    assert L.tw_stream_copy(b.data_ptr() + 4, a.data_ptr(), 16, s) == -1
    assert L.tw_stream_copy(a.data_ptr() + 16, a.data_ptr(), 64, s) == -1
    assert "overlap" in _lib.last_error()
    torch.cuda.synchronize()
    assert not torch.any(b != 0)


def test_backward_after_a_later_forward_raises():
    """The saved wavefield storage of a forward is the module's own buffer
    (as in the reference); a backward whose forward is no longer the
    module's latest raises instead of returning a wrong gradient, and the
    latest forward's backward still works and equals the oracle."""
    import torchwi_amd as torchwi
    order = 8
    c = small_case(order, nx=40, ny=24, nt=60, seed=31)
    w = torch.from_numpy(c["w"])
    pos = [torch.tensor([c["sx"]]), torch.tensor([c["sy"]]), torch.tensor([c["ry"]])]
    m = torchwi.Time2d(c["nx"], c["ny"], c["h"], w, c["dt"], order, DEV)
    v = torch.nn.Parameter(torch.from_numpy(c["vel"]).to(DEV))
    f1 = m(v, *pos)
    l1 = f1.pow(2).sum()
    f2 = m(v, *pos)
    l2 = f2.pow(2).sum()
    with pytest.raises(RuntimeError, match="another forward"):
        l1.backward()
    l2.backward()
    vpad = O.pad_vel(c["vel"], order)
    args = (order, c["dimx"], c["dimy"], c["h"], c["dt"])
    frd_o, virt_o = O.forward(vpad, c["w"], *args, c["sx"], c["sy"], c["ry"])
    g_o = O.crop_grad(O.backward(virt_o, vpad, 2 * frd_o, *args, c["ry"]), order, c["nx"], c["ny"])
    assert np.array_equal(v.grad.cpu().numpy(), g_o)
    # the batched module keeps the same rule
    mb = torchwi.Time2dBatch(c["nx"], c["ny"], c["h"], w, c["dt"], order, DEV, ckpt_every=0)
    vb = torch.nn.Parameter(torch.from_numpy(c["vel"]).to(DEV))
    s3 = [c["sx"]], [c["sy"]], [c["ry"]]
    lb1 = mb(vb, *s3).pow(2).sum()
    lb2 = mb(vb, *s3).pow(2).sum()
    with pytest.raises(RuntimeError, match="another forward"):
        lb1.backward()
    lb2.backward()
    assert np.array_equal(vb.grad.cpu().numpy(), g_o)


@pytest.mark.parametrize("order,S,tb", [(8, 4, "auto"), (4, 3, 3), (2, 5, 1)])
def test_forward_batch_equals_time2dforward_per_shot(order, S, tb):
    """Time2dForwardBatch (a forward-only plan: one step per launch or the
    blocked kernel) gives every shot the gather of the drop-in
    Time2dForward module, bit for bit, and a second call with other
    positions re-plans."""
    import torchwi_amd as torchwi
    c = small_case(order, nx=120, ny=50, nt=90, seed=S * 3 + order)
    h = c["h"]
    w = torch.from_numpy(c["w"])
    vel = torch.from_numpy(c["vel"]).to(DEV)
    sx = [(5 + 25 * s + 0.5) * h for s in range(S)]
    sy = [(s % 3 + 0.5) * h for s in range(S)]
    ry = [(2 + s % 2 + 0.5) * h for s in range(S)]
    print(f"case forward batch order={order} S={S} time_block={tb}", flush=True)
    fb = torchwi.Time2dForwardBatch(c["nx"], c["ny"], h, w, c["dt"], order, DEV, time_block=tb)
    one = torchwi.Time2dForward(c["nx"], c["ny"], h, w, c["dt"], order, DEV)
    for shift in (0, 1):
        xs = [x + shift * h for x in sx]
        frd = fb(vel, xs, sy, ry).clone()
        assert frd.shape == (S, c["nt"], c["nx"])
        for s in range(S):
            g = one(vel, torch.tensor([xs[s]]), torch.tensor([sy[s]]), torch.tensor([ry[s]]))
            assert torch.equal(frd[s], g), f"shot {s} shift {shift}"
