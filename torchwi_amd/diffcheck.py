"""Gradient verification at scale: checkpointed vs full storage, and central
finite differences along seeded smooth perturbation directions.

    python -m torchwi_amd.diffcheck [--shots 64] [--nt 8000] [--dirs 5] [--eps 1e-3]

Workload (defaults): the 1000 x 3000 dipping-layer model at h = 10 m,
dt = 0.8 ms, nt = 8000, order 8, 64 ricker(6) shots at z = 2, 3000 receivers;
observed data from a 2% smooth perturbation of the model.  Steps:

1. the gradient of shot 0 with full virtual-source storage (97 GB);
2. the gradient of all shots with checkpoints every ceil(sqrt(nt)) steps and
   recomputation; its shot-0 part must equal step 1 bit for bit;
3. for each direction d_k (seed 10+k: white noise, Gaussian sigma = 20
   cells, unit max-norm, zero within 10 cells of the left/right/bottom
   edges) the central difference (J(v(1+eps d)) - J(v(1-eps d))) / (2 eps),
   J = 0.5 ||d(v) - d_obs||^2 accumulated in FP64, against <g, v d>.

FP32 forward modelling limits the finite-difference agreement to about
1e-3 at eps = 1e-3 (rounding noise over a 1e-3 perturbation); the exactness
of the discrete gradient itself is pinned in FP64 by the test suite.
Prints one JSON line.
"""
import argparse
import json
import sys
import time

import numpy as np
import torch
from scipy import ndimage

from . import Time2dBatch
from . import workloads as W
from .shots import auto_time_block, default_ckpt_interval


def direction(nx, ny, seed, sigma=20.0, margin=10):
    """Smooth random direction in the reference layout (nx, ny)."""
    rng = np.random.default_rng(seed)
    d = ndimage.gaussian_filter(rng.standard_normal((ny, nx)), sigma, mode="wrap")
    d /= np.abs(d).max()
    d[:, :margin] = 0
    d[:, nx - margin:] = 0
    d[ny - margin:, :] = 0
    return np.ascontiguousarray(d.T.astype(np.float32))


def progress(msg):
    print("diffcheck: " + msg, file=sys.stderr, flush=True)


def misfit64(frd, d_obs):
    J = 0.0
    for s in range(frd.shape[0]):
        r = (frd[s] - d_obs[s]).double()
        J += 0.5 * float(torch.dot(r.view(-1), r.view(-1)))
    return J


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--shots", type=int, default=64)
    ap.add_argument("--nt", type=int, default=8000)
    ap.add_argument("--nz", type=int, default=1000)
    ap.add_argument("--nx", type=int, default=3000)
    ap.add_argument("--dirs", type=int, default=5)
    ap.add_argument("--eps", type=float, default=1e-3)
    ap.add_argument("--ckpt-every", type=int, default=0)
    ap.add_argument("--no-full", action="store_true", help="skip the full-storage comparison")
    ap.add_argument("--time-block", type=int, default=0,
                    help="steps per launch (1: one step; 2..4: temporally blocked kernel; 0: automatic by size)")
    ap.add_argument("--repeat", type=int, default=1, help="time this many gradients in a row")
    ap.add_argument("--tile-rows", type=int, default=0)
    ap.add_argument("--tile-cols", type=int, default=0)
    args = ap.parse_args(argv)
    dev = torch.device("cuda:0")
    c = W.dipping_1000x3000(nshots=args.shots, nt=args.nt)
    if (args.nz, args.nx) != (1000, 3000):
        v = W.dipping_layers(args.nz, args.nx, seed=2)
        xs = (np.arange(args.shots) + 0.5) * args.nx / args.shots
        c.update(nx=args.nx, ny=args.nz, vel=W.to_user(v),
                 sx=[float(W.cell_pos(int(x), c["h"])) for x in xs])
    order, nx, ny, nt, h, dt = c["order"], c["nx"], c["ny"], c["nt"], c["h"], c["dt"]
    tb = args.time_block or auto_time_block(args.shots * (nx + order) * (ny + order))
    K = args.ckpt_every or default_ckpt_interval(nt, tb)
    kcfg = dict(time_block=tb, tile_rows=args.tile_rows, tile_cols=args.tile_cols)
    w = torch.from_numpy(c["w"])
    vel = torch.from_numpy(c["vel"]).to(dev)
    pert = direction(nx, ny, seed=99, sigma=30.0)
    vtrue = vel * (1.0 + 0.02 * torch.from_numpy(pert).to(dev))
    out = {"workload": "%dx%d model, h=%g m, dt=%g s, nt=%d, order %d, %d ricker(6) shots, %d receivers"
                       % (ny, nx, h, dt, nt, order, args.shots, nx), "ckpt_every": K, "kernel": kcfg}

    # 1. shot 0 with full storage (before the batch plan: both would not fit)
    g0_full = d0 = None
    if not args.no_full:
        mf = Time2dBatch(nx, ny, h, w, dt, order, dev, ckpt_every=0, **kcfg)
        one = (c["sx"][:1], c["sy"][:1], c["ry"][:1])
        d0 = mf.modeling(vtrue, *one).clone()
        t0 = time.perf_counter()
        mf.misfit_gradient(vel, *one, d0)
        torch.cuda.synchronize()
        out["full_storage_1shot_s"] = round(time.perf_counter() - t0, 3)
        g0_full = mf.plan.grad_of(0).clone()
        progress("full-storage shot-0 gradient done (%.2f s)" % out["full_storage_1shot_s"])
        out["full_storage_bytes_per_shot"] = int(mf.plan.virt.numel() * 4)
        del mf
        torch.cuda.empty_cache()

    m = Time2dBatch(nx, ny, h, w, dt, order, dev, ckpt_every=K, **kcfg)
    t0 = time.perf_counter()
    d_obs = m.modeling(vtrue, c["sx"], c["sy"], c["ry"]).clone()
    torch.cuda.synchronize()
    out["forward_s"] = round(time.perf_counter() - t0, 3)
    progress("observed data modelled (%.2f s)" % out["forward_s"])
    if d0 is not None:
        out["batched_gather_equals_single_shot0"] = bool(torch.equal(d_obs[0], d0[0]))

    times = []
    for _ in range(max(1, args.repeat)):
        t0 = time.perf_counter()
        loss, g = m.misfit_gradient(vel, c["sx"], c["sy"], c["ry"], d_obs)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        progress("gradient %d: %.3f s" % (len(times), times[-1]))
    tg = min(times)
    cells = 2.0 * args.shots * (nx + order) * (ny + order) * nt
    out.update(gradient_s=round(tg, 3), gradient_runs_s=[round(t, 3) for t in times],
               gradient_Gcell_updates_per_s=round(cells / tg / 1e9, 2),
               checkpoint_plan_GB=round(sum(t.numel() for t in (m.plan.u, m.plan.lam, m.plan.virt, m.plan.ckpt,
                                                                   m.plan.grad)) * 4 / 1e9, 2))
    progress("checkpointed gradient of %d shots done (%.2f s)" % (args.shots, tg))
    if g0_full is not None:
        out["checkpointed_equals_full_storage_shot0"] = bool(torch.equal(m.plan.grad_of(0), g0_full))
    g = g.clone()
    frd_dirs = []
    for k in range(args.dirs):
        d = torch.from_numpy(direction(nx, ny, seed=10 + k)).to(dev)
        Js = []
        for sgn in (1.0, -1.0):
            vk = vel * (1.0 + sgn * args.eps * d)
            frd = m.modeling(vk, c["sx"], c["sy"], c["ry"])
            Js.append(misfit64(frd, d_obs))
        fd = (Js[0] - Js[1]) / (2 * args.eps)
        an = float(torch.sum((g * vel * d).double()))
        frd_dirs.append({"seed": 10 + k, "fd": fd, "adjoint": an, "rel_err": abs(fd - an) / abs(an)})
        progress("direction %d: fd %.6e adjoint %.6e rel %.2e" % (k, fd, an, frd_dirs[-1]["rel_err"]))
    out["fd_checks"] = frd_dirs
    out["max_fd_rel_err"] = max(x["rel_err"] for x in frd_dirs) if frd_dirs else None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
