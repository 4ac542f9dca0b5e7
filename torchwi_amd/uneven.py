"""Batched uneven shots in one plan (256 receiver windows of a 600 x 4000 model), with tile-shape autotuning.

    python -m torchwi_amd.uneven [--shots 256] [--nt 6000]

256 shots on a 600 x 4000 model, each on its own aperture window (receiver
window of 400..2000 cells plus 40 cells each side), all marched by one plan
(in lockstep within each shot group), with the step kernel picked by
torchwi_amd.autotune: one step per launch replayed as hipGraph blocks of
100 steps in one of its tile shapes, or the temporally blocked kernel (3
steps per direct launch), whichever its probe measures faster.  Reports the
gradient's throughput and checks two shots against the same shots run alone
(bitwise).  Prints one JSON line.
"""
import argparse
import json
import sys
import time

import numpy as np
import torch

from . import Time2dWindows, workloads as W
from .autotune import autotune


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--shots", type=int, default=256)
    ap.add_argument("--nt", type=int, default=6000)
    ap.add_argument("--no-autotune", action="store_true")
    args = ap.parse_args(argv)
    dev = torch.device("cuda:0")
    c = W.uneven_600x4000(nshots=args.shots, nt=args.nt)
    order, nx, ny, nt, h, dt = c["order"], c["nx"], c["ny"], c["nt"], c["h"], c["dt"]
    wins = c["windows"]
    out = {"workload": "%d shots on windows of a %dx%d model, h=%g m, dt=%g s, nt=%d, order %d, "
                       "receiver windows %d..%d cells" % (len(wins), ny, nx, h, dt, nt, order,
                                                          min(q["nrec"] for q in wins), max(q["nrec"] for q in wins))}
    cfg = dict(tile_cols=0, tile_rows=0, time_block=1)
    if not args.no_autotune:
        models = [(q["x1"] - q["x0"], ny) for q in wins]
        shots = [dict(model=i, sx=q["sx"], sy=q["sy"], ry=q["ry"], rx0=q["rx0"], nrec=q["nrec"])
                 for i, q in enumerate(wins)]
        cfg, times = autotune(order, h, dt, models, shots, dev)
        out["autotune"] = {"picked": cfg, "probe_s": times}
        print("uneven: autotune picked %s (%s)" % (cfg, times), file=sys.stderr, flush=True)
    w = torch.from_numpy(c["w"])
    vel = torch.from_numpy(c["vel"]).to(dev)
    rng = np.random.default_rng(1)
    vtrue = vel * (1.0 + 0.02 * torch.from_numpy(rng.standard_normal(c["vel"].shape).astype(np.float32)).to(dev))
    m = Time2dWindows(nx, ny, h, w, dt, order, wins, dev, **cfg)
    t0 = time.perf_counter()
    d_obs = [f.clone() for f in m.modeling(vtrue)]
    torch.cuda.synchronize()
    out["forward_s"] = round(time.perf_counter() - t0, 3)
    print("uneven: observed data modelled (%.2f s)" % out["forward_s"], file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    loss, g = m.misfit_gradient(vel, d_obs)
    torch.cuda.synchronize()
    tg = time.perf_counter() - t0
    cells = sum(2.0 * (q["x1"] - q["x0"] + order) * (ny + order) * nt for q in wins)
    out.update(ckpt_every=m.plan.K, time_block=m.plan.time_block, gradient_s=round(tg, 3),
               gradient_Gcell_updates_per_s=round(cells / tg / 1e9, 2), loss=float(loss))
    print("uneven: gradient done (%.2f s)" % tg, file=sys.stderr, flush=True)
    # two shots alone, same configuration: bitwise equal to their part of the batch
    same = []
    for s in (0, len(wins) // 2):
        g_batch = m.plan.grad_of(s).clone()
        one = Time2dWindows(nx, ny, h, w, dt, order, [wins[s]], dev, ckpt_every=m.plan.K, **cfg)
        one.modeling(vtrue)
        one.misfit_gradient(vel, [d_obs[s]])
        same.append(bool(torch.equal(one.plan.grad_of(0), g_batch)))
        del one
    out["alone_equals_batched"] = same
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
