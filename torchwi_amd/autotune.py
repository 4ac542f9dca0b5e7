"""Pick the step-kernel configuration of a shot plan by timing candidates.

The candidates are tile shapes of the one-step kernel (replayed as hipGraph
blocks) and the temporally blocked kernel (several steps per launch).  Each candidate runs one misfit gradient
(forward + adjoint) on a proxy plan with the same shots, grids and order but
a short time axis, so the choice reflects the real launch shapes at a small
fraction of the cost.  All candidates compute bit-identical results; only
speed differs.
"""
import time

import torch

from .shots import ShotPlan, default_ckpt_interval

CANDIDATES = (
    dict(tile_cols=64, tile_rows=16, time_block=1),
    dict(tile_cols=64, tile_rows=32, time_block=1),
    dict(tile_cols=128, tile_rows=8, time_block=1),
    dict(tile_cols=0, tile_rows=0, time_block=3),
)


def autotune(order, h, dt, models, shots, device, nt_probe=96, candidates=CANDIDATES, repeats=2):
    """Return (best candidate dict, {label: seconds per probe gradient})."""
    w = torch.zeros(nt_probe, dtype=torch.float32)
    w[5] = 1.0
    probe_shots = [dict(q, w=w) for q in shots]
    times = {}
    best, best_t = None, None
    for cand in candidates:
        K = default_ckpt_interval(nt_probe, cand["time_block"])
        p = ShotPlan(order, nt_probe, h, dt, [(nx, ny) for nx, ny in models], probe_shots, K, device,
                     tile_rows=cand["tile_rows"], time_block=cand["time_block"],
                     tile_cols=cand["tile_cols"])
        for i in range(len(models)):
            p.models[i].vel.fill_(2000.0)
        p.forward()
        p.backward()
        torch.cuda.synchronize(device)
        t0 = time.perf_counter()
        for _ in range(repeats):
            p.forward()
            p.backward()
        torch.cuda.synchronize(device)
        dt_s = (time.perf_counter() - t0) / repeats
        label = "tile=%dx%d,time_block=%d" % (cand["tile_cols"], cand["tile_rows"], cand["time_block"])
        times[label] = round(dt_s, 5)
        if best_t is None or dt_s < best_t:
            best, best_t = cand, dt_s
        del p
        torch.cuda.empty_cache()
    return dict(best), times
