"""Benchmark: batched forward + adjoint gradient of the time-domain 2D
acoustic propagator on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode auto|full|ckpt]

Workload (one "step" = one misfit gradient over every shot): the 384 x 1024
model at h = 12.5 m, dt = 1 ms, nt = 4000, order 8, 32 ricker(8) shots at
z = 2 with 1024 receivers, eight dipping layers 1500-4500 m/s with a +-5%
correlated perturbation; observed data are modelled from a second seeded
perturbation (untimed).  Inputs are resident in HBM before timing starts.

Throughput is Gcell-updates/s with one cell-update = one padded-grid cell of
one wavefield advanced one step for one shot; a gradient step is
2 * shots * dimx * dimy * nt cell-updates (checkpoint recomputation is not
credited).  Prints one JSON line.  Besides the headline (full storage, which
fits in HBM here), the line carries the same workload checkpointed, the plain
one-shot-per-launch form, and the checkpointed gradient of the 1000 x 3000
model with 64 shots (nt = 8000), where checkpointing is required.
"""
import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0            # MI355X HBM3E, AMD spec
BYTES_FWD = 16                   # u_prev, u_cur, model read; u_new write
BYTES_FWD_VIRT = 20              # + virtual-source write
BYTES_ADJ = 28                   # lambda x2 + model + virt + grad read; lambda + grad write
BYTES_GRAD_STEP = BYTES_FWD + BYTES_ADJ   # credited per cell*step*shot (checkpointed form)
METRIC = "Gcell-updates/s fwd+adj FP32 vs HBM roofline; gradient p50/p99 latency"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--mode", default="auto", choices=["auto", "full", "ckpt"])
    ap.add_argument("--ckpt-every", type=int, default=0, help="K for --mode ckpt (0: ceil(sqrt(nt)))")
    ap.add_argument("--nshots", type=int, default=32)
    ap.add_argument("--nt", type=int, default=4000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-timing", action="store_true", help="skip per-launch HIP-event timing")
    ap.add_argument("--time-block", type=int, default=1,
                    help="time steps per launch (2..4: temporally blocked kernel)")
    ap.add_argument("--tile-rows", type=int, default=0,
                    help="tile height (one step per launch: 32 or 64; blocked kernel: rows per chunk); 0 automatic")
    ap.add_argument("--grad-group", type=int, default=1,
                    help="shots per gradient group of the one-step adjoint (1: one gradient per shot; > 1 is refused by the library for now)")
    ap.add_argument("--tile-cols", type=int, default=0,
                    help="one-step kernel tile width: 64 or 128; 0 automatic")
    ap.add_argument("--graph-steps", type=int, default=100,
                    help="time steps per captured hipGraph block (0: one direct launch per step)")
    ap.add_argument("--compare", type=int, default=1,
                    help="also time the checkpointed form and the plain one-shot-per-launch form")
    ap.add_argument("--no-scale", action="store_true",
                    help="skip the checkpointed gradient of the 1000x3000 model with 64 shots (215 GB)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MB",
                    help="first time a device-to-device copy of MB megabytes (achievable-bandwidth "
                         "reference and PMC byte-counter calibration)")
    return ap.parse_args()


def copy_probe(mb, reps=10):
    """Achieved GB/s (read + write bytes) of a device copy of mb MiB: torch's
    copy, and the library's streaming copy (tw_stream_copy: one 16-byte load
    and store per thread), the achievable HBM rate of this device."""
    import torch
    from torchwi_amd import _lib
    n = mb * (1 << 20) // 4
    a = torch.rand(n, device="cuda:0")
    b = torch.empty_like(a)
    L = _lib.lib()

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    ms = timed(lambda: b.copy_(a))
    stream = torch.cuda.current_stream().cuda_stream
    ms_s = timed(lambda: _lib.check(L.tw_stream_copy(b.data_ptr(), a.data_ptr(), 4 * n, stream), "tw_stream_copy"))
    assert torch.equal(a, b)
    gbs = lambda t: round(2 * n * 4 / (t * 1e-3) / 1e9, 1)  # noqa: E731
    return {"bytes_per_copy": 2 * n * 4, "avg_ms": round(ms, 4), "GBs": gbs(ms),
            "stream_copy_avg_ms": round(ms_s, 4), "stream_copy_GBs": gbs(ms_s)}


def cpu_baseline():
    """Config-1 forward + gradient on the host cores: the reference's own CPU
    path when it was built (oracle/_ref), else the oracle's restatement."""
    from oracle import oracle as O
    from torchwi_amd import workloads as W
    cores = len(os.sched_getaffinity(0))
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))
    c = W.three_layer_100x200()
    order, nx, ny, nt, h, dt = c["order"], c["nx"], c["ny"], c["nt"], c["h"], c["dt"]
    dimx, dimy = nx + order, ny + order
    vpad = O.pad_vel(c["vel"], order)
    d_obs, _ = O.forward(O.pad_vel(c["vel_true"], order), c["w"], order, dimx, dimy, h, dt,
                         c["sx"][0], c["sy"][0], c["ry"][0], store_virt=False)
    impl = "ref" if O.ref_available() else "oracle"

    def run():
        frd, virt = O.forward(vpad, c["w"], order, dimx, dimy, h, dt, c["sx"][0], c["sy"][0],
                              c["ry"][0], impl=impl)
        O.backward(virt, vpad, frd - d_obs, order, dimx, dimy, h, dt, c["ry"][0], impl=impl)

    run()
    ts = []
    for _ in range(5):
        t0 = time.perf_counter()
        run()
        ts.append(time.perf_counter() - t0)
    t = statistics.median(ts)
    cells = 2.0 * dimx * dimy * nt
    return {"value": round(cells / t / 1e9, 4), "unit": "Gcell-updates/s", "cores": int(os.environ.get("OMP_NUM_THREADS", cores)),
            "kind": "reference" if impl == "ref" else "port",
            "sample": "100x200 model + 20-cell pad (248x128 padded, order 8), nt=1000, 1 shot, "
                      "forward with full storage + adjoint gradient; median of 5 runs (%.3f s each)" % t}


def run_batched(c, K, steps, warmup, timing, graph_steps, dev, time_block=1, tile_rows=0, tile_cols=0,
                grad_group=1):
    """Time `steps` misfit gradients of the batched engine; returns stats."""
    import torch
    from torchwi_amd import Time2dBatch
    order, nx, ny, nt, h, dt = c["order"], c["nx"], c["ny"], c["nt"], c["h"], c["dt"]
    S = len(c["sx"])
    model = Time2dBatch(nx, ny, h, torch.from_numpy(c["w"]), dt, order, dev, ckpt_every=K,
                        graph_steps=graph_steps, time_block=time_block, tile_rows=tile_rows,
                        tile_cols=tile_cols, grad_group=grad_group)
    vel = torch.from_numpy(c["vel"]).to(dev)
    rng = np.random.default_rng(7)
    vtrue = torch.from_numpy(c["vel"] * (1.0 + 0.02 * rng.standard_normal(c["vel"].shape)).astype(np.float32)).to(dev)
    d_obs = model.modeling(vtrue, c["sx"], c["sy"], c["ry"]).clone()
    st = model.plan
    mg = st.models[0]
    torch.cuda.synchronize()
    for _ in range(warmup):
        model.misfit_gradient(vel, c["sx"], c["sy"], c["ry"], d_obs)
    torch.cuda.synchronize()
    if timing:
        st.set_timing(True)
        st.timing(reset=True)
    lat = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ts = time.perf_counter()
        model.misfit_gradient(vel, c["sx"], c["sy"], c["ry"], d_obs)
        torch.cuda.synchronize()
        lat.append((time.perf_counter() - ts) * 1e3)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    cells_step = 2.0 * S * mg.dimx * mg.dimy * nt
    r = {"K": st.K, "S": S, "dimx": mg.dimx, "dimy": mg.dimy, "elapsed": elapsed, "lat": lat,
         "value": cells_step * steps / elapsed / 1e9, "shots_per_s": S * steps / elapsed}
    if timing:
        r["timing"] = st.timing()
        st.set_timing(False)
    del model, st
    torch.cuda.empty_cache()
    return r


def run_plain(c, nshots, dev):
    """The plain form: one shot at a time through the drop-in Time2d module
    (full storage, one launch per step), as the reference's loop calls it
    (inv_torch.py:128-137).  Returns gradient-shots/s."""
    import torch
    import torchwi_amd as torchwi
    order, nx, ny, nt, h, dt = c["order"], c["nx"], c["ny"], c["nt"], c["h"], c["dt"]
    model = torchwi.Time2d(nx, ny, h, torch.from_numpy(c["w"]), dt, order, "cuda")
    vel = torch.nn.Parameter(torch.from_numpy(c["vel"]).to(dev))
    true = torch.zeros(nt, nx, device=dev)

    def one(k):
        frd = model(vel, torch.tensor([c["sx"][k]]), torch.tensor([c["sy"][k]]), torch.tensor([c["ry"][k]]))
        loss = 0.5 * torch.sum((frd - true) ** 2)
        loss.backward()

    one(0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(nshots):
        one(k % len(c["sx"]))
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    del model
    torch.cuda.empty_cache()
    return {"gradient_shots_per_s": round(nshots / el, 3), "shots_timed": nshots,
            "s_per_shot": round(el / nshots, 4)}


def run_scale(dev):
    """Checkpointed misfit gradient of the 1000 x 3000 model (h = 10 m,
    dt = 0.8 ms, nt = 8000, order 8, 64 ricker(6) shots, 3000 receivers),
    where full storage (97 GB per shot) cannot be used: one untimed modelling
    of the observed data, then one timed gradient.  The step kernel is chosen
    by plan size as Time2dBatch does (the temporally blocked kernel here)."""
    import torch
    from torchwi_amd import Time2dBatch
    from torchwi_amd import workloads as W
    from torchwi_amd.shots import auto_time_block, default_ckpt_interval
    c = W.dipping_1000x3000(nshots=64, nt=8000)
    order, nx, ny, nt, h, dt = c["order"], c["nx"], c["ny"], c["nt"], c["h"], c["dt"]
    S = len(c["sx"])
    tb = auto_time_block(S * (nx + order) * (ny + order))
    K = default_ckpt_interval(nt, tb)
    m = Time2dBatch(nx, ny, h, torch.from_numpy(c["w"]), dt, order, dev, ckpt_every=K, time_block=tb)
    vel = torch.from_numpy(c["vel"]).to(dev)
    d_obs = m.modeling(vel * 1.01, c["sx"], c["sy"], c["ry"]).clone()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    m.misfit_gradient(vel, c["sx"], c["sy"], c["ry"], d_obs)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    cells = 2.0 * S * (nx + order) * (ny + order) * nt
    v = cells / t / 1e9
    del m, d_obs
    torch.cuda.empty_cache()
    return {"workload": "checkpointed misfit gradient, %d shots, %dx%d model (padded %dx%d), h=%g m, dt=%g s, "
                        "nt=%d, order %d, checkpoint every %d steps + recompute, %d steps per launch"
                        % (S, ny, nx, ny + order, nx + order, h, dt, nt, order, K, tb),
            "value": round(v, 3), "unit": "Gcell-updates/s", "s_per_gradient": round(t, 3),
            "gradient_shots_per_s": round(S / t, 3), "ckpt_every": K, "time_block": tb,
            "effective_hbm_frac": round(v * 1e9 * (BYTES_GRAD_STEP / 2.0) / (HBM_PEAK_GBS * 1e9), 4)}


def load_traffic(mode_key, kernel_key):
    """HBM bytes per launch of a step kernel at this workload, from the
    calibrated rocprofv3 FETCH_SIZE / WRITE_SIZE runs kept in profiles/
    (mode_key: "full_storage" or "checkpointed")."""
    path = os.path.join(ROOT, "profiles", "traffic_k_vstep_384x1024_32shots.json")
    try:
        with open(path) as f:
            return json.load(f)[mode_key].get(kernel_key)
    except (OSError, ValueError, KeyError):
        return None


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU (one MI355X); sharding shots over several GPUs is designed "
              "(one gradient all-reduce per batch) but not run (--gpus %d refused)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    import torch
    from torchwi_amd import workloads as W
    from torchwi_amd.shots import default_ckpt_interval

    dev = torch.device("cuda:0")
    probe = copy_probe(args.copy_probe) if args.copy_probe else None
    c = W.dipping_384x1024(nshots=args.nshots, nt=args.nt)
    order, nx, ny, nt, h, dt = c["order"], c["nx"], c["ny"], c["nt"], c["h"], c["dt"]
    tb = args.time_block
    kck = args.ckpt_every or default_ckpt_interval(nt, tb)
    K = {"full": 0, "ckpt": kck, "auto": "auto"}[args.mode]
    gg = args.grad_group if tb == 1 else 1
    r = run_batched(c, K, args.steps, args.warmup, not args.no_timing, args.graph_steps, dev, tb, args.tile_rows,
                    args.tile_cols, gg)
    S, kstate, value, elapsed, lat = r["S"], r["K"], r["value"], r["elapsed"], r["lat"]
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Gcell-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "batched misfit gradient (forward + adjoint), %d shots, %dx%d model "
                               "(padded %dx%d), h=%g m, dt=%g s, nt=%d, order %d, %s"
                               % (S, ny, nx, r["dimy"], r["dimx"], h, dt, nt, order,
                                  "full virtual-source storage" if kstate == 0 else "checkpoint every %d steps + recompute" % kstate),
                   "shots": S, "nz": ny, "nx": nx, "nt": nt, "order": order, "ckpt_every": kstate,
                   "graph_steps": args.graph_steps, "time_block": tb, "tile_rows": args.tile_rows,
                   "tile_cols": args.tile_cols, "grad_group": gg},
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3), "p99": round(float(np.percentile(lat, 99)), 3)},
        "gradient_shots_per_s": round(r["shots_per_s"], 3),
        "effective_hbm_frac": round(value * 1e9 * (BYTES_GRAD_STEP / 2.0) / (HBM_PEAK_GBS * 1e9), 4),
    }
    if not args.no_timing:
        tm = r["timing"]
        per = {}
        cells_launch = S * r["dimx"] * r["dimy"]
        for kind, b in (("fwd", BYTES_FWD), ("fwd_virt", BYTES_FWD_VIRT), ("recomp", BYTES_FWD_VIRT), ("adj", BYTES_ADJ)):
            ms, n = tm[kind]
            if n:
                avg = ms / n
                per[kind] = {"avg_us": round(avg * 1e3, 3), "launches": n,
                             "achieved_GBs": round(b * cells_launch / (avg * 1e-3) / 1e9, 1)}
        out["kernels"] = per
        dom = max(per, key=lambda k: per[k]["avg_us"] * per[k]["launches"])
        b = {"fwd": BYTES_FWD, "fwd_virt": BYTES_FWD_VIRT, "recomp": BYTES_FWD_VIRT, "adj": BYTES_ADJ}[dom]
        ach = per[dom]["achieved_GBs"]
        tr = (load_traffic("full_storage" if kstate == 0 else "checkpointed", dom)
              if (nt, S, order, tb, args.tile_rows, args.tile_cols, gg) == (4000, 32, 8, 1, 0, 0, 1) else None)
        kname = ("k_pipe<4,%s,%d>" % (dom, tb) if tb > 1 else
                 "k_vadj_group<4> (%d shots per group)" % gg if dom == "adj" and gg > 1 else "k_vstep<4,%s>" % dom)
        out["roofline"] = {"bound": "hbm", "kernel": kname,
                           "achieved": ach,
                           "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                           "traffic": tr, "bytes_per_cell": b, "cells_per_launch": cells_launch}
        if tr and probe:
            # measured HBM bytes of the launch over its duration, against the
            # streaming copy of the same device
            real = tr["bytes_per_cell"] * cells_launch / (per[dom]["avg_us"] * 1e-6) / 1e9
            out["roofline"]["traffic_GBs"] = round(real, 1)
            out["roofline"]["frac_of_stream_copy"] = round(real / probe["stream_copy_GBs"], 4)
    if args.compare:
        if kstate == 0:
            rc = run_batched(c, kck, 1, 0, False, args.graph_steps, dev, tb, args.tile_rows, args.tile_cols, gg)
            out["checkpointed"] = {"ckpt_every": rc["K"], "value": round(rc["value"], 3),
                                   "ms_per_step": round(rc["elapsed"] * 1e3, 3),
                                   "gradient_shots_per_s": round(rc["shots_per_s"], 3),
                                   "effective_hbm_frac": round(rc["value"] * 1e9 * (BYTES_GRAD_STEP / 2.0) / (HBM_PEAK_GBS * 1e9), 4)}
            ck_rate = rc["shots_per_s"]
        else:
            ck_rate = r["shots_per_s"]
        pl = run_plain(c, 2, dev)
        out["plain_one_shot"] = pl
        out["speedup_vs_plain"] = {"best": round(r["shots_per_s"] / pl["gradient_shots_per_s"], 2),
                                   "checkpointed": round(ck_rate / pl["gradient_shots_per_s"], 2)}
    if not args.no_scale:
        out["checkpointed_at_scale"] = run_scale(dev)
    if probe:
        out["copy_probe"] = probe
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
