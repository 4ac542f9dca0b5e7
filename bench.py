"""Benchmark: checkpointed batched forward + adjoint gradient of the
time-domain 2D acoustic propagator on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--nt NT] [--compare 0|1|LEGS] ...

Headline workload (one "step" = one misfit gradient over every shot): 64
ricker(6) shots on the 1000 x 3000 dipping-layer model (padded 1008 x 3008,
eight layers 1500-4500 m/s with a +-5% correlated perturbation), h = 10 m,
dt = 0.8 ms, nt = 8000, order 8, 3000 receivers at z = 2.  Full
virtual-source storage would need 97 GB per shot, so the gradient
checkpoints the wavefield every 90 steps and recomputes each segment; the
step kernel is the temporally blocked one (3 time steps per launch).
Observed data are modelled from a 1%-faster model (untimed).  Inputs are
resident in HBM before timing starts.

Throughput is Gcell-updates/s with one cell-update = one padded-grid cell of
one wavefield advanced one step for one shot; a gradient is
2 * shots * dimx * dimy * nt cell-updates (checkpoint recomputation is not
credited).  Prints one JSON line.  Besides the headline the line carries the
per-kernel roofline of the blocked kernel (against 8 TB/s and against the rate
a streaming copy of 1 GiB reaches on the same device), the 384 x 1024 x
32-shot gradient (full storage and checkpointed, and the plain
one-shot-per-launch form), the p50/p99 latency of a serving stream, and the
reference CPU path on the host cores.
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
METRIC = "Gcell-updates/s fwd+adj FP32 vs HBM roofline; gradient p50/p99 latency"
# Per-unit algorithmic bytes of the one-step kernels (one cell-update):
# forward u_prev, u_cur, model read + u_new write; + virtual-source write;
# adjoint lambda x2, model, virt, grad read + lambda, grad write.
STEP_BYTES = {"fwd": 16, "fwd_virt": 20, "recomp": 20, "adj": 28}
BYTES_GRAD_STEP = 16 + 28        # credited per cell*step*shot of a gradient (22 B per cell-update)


def pipe_bytes(T):
    """Compulsory HBM bytes per cell of one launch of the temporally blocked
    kernel advancing T steps (the intermediate levels stay on chip):
      forward        two fields in, model m, two fields out        = 20
      recomputation  + model c3, + T virtual sources written       = 24 + 4T
      adjoint        + T virtual sources read, gradient in and out = 28 + 4T"""
    return {"fwd": 20, "fwd_virt": 24 + 4 * T, "recomp": 24 + 4 * T, "adj": 28 + 4 * T}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nshots", type=int, default=64)
    ap.add_argument("--nt", type=int, default=8000)
    ap.add_argument("--ckpt-every", type=int, default=0, help="K (0: ceil(sqrt(nt)) rounded to the time block)")
    ap.add_argument("--time-block", type=int, default=0,
                    help="time steps per launch (1: one step; 2..4: blocked kernel; 0: automatic by size)")
    ap.add_argument("--tile-rows", type=int, default=0,
                    help="row-chunk height of the blocked kernel (0: automatic)")
    ap.add_argument("--streams", type=int, default=0,
                    help="blocked kernel: 1 one stream, 2 two shot groups on two streams, 0 automatic")
    ap.add_argument("--no-timing", action="store_true",
                    help="skip the extra gradient timed per launch (the kernels' roofline)")
    ap.add_argument("--compare", default="1",
                    help="also time the 384 x 1024 x 32-shot gradient: 1 every leg (full storage, checkpointed, "
                         "plain form), 0 none, or a comma list of full, ckpt, plain")
    ap.add_argument("--compare-nt", type=int, default=0, help="nt of the 384 x 1024 legs (0: the workload's 4000)")
    ap.add_argument("--compare-streams", type=int, default=0,
                    help="shot groups of the 384 x 1024 batched legs (0 automatic, 1 one stream, 2 two groups)")
    ap.add_argument("--headline", type=int, default=1,
                    help="0: skip the headline gradient (profiling runs of the other legs; value is then null)")
    ap.add_argument("--serving-requests", type=int, default=12,
                    help="requests of the serving-stream leg (0: skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=1024, metavar="MB",
                    help="first time a device copy of MB megabytes (the achievable-bandwidth reference the "
                         "roofline is also stated against, and the PMC byte-counter calibration; 0: skip)")
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
    out = {"bytes_per_copy": 2 * n * 4, "avg_ms": round(ms, 4), "GBs": gbs(ms),
           "stream_copy_avg_ms": round(ms_s, 4), "stream_copy_GBs": gbs(ms_s)}
    del a, b
    torch.cuda.empty_cache()
    return out


# time steps of the headline workload's shot 0 timed on the host cores: the reference's CPU path indexes
# its virtual sources with int (propagator_cpu.cpp:71,144), so nt * dimx * dimy must stay below 2^31
# (708 steps of the 3008 x 1008 padded grid)
CPU_SAMPLE_STEPS = 700


def _cpu_run(c, impl, nt_sample=None, d_obs_model=None):
    """One forward (full virtual-source storage) + adjoint gradient of shot 0
    of workload c on the host cores (the reference's own CPU path, impl
    "ref", or the oracle's restatement, "port"); returns (seconds per run,
    the median of five after one warm-up, credited cell-updates per run)."""
    from oracle import oracle as O
    order, nx, ny, h, dt = c["order"], c["nx"], c["ny"], c["h"], c["dt"]
    w = c["w"] if nt_sample is None else c["w"][:nt_sample]
    nt = len(w)
    dimx, dimy = nx + order, ny + order
    if impl == "ref" and nt * dimx * dimy >= 2 ** 31:
        raise ValueError("the reference's CPU path indexes virt with int: nt * dimx * dimy = %d >= 2^31"
                         % (nt * dimx * dimy))
    vpad = O.pad_vel(c["vel"], order)
    if d_obs_model is not None:
        d_obs, _ = O.forward(O.pad_vel(d_obs_model, order), w, order, dimx, dimy, h, dt,
                             c["sx"][0], c["sy"][0], c["ry"][0], store_virt=False)
    else:
        d_obs = np.zeros((nt, nx), np.float32)

    def run():
        frd, virt = O.forward(vpad, w, order, dimx, dimy, h, dt, c["sx"][0], c["sy"][0], c["ry"][0], impl=impl)
        O.backward(virt, vpad, frd - d_obs, order, dimx, dimy, h, dt, c["ry"][0], impl=impl)

    run()
    ts = []
    for _ in range(5):
        t0 = time.perf_counter()
        run()
        ts.append(time.perf_counter() - t0)
    return statistics.median(ts), 2.0 * dimx * dimy * nt


def cpu_baseline():
    """The reference's own CPU path (oracle/_ref, OpenMP on every host core;
    the oracle's restatement where it was not built) on a bounded sample of
    the headline workload: shot 0 of the 1000 x 3000 model, its first
    CPU_SAMPLE_STEPS time steps, forward with full virtual-source storage
    (the reference's CPU path has no checkpointing) + adjoint gradient.  Also
    the three-layer 100 x 200 model (20-cell pad, nt = 1000), whose grid fits
    in the host caches."""
    from oracle import oracle as O
    from torchwi_amd import workloads as W
    cores = len(os.sched_getaffinity(0))
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))
    impl = "ref" if O.ref_available() else "oracle"
    c = W.dipping_1000x3000(nshots=64)
    t, cells = _cpu_run(c, impl, nt_sample=CPU_SAMPLE_STEPS)
    small = W.three_layer_100x200()
    ts, cells_s = _cpu_run(small, impl, d_obs_model=small["vel_true"])
    return {"value": round(cells / t / 1e9, 4), "unit": "Gcell-updates/s",
            "cores": int(os.environ.get("OMP_NUM_THREADS", cores)),
            "kind": "reference" if impl == "ref" else "port",
            "sample": "shot 0 of the 1000x3000 model (1008x3008 padded, order 8), its first %d of 8000 time steps, "
                      "forward with full storage + adjoint gradient; median of 5 runs after one warm-up "
                      "(%.2f s each, about %.0f s of CPU work timed)" % (CPU_SAMPLE_STEPS, t, 5 * t),
            "small_model": {"value": round(cells_s / ts / 1e9, 4),
                            "sample": "100x200 model + 20-cell pad (248x128 padded, order 8), nt=1000, 1 shot, "
                                      "forward with full storage + adjoint gradient; median of 5 runs (%.3f s each)"
                                      % ts}}


def kernel_table(tm, cells, T):
    """Per step-kernel kind: average launch duration and algorithmic GB/s
    (the blocked kernel's compulsory bytes per launch, or the one-step
    kernels' per-unit bytes) from the plan's HIP-event timing."""
    per = {}
    nb = pipe_bytes(T) if T > 1 else STEP_BYTES
    for kind in ("fwd", "fwd_virt", "recomp", "adj"):
        ms, steps = tm[kind]
        if not steps:
            continue
        ms_step = ms / steps
        b = nb[kind]
        gbs = (b / T) * cells / (ms_step * 1e-3) / 1e9
        per[kind] = {"avg_us_per_launch": round(ms_step * T * 1e3, 2), "steps": steps,
                     "bytes_per_cell_per_launch": b, "achieved_GBs": round(gbs, 1),
                     "frac": round(gbs / HBM_PEAK_GBS, 4), "share_of_time": ms}
    tot = sum(v["share_of_time"] for v in per.values())
    for v in per.values():
        v["share_of_time"] = round(v["share_of_time"] / tot, 4)
    return per


def load_traffic(name):
    """Calibrated rocprofv3 FETCH_SIZE / WRITE_SIZE bytes per launch, kept in
    profiles/ (see torchwi_amd.profiling)."""
    try:
        with open(os.path.join(ROOT, "profiles", name)) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def run_scale(args, dev):
    """The headline: `steps` timed checkpointed misfit gradients of the
    1000 x 3000 model after `warmup` untimed ones; then, for the kernels'
    roofline, one more gradient with every shot on one stream and an event
    pair around each launch (the headline's two shot groups overlap their
    launches on two streams, so their per-launch durations would not describe
    one kernel's work)."""
    import torch
    from torchwi_amd import Time2dBatch
    from torchwi_amd import workloads as W
    from torchwi_amd.shots import auto_time_block, default_ckpt_interval
    c = W.dipping_1000x3000(nshots=args.nshots, nt=args.nt)
    order, nx, ny, nt, h, dt = c["order"], c["nx"], c["ny"], c["nt"], c["h"], c["dt"]
    S = len(c["sx"])
    dimx, dimy = nx + order, ny + order
    tb = args.time_block or auto_time_block(S * dimx * dimy)
    K = args.ckpt_every or default_ckpt_interval(nt, tb)
    mk = lambda streams: Time2dBatch(nx, ny, h, torch.from_numpy(c["w"]), dt, order, dev,  # noqa: E731
                                     ckpt_every=K, time_block=tb, tile_rows=args.tile_rows, streams=streams)
    m = mk(args.streams)
    vel = torch.from_numpy(c["vel"]).to(dev)
    d_obs = m.modeling(vel * 1.01, c["sx"], c["sy"], c["ry"]).clone()
    torch.cuda.synchronize()
    for k in range(args.warmup):
        m.misfit_gradient(vel, c["sx"], c["sy"], c["ry"], d_obs)
        torch.cuda.synchronize()
        print("bench: warm-up gradient %d done" % (k + 1), file=sys.stderr, flush=True)
    groups = m.plan.ngroups
    lat = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        ts = time.perf_counter()
        loss, g = m.misfit_gradient(vel, c["sx"], c["sy"], c["ry"], d_obs)
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - ts)
        print("bench: gradient %d: %.3f s" % (k + 1, lat[-1]), file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    assert np.isfinite(float(loss)) and bool(torch.isfinite(g).all())
    r = {"S": S, "K": m.plan.K, "T": m.plan.time_block, "dimx": dimx, "dimy": dimy, "elapsed": elapsed,
         "lat": lat, "groups": groups,
         "cells_step": 2.0 * S * dimx * dimy * nt, "h": h, "dt": dt, "nt": nt, "order": order, "nx": nx, "ny": ny}
    if not args.no_timing:
        if groups != 1:
            g_ref = g.clone()
            del m
            torch.cuda.empty_cache()
            m = mk(1)
            m.prepare(vel, c["sx"], c["sy"], c["ry"])
        st = m.plan
        st.set_timing(True)
        st.timing(reset=True)
        t1 = time.perf_counter()
        _, g1 = m.misfit_gradient(vel, c["sx"], c["sy"], c["ry"], d_obs)
        torch.cuda.synchronize()
        r["one_stream_s"] = time.perf_counter() - t1
        r["timing"] = st.timing()
        st.set_timing(False)
        if groups != 1:
            assert torch.equal(g1, g_ref), "one stream and two shot groups must give the same gradient"
        print("bench: one-stream gradient timed per launch: %.3f s" % r["one_stream_s"], file=sys.stderr, flush=True)
        del st
    del m, d_obs
    torch.cuda.empty_cache()
    return r


def run_batched(c, K, steps, warmup, dev, time_block=1, streams=0):
    """Time `steps` misfit gradients of the batched engine on workload c."""
    import torch
    from torchwi_amd import Time2dBatch
    order, nx, ny, nt, h, dt = c["order"], c["nx"], c["ny"], c["nt"], c["h"], c["dt"]
    S = len(c["sx"])
    model = Time2dBatch(nx, ny, h, torch.from_numpy(c["w"]), dt, order, dev, ckpt_every=K,
                        time_block=time_block, streams=streams)
    vel = torch.from_numpy(c["vel"]).to(dev)
    rng = np.random.default_rng(7)
    vtrue = torch.from_numpy(c["vel"] * (1.0 + 0.02 * rng.standard_normal(c["vel"].shape)).astype(np.float32)).to(dev)
    d_obs = model.modeling(vtrue, c["sx"], c["sy"], c["ry"]).clone()
    for _ in range(warmup):
        model.misfit_gradient(vel, c["sx"], c["sy"], c["ry"], d_obs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        model.misfit_gradient(vel, c["sx"], c["sy"], c["ry"], d_obs)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    mg = model.plan.models[0]
    cells = 2.0 * S * mg.dimx * mg.dimy * nt * steps
    r = {"ckpt_every": model.plan.K, "time_block": model.plan.time_block, "shot_groups": model.plan.ngroups,
         "value": round(cells / elapsed / 1e9, 3), "ms_per_gradient": round(elapsed / steps * 1e3, 3),
         "gradient_shots_per_s": round(S * steps / elapsed, 3),
         "effective_hbm_frac": round(cells / elapsed * (BYTES_GRAD_STEP / 2.0) / (HBM_PEAK_GBS * 1e9), 4)}
    del model, d_obs
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
    return {"gradient_shots_per_s": round(nshots / el, 3), "shots_timed": nshots, "s_per_shot": round(el / nshots, 4)}


def run_compare(dev, legs=("full", "ckpt", "plain"), nt=0, streams=0):
    """The 384 x 1024 model, 32 ricker(8) shots, nt = 4000: batched gradient
    with full storage (it fits: 212 GB) and checkpointed, against the plain
    one-shot-per-launch form."""
    from torchwi_amd import workloads as W
    from torchwi_amd.shots import default_ckpt_interval
    c = W.dipping_384x1024(nt=nt) if nt else W.dipping_384x1024()
    out = {"workload": "32 ricker(8) shots, 384x1024 model (padded 392x1032), h=12.5 m, dt=1 ms, nt=%d, order 8, "
                       "one step per launch in hipGraph blocks of 100 steps" % c["nt"]}
    if "full" in legs:
        out["full_storage"] = run_batched(c, 0, 2, 1, dev, streams=streams)
    if "ckpt" in legs:
        out["checkpointed"] = run_batched(c, default_ckpt_interval(c["nt"]), 1, 0, dev, streams=streams)
    if "plain" in legs:
        pl = out["plain_one_shot"] = run_plain(c, 2, dev)
        sp = {}
        for name, key in (("full_storage", "full_storage"), ("checkpointed", "checkpointed")):
            if key in out:
                sp[name] = round(out[key]["gradient_shots_per_s"] / pl["gradient_shots_per_s"], 2)
        out["speedup_vs_plain"] = sp
    return out


def run_serving(n, dev):
    """A short serving stream on the four serving grid classes (256 x 512 to
    1200 x 3600): n Poisson requests at
    80% of the measured back-to-back capacity; p50/p99 latency."""
    import torch
    from torchwi_amd import serving as SV
    kw = dict(seed=5, nt_range=(2000, 6000), shots_range=(1, 16))
    probe = SV.make_stream(n, 1.0, **kw)
    srv = SV.Server(dev, slots=4)
    t0 = time.perf_counter()
    srv.run(probe, realtime=False)
    capacity = n / (time.perf_counter() - t0)
    del srv, probe
    torch.cuda.empty_cache()
    rate = 0.8 * capacity
    reqs = SV.make_stream(n, rate, **kw)
    srv = SV.Server(dev, slots=4)
    t0 = time.perf_counter()
    res = srv.run(reqs, realtime=True)
    wall = time.perf_counter() - t0
    out = {"workload": "%d Poisson requests (seed 5) at %.3f req/s (80%% of the %.3f req/s back-to-back capacity), "
                       "grids %s, nt 2000-6000, 1-16 shots, 50/50 forward/gradient"
                       % (n, rate, capacity, [list(g) for g in SV.GRIDS])}
    out.update(SV.summarize(res, wall))
    out["max_concurrent"] = srv.max_concurrent
    del srv, reqs, res
    torch.cuda.empty_cache()
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU (one MI355X); sharding shots over several GPUs is designed "
              "(one gradient all-reduce per batch) but not run (--gpus %d refused)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    import torch

    dev = torch.device("cuda:0")
    probe = copy_probe(args.copy_probe) if args.copy_probe else None
    legs = {"1": ("full", "ckpt", "plain"), "0": ()}.get(args.compare, tuple(args.compare.split(",")))
    if any(x not in ("full", "ckpt", "plain") for x in legs):
        print("--compare takes 0, 1 or a comma list of full, ckpt, plain (got %r)" % args.compare, file=sys.stderr)
        sys.exit(2)
    if not args.headline:
        out = {"metric": METRIC, "value": None, "unit": "Gcell-updates/s", "n_gpus": 1,
               "note": "headline gradient skipped (--headline 0)"}
        if legs:
            out["batched_384x1024"] = run_compare(dev, legs, args.compare_nt, args.compare_streams)
        if args.serving_requests:
            out["serving"] = run_serving(args.serving_requests, dev)
        if probe:
            out["copy_probe"] = probe
        print(json.dumps(out), flush=True)
        return
    r = run_scale(args, dev)
    S, T, elapsed, lat = r["S"], r["T"], r["elapsed"], r["lat"]
    value = r["cells_step"] * args.steps / elapsed / 1e9
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Gcell-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "checkpointed batched misfit gradient (forward + adjoint), %d shots, %dx%d model "
                               "(padded %dx%d), h=%g m, dt=%g s, nt=%d, order %d, checkpoint every %d steps + "
                               "recompute, %d steps per launch"
                               % (S, r["ny"], r["nx"], r["dimy"], r["dimx"], r["h"], r["dt"], r["nt"], r["order"],
                                  r["K"], T),
                   "shots": S, "nz": r["ny"], "nx": r["nx"], "nt": r["nt"], "order": r["order"],
                   "ckpt_every": r["K"], "time_block": T, "tile_rows": args.tile_rows,
                   "shot_groups": r["groups"]},
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 1),
                       "p99": round(float(np.percentile(lat, 99)) * 1e3, 1)},
        "gradient_shots_per_s": round(S * args.steps / elapsed, 3),
        # credited updates at 22 B each (a one-step gradient's compulsory bytes) against 8 TB/s
        "effective_hbm_frac": round(value * 1e9 * (BYTES_GRAD_STEP / 2.0) / (HBM_PEAK_GBS * 1e9), 4),
    }
    if not args.no_timing:
        cells = S * r["dimx"] * r["dimy"]
        per = kernel_table(r["timing"], cells, T)
        out["kernels"] = per
        dom = max(per, key=lambda k: per[k]["share_of_time"])
        kname = ("k_pipe<4,%s,%d>" % (dom, T)) if T > 1 else "k_vstep<4,%s>" % dom
        out["roofline"] = {"bound": "hbm", "kernel": kname, "achieved": per[dom]["achieved_GBs"],
                           "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": per[dom]["frac"], "traffic": None,
                           "bytes_per_cell_per_launch": per[dom]["bytes_per_cell_per_launch"],
                           "cells_per_launch": cells, "steps_per_launch": T,
                           "measured_on": "one extra gradient, all shots on one stream, %.3f s" % r["one_stream_s"]}
        tr = load_traffic("traffic_k_pipe_1000x3000_64shots.json")
        if (tr and (S, r["nt"] >= 300, T, args.tile_rows) == (64, True, 3, 0)
                and dom in tr.get("kernels", {})):
            t = tr["kernels"][dom]
            out["roofline"]["traffic"] = t["total_bytes"]
            out["roofline"]["traffic_bytes_per_cell"] = t["bytes_per_cell"]
            out["roofline"]["traffic_GBs"] = round(t["total_bytes"] / (per[dom]["avg_us_per_launch"] * 1e-6) / 1e9, 1)
        if probe:
            # the same achieved rate against what a streaming copy reaches on this device
            out["roofline"]["stream_copy_GBs"] = probe["stream_copy_GBs"]
            out["roofline"]["frac_of_stream_copy"] = round(per[dom]["achieved_GBs"] / probe["stream_copy_GBs"], 4)
    if legs:
        out["batched_384x1024"] = run_compare(dev, legs, args.compare_nt, args.compare_streams)
    if args.serving_requests:
        out["serving"] = run_serving(args.serving_requests, dev)
    if probe:
        out["copy_probe"] = probe
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
