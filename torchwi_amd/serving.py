"""Continuous-batching request loop for forward-modelling and misfit-gradient
jobs on one GPU.

Requests (a model, its shots, a wavelet, and -- for gradient jobs -- the
observed gathers) arrive over time.  The server keeps up to ``slots``
requests on the device at once, each on its own HIP stream so their time
steps overlap on the GPU, and admits a waiting request as soon as a slot and
enough of the HBM budget are free (footprint = its largest shot chunk's
plan plus its inputs and results).  Wavefields, checkpoints, gathers and
gradients stay device-resident for the whole request.

A request's shots run in chunks whose size depends only on the request (a
fixed per-request byte cap), never on what else is running, so each result
is bit-identical alone and in a mix.  A request whose smallest chunk cannot
fit the budget is rejected with an error instead of blocking the queue.

There is no counterpart in the reference, whose shot loop is one process per
GPU under Horovod (examples/time/inv_torch.py:123-143).
"""
import math
import sys
import time
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np
import torch

from .shots import GRAPH_STEPS, ShotPlan, auto_time_block, default_ckpt_interval, plan_bytes


@dataclass
class Request:
    """One job.  vel: (nx, ny) model (reference layout, depth fastest);
    w: (nt,) float32 wavelet; sx, sy, ry: per-shot physical positions;
    kind: "forward" or "gradient"; d_obs: (S, nt, nx) observed gathers
    (gradient jobs)."""
    kind: str
    vel: torch.Tensor
    w: torch.Tensor
    sx: List[float]
    sy: List[float]
    ry: List[float]
    h: float
    dt: float
    order: int = 8
    d_obs: Optional[torch.Tensor] = None
    arrival: float = 0.0
    rid: int = -1

    @property
    def nx(self):
        return int(self.vel.shape[0])

    @property
    def ny(self):
        return int(self.vel.shape[1])

    @property
    def nt(self):
        return int(self.w.numel())

    @property
    def nshots(self):
        return len(self.sx)


@dataclass
class Result:
    rid: int
    kind: str
    arrival: float
    admitted: float = math.nan
    finished: float = math.nan
    frd: Optional[torch.Tensor] = None     # forward: (S, nt, nx)
    loss: Optional[float] = None           # gradient
    grad: Optional[torch.Tensor] = None    # gradient: (nx, ny)
    error: Optional[str] = None
    chunks: int = 0
    cell_updates: float = 0.0              # credited work (2 passes for gradients)

    @property
    def latency(self):
        return self.finished - self.arrival


def chunk_plan(req, cap_bytes):
    """(shots per chunk, ckpt K, bytes of one chunk plan) for a request under
    a per-request cap -- a function of the request alone."""
    models = [(req.nx, req.ny, req.order)]
    shot = dict(model=0, nrec=req.nx)

    fwd = req.kind == "forward"

    def nbytes(n, K):
        return plan_bytes(models, [shot] * n, req.nt, K, forward_only=fwd)

    if req.kind == "gradient" and nbytes(req.nshots, 0) <= cap_bytes:
        # full storage when the whole request fits under the cap: less traffic
        return req.nshots, 0, nbytes(req.nshots, 0)
    # a multiple of the blocked kernel's 3 steps, so either step kernel can run the chunk
    K = 0 if req.kind == "forward" else default_ckpt_interval(req.nt, 3)
    base, one = nbytes(0, K), nbytes(1, K) - nbytes(0, K)
    per = max(1, min(req.nshots, int((cap_bytes - base) // max(one, 1))))
    return per, K, nbytes(per, K)


def chunk_time_block(req, nshots, K):
    """Steps per launch for a chunk of nshots shots: gradients by plan size
    (shots.auto_time_block), forward modelling one step per launch (measured
    faster at every size on MI355X, e.g. 12 shots of 3608 x 1208: 355 vs 307
    Gcell-updates/s).  Either gives the same bits."""
    if req.kind != "gradient" or K % 3:
        return 1
    return auto_time_block(nshots * (req.nx + req.order) * (req.ny + req.order))


def io_bytes(req):
    S, nt, nx = req.nshots, req.nt, req.nx
    b = 4 * (req.nx * req.ny + nt)
    if req.kind == "gradient":
        b += 4 * (S * nt * nx + 2 * nx * req.ny)
    else:
        b += 4 * S * nt * nx
    return b


class _Active:
    def __init__(self, req, res, stream, per, K, cost):
        self.req, self.res, self.stream = req, res, stream
        self.per, self.K, self.cost = per, K, cost
        self.next_shot = 0
        self.event = None
        self.plan = None
        self.loss = None
        self.grad = None
        self.frd = None


class Server:
    """Continuous-batching executor.  budget_bytes: HBM the server may use;
    slots: requests in flight at once (one HIP stream each, default priority);
    cap_bytes: per-request chunk cap (default budget/slots)."""

    def __init__(self, device="cuda", budget_bytes=None, slots=4, cap_bytes=None,
                 graph_steps=GRAPH_STEPS, verbose=False):
        self.verbose = verbose
        self.device = torch.device(device)
        if budget_bytes is None:
            free, _ = torch.cuda.mem_get_info(self.device)
            budget_bytes = int(0.7 * free)
        self.budget = int(budget_bytes)
        self.slots = int(slots)
        self.cap = int(cap_bytes) if cap_bytes else self.budget // self.slots
        self.graph_steps = graph_steps
        self.streams = [torch.cuda.Stream(self.device) for _ in range(self.slots)]
        self.in_use = 0
        self.free_slots = list(range(self.slots))
        self.waiting = []
        self.active = []
        self.done = []
        self.peak_in_use = 0
        self.max_concurrent = 0

    # ---- queue ------------------------------------------------------------
    def submit(self, req):
        res = Result(rid=req.rid, kind=req.kind, arrival=req.arrival)
        if req.kind not in ("forward", "gradient"):
            res.error = "unknown kind %r" % (req.kind,)
        elif req.kind == "gradient" and req.d_obs is None:
            res.error = "gradient request without observed data"
        else:
            per, K, chunk = chunk_plan(req, self.cap)
            cost = chunk + io_bytes(req)
            if cost > self.budget:
                res.error = "request needs %d bytes of HBM, budget is %d" % (cost, self.budget)
            else:
                self.waiting.append((req, res, per, K, cost))
                return res
        res.finished = res.admitted = time.perf_counter()
        self.done.append(res)
        return res

    # ---- execution ----------------------------------------------------------
    def _admit(self, now):
        while self.waiting and self.free_slots:
            req, res, per, K, cost = self.waiting[0]
            if req.arrival > now:
                break
            if self.in_use + cost > self.budget:
                break  # FIFO: wait for memory instead of starving large jobs
            self.waiting.pop(0)
            slot = self.free_slots.pop(0)
            self.in_use += cost
            self.peak_in_use = max(self.peak_in_use, self.in_use)
            res.admitted = time.perf_counter()
            a = _Active(req, res, self.streams[slot], per, K, cost)
            a.slot = slot
            self.active.append(a)
            self.max_concurrent = max(self.max_concurrent, len(self.active))
            with torch.cuda.stream(a.stream):
                dev = self.device
                a.vel = req.vel.to(dev, torch.float32, non_blocking=False)
                if req.kind == "gradient":
                    a.loss = torch.zeros((), dtype=torch.float64, device=dev)
                    a.grad = torch.zeros(req.ny + req.order, req.nx + req.order, dtype=torch.float32, device=dev)
                else:
                    a.frd = torch.empty(req.nshots, req.nt, req.nx, dtype=torch.float32, device=dev)
            self._launch_chunk(a)

    def _launch_chunk(self, a):
        req = a.req
        s0 = a.next_shot
        s1 = min(req.nshots, s0 + a.per)
        shots = [dict(model=0, sx=req.sx[i], sy=req.sy[i], ry=req.ry[i], rx0=0, nrec=req.nx, w=req.w)
                 for i in range(s0, s1)]
        with torch.cuda.stream(a.stream):
            a.plan = None
            p = ShotPlan(req.order, req.nt, req.h, req.dt, [(req.nx, req.ny)], shots, a.K, self.device,
                         graph_steps=self.graph_steps, forward_only=req.kind == "forward",
                         time_block=chunk_time_block(req, s1 - s0, a.K))
            p.models[0].load(a.vel)
            p.forward()
            if req.kind == "gradient":
                r = p.resid_all()
                torch.sub(p.frd_all(), req.d_obs[s0:s1], out=r)
                rv = r.view(-1)
                a.loss += 0.5 * torch.dot(rv, rv).double()
                p.backward()
                for s in range(p.S):
                    a.grad += p.grad_of(s)[:, :req.nx + req.order]
            else:
                a.frd[s0:s1].copy_(p.frd_all())
            a.plan = p
            a.event = torch.cuda.Event()
            a.event.record(a.stream)
        a.res.chunks += 1
        cells = (req.nx + req.order) * (req.ny + req.order) * req.nt * (s1 - s0)
        a.res.cell_updates += cells * (2 if req.kind == "gradient" else 1)
        a.next_shot = s1

    def _retire(self):
        still = []
        for a in self.active:
            if not a.event.query():
                still.append(a)
                continue
            req = a.req
            if a.next_shot < req.nshots:
                self._launch_chunk(a)
                still.append(a)
                continue
            a.plan = None
            ne = req.order // 2
            if req.kind == "gradient":
                a.res.loss = float(a.loss)
                a.res.grad = a.grad[ne:ne + req.ny, ne:ne + req.nx].transpose(0, 1)
            else:
                a.res.frd = a.frd
            a.res.finished = time.perf_counter()
            if self.verbose:
                print("serving: request %d (%s) done, latency %.1f ms" % (req.rid, req.kind, 1e3 * a.res.latency),
                      file=sys.stderr, flush=True)
            self.in_use -= a.cost
            self.free_slots.append(a.slot)
            self.free_slots.sort()
            self.done.append(a.res)
        self.active = still

    def poll(self, now=None):
        now = time.perf_counter() if now is None else now
        self._retire()
        self._admit(now)

    def idle(self):
        return not self.waiting and not self.active

    def run(self, requests, realtime=True, poll_s=0.0005):
        """Serve requests (sorted by arrival offset in seconds from now when
        realtime, else all at once).  Returns the results in request order."""
        t0 = time.perf_counter()
        pending = sorted(requests, key=lambda r: r.arrival)
        results = {}
        for r in pending:
            r.arrival = (t0 + r.arrival) if realtime else t0
        i = 0
        while i < len(pending) or not self.idle():
            now = time.perf_counter()
            while i < len(pending) and pending[i].arrival <= now:
                res = self.submit(pending[i])
                results[pending[i].rid] = res
                i += 1
            self.poll(now)
            if not self.idle() or i < len(pending):
                time.sleep(poll_s)
        torch.cuda.synchronize(self.device)
        return [results[r.rid] for r in requests]


# ---- workload generator (the serving stream) ---------------------------------

GRIDS = ((256, 512), (512, 1024), (800, 2400), (1200, 3600))  # (nz, nx)


def make_stream(n, rate_per_s, seed=5, grids=GRIDS, nt_range=(2000, 6000), shots_range=(1, 64),
                freq_range=(5.0, 15.0), h=10.0, dt=1e-3, order=8, device="cuda", frac_gradient=0.5):
    """n requests with Poisson arrivals (seeded): grid from `grids`, nt uniform
    in nt_range, shots uniform in shots_range at z=2 with receivers on every
    column, ricker(freq uniform in freq_range), random layered + correlated
    perturbation model in 1500-4500 m/s, half forward-only and half gradient
    (observed data: the gathers of a 2%-perturbed model would double the
    work, so a seeded random field of the gathers' shape stands in)."""
    from .source.wavelets import ricker
    from .workloads import dipping_layers, to_user, cell_pos
    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    for k in range(n):
        t += rng.exponential(1.0 / rate_per_s)
        nz, nx = grids[rng.integers(len(grids))]
        nt = int(rng.integers(nt_range[0], nt_range[1] + 1))
        S = int(rng.integers(shots_range[0], shots_range[1] + 1))
        f = float(rng.uniform(*freq_range))
        kind = "gradient" if rng.random() < frac_gradient else "forward"
        v = dipping_layers(nz, nx, seed=int(rng.integers(1 << 30)), perturb=0.05)
        vel = torch.from_numpy(to_user(v))
        w = torch.from_numpy(ricker(f, dt, nt).astype(np.float32))
        xs = (np.arange(S) + 0.5) * nx / S
        sx = [float(cell_pos(int(x), h)) for x in xs]
        d_obs = None
        if kind == "gradient":
            g = torch.Generator(device="cpu").manual_seed(int(rng.integers(1 << 30)))
            d_obs = (1e-3 * torch.randn(S, nt, nx, generator=g)).to(device)
        reqs.append(Request(kind=kind, vel=vel, w=w, sx=sx, sy=[float(cell_pos(2, h))] * S,
                            ry=[float(cell_pos(2, h))] * S, h=h, dt=dt, order=order, d_obs=d_obs,
                            arrival=t, rid=k))
    return reqs


def summarize(results, wall_s):
    ok = [r for r in results if r.error is None]
    lat = np.array([r.latency for r in ok]) * 1e3 if ok else np.array([np.nan])
    cells = sum(r.cell_updates for r in ok)
    return {
        "requests": len(results), "completed": len(ok), "rejected": len(results) - len(ok),
        "p50_ms": round(float(np.percentile(lat, 50)), 2), "p99_ms": round(float(np.percentile(lat, 99)), 2),
        "mean_ms": round(float(np.mean(lat)), 2), "wall_s": round(wall_s, 3),
        "throughput_req_per_s": round(len(ok) / wall_s, 4),
        "throughput_Gcell_updates_per_s": round(cells / wall_s / 1e9, 2),
    }


def main(argv=None):
    """Serving benchmark: python -m torchwi_amd.serving [--requests N] [--rate R]"""
    import argparse
    import json
    ap = argparse.ArgumentParser()
    ap.add_argument("--requests", type=int, default=24)
    ap.add_argument("--rate", type=float, default=0.0, help="arrivals/s (0: 80% of measured capacity)")
    ap.add_argument("--slots", type=int, default=4)
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--scale", type=float, default=1.0, help="scale grids and nt (tests use < 1)")
    args = ap.parse_args(argv)
    sc = args.scale
    grids = tuple((max(32, int(a * sc)), max(64, int(b * sc))) for a, b in GRIDS)
    ntr = (max(50, int(2000 * sc)), max(60, int(6000 * sc)))
    shr = (1, max(1, int(64 * sc)))
    dev = torch.device("cuda:0")
    # capacity probe: serve the stream back to back (all arrivals at t=0)
    probe = make_stream(args.requests, 1.0, seed=args.seed, grids=grids, nt_range=ntr, shots_range=shr)
    srv = Server(dev, slots=args.slots, verbose=True)
    t0 = time.perf_counter()
    srv.run(probe, realtime=False)
    cap_s = time.perf_counter() - t0
    capacity = args.requests / cap_s
    del srv, probe
    torch.cuda.empty_cache()
    rate = args.rate or 0.8 * capacity
    reqs = make_stream(args.requests, rate, seed=args.seed, grids=grids, nt_range=ntr, shots_range=shr)
    srv = Server(dev, slots=args.slots, verbose=True)
    t0 = time.perf_counter()
    res = srv.run(reqs, realtime=True)
    wall = time.perf_counter() - t0
    out = {"workload": "serving stream: %d Poisson requests (seed %d) at %.3f req/s (80%% of the %.3f req/s "
                       "back-to-back capacity), grids %s, nt %s, shots %s, 50/50 forward/gradient"
                       % (args.requests, args.seed, rate, capacity, list(grids), ntr, shr),
           "slots": args.slots, "budget_GB": round(srv.budget / 1e9, 1),
           "peak_in_use_GB": round(srv.peak_in_use / 1e9, 2), "max_concurrent": srv.max_concurrent}
    out.update(summarize(res, wall))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
