"""Continuous-batching request loop for forward-modelling and misfit-gradient
jobs on one GPU.

Requests (a model, its shots, a wavelet, and -- for gradient jobs -- the
observed gathers) arrive over time.  The server keeps up to ``slots``
requests on the device at once and admits a waiting request as soon as a
slot and enough of the HBM budget are free (footprint = its largest shot
chunk's plan plus its inputs and results).  Wavefields, checkpoints,
gathers and gradients stay device-resident for the whole request.

Execution is in batches: every request whose next chunk of shots is due
joins a batch with the other due requests of the same kind, grid, order,
time axis and storage mode; a batch is one shot plan (each request's shots
on its own model) launched on one of the server's HIP streams, so the shots
of different requests march in lockstep in the same persistent launches
(one per time step, or per 3 steps with the temporally blocked kernel,
replayed from captured hipGraph blocks for one-step plans).  Batches of
different keys run concurrently on different streams.

A request's shots run in chunks whose size depends only on the request (a
fixed per-request byte cap), and every per-request reduction (the loss, the
gradient sum over its shots) is taken over that request's own buffers in
shot order, so each result is bit-identical alone and in a mix.  A request
whose smallest chunk cannot fit the budget is rejected at submission.  A
request whose chunk fails to build or launch (bad geometry, allocation
failure, a library error) is ended with ``error`` set: its slot and budget
are released and the other requests continue; a failed batch of several
requests is retried request by request to find the one at fault.
``Server.cancel(rid)`` ends a request the same way: at once while it waits,
else once its chunk on the device (if any) has run.

Admission is first come first served with bounded skip-ahead: when the
oldest waiting request does not fit the free budget, a later one that fits
may go first, at most ``max_skips`` times, after which nothing overtakes the
oldest until it fits.

There is no counterpart in the reference, whose shot loop is one process per
GPU under Horovod (examples/time/inv_torch.py:123-143).
"""
import math
import sys
import time
from dataclasses import dataclass
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
    cancelled: bool = False                # ended by Server.cancel
    chunks: int = 0
    shared_batches: int = 0                # chunks that ran in a batch with other requests
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
    """Steps per launch for a plan of nshots shots of req's grid, by plan size
    (shots.auto_time_block; forward modelling plans with their own
    threshold).  Either kernel gives the same bits."""
    cells = nshots * (req.nx + req.order) * (req.ny + req.order)
    if req.kind != "gradient":
        return auto_time_block(cells, forward_only=True)
    if K % 3:
        return 1
    return auto_time_block(cells)


def io_bytes(req):
    S, nt, nx = req.nshots, req.nt, req.nx
    b = 4 * (req.nx * req.ny + nt)
    if req.kind == "gradient":
        b += 4 * (S * nt * nx + 2 * nx * req.ny)
    else:
        b += 4 * S * nt * nx
    return b


def batch_key(req, K):
    """Requests whose chunks can share one plan: same kind, grid, order,
    time axis and storage mode (each keeps its own model and wavelet)."""
    return (req.kind, req.nx, req.ny, req.order, req.nt, float(req.h), float(req.dt), K)


def select_admissions(waiting, free_slots, free_bytes, now, max_skips):
    """Indices of waiting entries to admit, in order.  waiting: sequence of
    (arrival, cost, skips) in arrival order.  First come first served with
    bounded skip-ahead: an entry that fits may overtake an older one that
    does not, unless that older one has already been overtaken max_skips
    times.  Returns (indices, overtaken) where overtaken lists the
    positions whose skip count grows by one per admitted overtaker."""
    picked, overtaken = [], []
    blocked = []   # positions of older entries that did not fit
    for i, (arrival, cost, skips) in enumerate(waiting):
        if len(picked) >= free_slots or arrival > now:
            break
        if cost <= free_bytes and all(waiting[j][2] + overtaken.count(j) < max_skips for j in blocked):
            picked.append(i)
            free_bytes -= cost
            overtaken.extend(blocked)
        elif cost > free_bytes:
            blocked.append(i)
        else:
            break  # an older entry has been overtaken too often: wait for it
    return picked, overtaken


class _Active:
    """A request on the device: its accumulators and progress."""

    def __init__(self, req, res, per, K, cost, slot):
        self.req, self.res = req, res
        self.per, self.K, self.cost, self.slot = per, K, cost, slot
        self.next_shot = 0
        self.loss = self.grad = self.frd = self.vel = None
        self.failed = False      # ended (error or cancel): no further chunk, results dropped
        self.cancelled = False


class _Batch:
    """One shot plan holding a chunk of each member request."""

    def __init__(self, stream, members):
        self.stream, self.members = stream, members   # members: [(active, s0, s1)]
        self.plan = None
        self.event = None


class Server:
    """Continuous-batching executor.  budget_bytes: HBM the server may use;
    slots: requests in flight at once; cap_bytes: per-request chunk cap
    (default budget/slots); merge: let due chunks of compatible requests
    share a plan; max_skips: how often a waiting request may be overtaken by
    later ones that fit."""

    def __init__(self, device="cuda", budget_bytes=None, slots=4, cap_bytes=None,
                 graph_steps=GRAPH_STEPS, verbose=False, merge=True, max_skips=4):
        self.verbose = verbose
        self.device = torch.device(device)
        if budget_bytes is None:
            free, _ = torch.cuda.mem_get_info(self.device)
            budget_bytes = int(0.7 * free)
        self.budget = int(budget_bytes)
        self.slots = int(slots)
        self.cap = int(cap_bytes) if cap_bytes else self.budget // self.slots
        self.graph_steps = graph_steps
        self.merge = bool(merge)
        self.max_skips = int(max_skips)
        self.streams = [torch.cuda.Stream(self.device) for _ in range(self.slots)]
        self.free_streams = list(range(self.slots))
        self.in_use = 0
        self.free_slots = list(range(self.slots))
        self.waiting = []     # [req, res, per, K, cost, skips]
        self.due = []         # actives whose next chunk should be launched
        self.batches = []
        self.active = []
        self.done = []
        self.peak_in_use = 0
        self.max_concurrent = 0
        self.max_batch_requests = 0

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
                self.waiting.append([req, res, per, K, cost, 0])
                return res
        res.finished = res.admitted = time.perf_counter()
        self.done.append(res)
        return res

    # ---- admission ----------------------------------------------------------
    def _admit(self, now):
        view = [(w[0].arrival, w[4], w[5]) for w in self.waiting]
        picked, overtaken = select_admissions(view, len(self.free_slots), self.budget - self.in_use, now,
                                              self.max_skips)
        for j in overtaken:
            self.waiting[j][5] += 1
        for i in picked:
            req, res, per, K, cost, _ = self.waiting[i]
            slot = self.free_slots.pop(0)
            self.in_use += cost
            self.peak_in_use = max(self.peak_in_use, self.in_use)
            res.admitted = time.perf_counter()
            a = _Active(req, res, per, K, cost, slot)
            self.active.append(a)
            self.max_concurrent = max(self.max_concurrent, len(self.active))
            try:
                dev = self.device
                a.vel = req.vel.to(dev, torch.float32)
                if req.kind == "gradient":
                    a.loss = torch.zeros((), dtype=torch.float64, device=dev)
                    a.grad = torch.zeros(req.ny + req.order, req.nx + req.order, dtype=torch.float32, device=dev)
                else:
                    a.frd = torch.empty(req.nshots, req.nt, req.nx, dtype=torch.float32, device=dev)
            except Exception as e:  # noqa: BLE001 -- any failure ends this request only
                self._fail(a, e)
                continue
            self.due.append(a)
        for i in sorted(picked, reverse=True):
            self.waiting.pop(i)

    def cancel(self, rid):
        """Cancel request rid.  A waiting request leaves the queue at once; an
        admitted one launches no further chunk and ends, its slot and HBM
        budget released, as soon as none of its chunks is on the device (a
        chunk in flight, possibly in a plan shared with other requests, runs
        to its end; those requests are unaffected).  The result gets
        cancelled = True and error "cancelled".  Returns False when rid is not
        waiting or in flight."""
        for i, w in enumerate(self.waiting):
            if w[0].rid == rid:
                self.waiting.pop(i)
                res = w[1]
                res.cancelled = True
                res.error = "cancelled"
                res.finished = time.perf_counter()
                self.done.append(res)
                return True
        for a in self.active:
            if a.req.rid == rid and not a.failed:
                a.cancelled = True
                if not any(m[0] is a for b in self.batches for m in b.members):
                    self.due = [x for x in self.due if x is not a]
                    self._end(a, "cancelled")
                return True
        return False

    # ---- execution ----------------------------------------------------------
    def _end(self, a, error):
        """End an admitted request without a result (failure or cancel)."""
        if a.failed:
            return
        a.failed = True
        a.res.error = error
        a.res.cancelled = a.cancelled
        a.res.finished = time.perf_counter()
        if self.verbose:
            print("serving: request %d ended: %s" % (a.req.rid, error), file=sys.stderr, flush=True)
        self._release(a)

    def _fail(self, a, err):
        self._end(a, "%s: %s" % (type(err).__name__, err))

    def _release(self, a):
        self.in_use -= a.cost
        self.free_slots.append(a.slot)
        self.free_slots.sort()
        self.active = [x for x in self.active if x is not a]
        a.loss = a.grad = a.frd = a.vel = None
        self.done.append(a.res)

    def _run_batch(self, members):
        """Build and launch one plan over the members' next chunks (on a free
        stream); results are accumulated per member on that stream."""
        sid = self.free_streams.pop(0)
        b = _Batch(sid, members)
        stream = self.streams[sid]
        try:
            req0 = members[0][0].req
            models, shots = [], []
            for i, (a, s0, s1) in enumerate(members):
                r = a.req
                models.append((r.nx, r.ny))
                shots += [dict(model=i, sx=r.sx[k], sy=r.sy[k], ry=r.ry[k], rx0=0, nrec=r.nx, w=r.w)
                          for k in range(s0, s1)]
            K = members[0][0].K
            with torch.cuda.stream(stream):
                p = ShotPlan(req0.order, req0.nt, req0.h, req0.dt, models, shots, K, self.device,
                             graph_steps=self.graph_steps, forward_only=req0.kind == "forward",
                             time_block=chunk_time_block(req0, len(shots), K))
                for i, (a, _, _) in enumerate(members):
                    p.models[i].load(a.vel)
                p.forward()
                o = 0
                if req0.kind == "gradient":
                    rall = p.resid_all()
                    for a, s0, s1 in members:
                        torch.sub(p.frd_all()[o:o + s1 - s0], a.req.d_obs[s0:s1], out=rall[o:o + s1 - s0])
                        o += s1 - s0
                    p.backward()
        except Exception:
            self.free_streams.append(sid)
            self.free_streams.sort()
            raise
        # Accumulate into the members' results.  A failure past this point
        # would leave partial sums, so it ends every member instead of being
        # retried.
        try:
            with torch.cuda.stream(stream):
                o = 0
                for a, s0, s1 in members:
                    n = s1 - s0
                    if req0.kind == "gradient":
                        # the request's own residual rows, in a buffer of their own
                        rv = rall[o:o + n].clone().view(-1)
                        a.loss += 0.5 * torch.dot(rv, rv).double()
                        for s in range(o, o + n):
                            a.grad += p.grad_of(s)[:, :a.req.nx + a.req.order]
                    else:
                        a.frd[s0:s1].copy_(p.frd_all()[o:o + n])
                    o += n
                b.plan = p
                b.event = torch.cuda.Event()
                b.event.record(stream)
        except Exception as e:  # noqa: BLE001
            self.free_streams.append(sid)
            self.free_streams.sort()
            for a, _, _ in members:
                self._fail(a, e)
            return
        for a, s0, s1 in members:
            a.res.chunks += 1
            if len(members) > 1:
                a.res.shared_batches += 1
            r = a.req
            cells = (r.nx + r.order) * (r.ny + r.order) * r.nt * (s1 - s0)
            a.res.cell_updates += cells * (2 if r.kind == "gradient" else 1)
            a.next_shot = s1
        self.max_batch_requests = max(self.max_batch_requests, len(members))
        self.batches.append(b)

    def _launch_due(self):
        groups = {}
        for a in self.due:
            if not a.failed and not a.cancelled:
                groups.setdefault(batch_key(a.req, a.K) if self.merge else id(a), []).append(a)
        self.due = []
        for group in groups.values():
            members = [(a, a.next_shot, min(a.req.nshots, a.next_shot + a.per)) for a in group]
            try:
                self._run_batch(members)
            except Exception as e:  # noqa: BLE001
                if len(members) == 1:
                    self._fail(members[0][0], e)
                    continue
                # find the request at fault: run each member's chunk alone
                for m in members:
                    try:
                        self._run_batch([m])
                    except Exception as e1:  # noqa: BLE001
                        self._fail(m[0], e1)

    def _retire(self):
        still = []
        for b in self.batches:
            try:
                if not b.event.query():
                    still.append(b)
                    continue
            except Exception as e:  # noqa: BLE001
                for a, _, _ in b.members:
                    self._fail(a, e)
                b.plan = None
                self.free_streams.append(b.stream)
                continue
            b.plan = None
            self.free_streams.append(b.stream)
            for a, _, _ in b.members:
                if a.failed:
                    continue
                if a.cancelled:
                    self._end(a, "cancelled")
                elif a.next_shot < a.req.nshots:
                    self.due.append(a)
                else:
                    self._finish(a)
        self.free_streams.sort()
        self.batches = still

    def _finish(self, a):
        req = a.req
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
        self.active = [x for x in self.active if x is not a]
        self.done.append(a.res)

    def poll(self, now=None):
        now = time.perf_counter() if now is None else now
        self._retire()
        self._admit(now)
        self._launch_due()

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


def make_request(kind, nz, nx, nt, nshots, freq, seed, h=10.0, dt=1e-3, order=8, device="cuda",
                 arrival=0.0, rid=-1):
    """One request of the serving stream: a random layered + correlated
    perturbation model in 1500-4500 m/s (seeded), nshots ricker(freq) shots
    evenly spaced at z=2 with receivers on every column; gradient requests
    get a seeded random field of the gathers' shape as observed data
    (modelling a perturbed model would double the work)."""
    from .source.wavelets import ricker
    from .workloads import dipping_layers, to_user, cell_pos
    rng = np.random.default_rng(seed)
    v = dipping_layers(nz, nx, seed=int(rng.integers(1 << 30)), perturb=0.05)
    vel = torch.from_numpy(to_user(v))
    w = torch.from_numpy(ricker(freq, dt, nt).astype(np.float32))
    xs = (np.arange(nshots) + 0.5) * nx / nshots
    sx = [float(cell_pos(int(x), h)) for x in xs]
    d_obs = None
    if kind == "gradient":
        g = torch.Generator(device="cpu").manual_seed(int(rng.integers(1 << 30)))
        d_obs = (1e-3 * torch.randn(nshots, nt, nx, generator=g)).to(device)
    return Request(kind=kind, vel=vel, w=w, sx=sx, sy=[float(cell_pos(2, h))] * nshots,
                   ry=[float(cell_pos(2, h))] * nshots, h=h, dt=dt, order=order, d_obs=d_obs,
                   arrival=arrival, rid=rid)


def make_stream(n, rate_per_s, seed=5, grids=GRIDS, nt_range=(2000, 6000), shots_range=(1, 64),
                freq_range=(5.0, 15.0), h=10.0, dt=1e-3, order=8, device="cuda", frac_gradient=0.5):
    """n requests with Poisson arrivals (seeded): grid from `grids`, nt uniform
    in nt_range, shots uniform in shots_range, ricker(freq uniform in
    freq_range), half forward-only and half gradient (make_request)."""
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
        reqs.append(make_request(kind, nz, nx, nt, S, f, int(rng.integers(1 << 30)), h=h, dt=dt, order=order,
                                 device=device, arrival=t, rid=k))
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
        "chunks_in_shared_batches": int(sum(r.shared_batches for r in results)),
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
    ap.add_argument("--no-merge", action="store_true", help="one plan per request chunk")
    args = ap.parse_args(argv)
    sc = args.scale
    grids = tuple((max(32, int(a * sc)), max(64, int(b * sc))) for a, b in GRIDS)
    ntr = (max(50, int(2000 * sc)), max(60, int(6000 * sc)))
    shr = (1, max(1, int(64 * sc)))
    dev = torch.device("cuda:0")
    merge = not args.no_merge
    # capacity probe: serve the stream back to back (all arrivals at t=0)
    probe = make_stream(args.requests, 1.0, seed=args.seed, grids=grids, nt_range=ntr, shots_range=shr)
    srv = Server(dev, slots=args.slots, verbose=True, merge=merge)
    t0 = time.perf_counter()
    srv.run(probe, realtime=False)
    cap_s = time.perf_counter() - t0
    capacity = args.requests / cap_s
    del srv, probe
    torch.cuda.empty_cache()
    rate = args.rate or 0.8 * capacity
    reqs = make_stream(args.requests, rate, seed=args.seed, grids=grids, nt_range=ntr, shots_range=shr)
    srv = Server(dev, slots=args.slots, verbose=True, merge=merge)
    t0 = time.perf_counter()
    res = srv.run(reqs, realtime=True)
    wall = time.perf_counter() - t0
    out = {"workload": "serving stream: %d Poisson requests (seed %d) at %.3f req/s (80%% of the %.3f req/s "
                       "back-to-back capacity), grids %s, nt %s, shots %s, 50/50 forward/gradient"
                       % (args.requests, args.seed, rate, capacity, list(grids), ntr, shr),
           "slots": args.slots, "merge": merge, "budget_GB": round(srv.budget / 1e9, 1),
           "peak_in_use_GB": round(srv.peak_in_use / 1e9, 2), "max_concurrent": srv.max_concurrent,
           "max_requests_per_batch": srv.max_batch_requests}
    out.update(summarize(res, wall))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
