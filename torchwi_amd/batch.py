"""In-GPU shot batching for the time-domain 2D propagator.

The reference computes one shot per call and spreads shots over processes
(examples/time/inv_torch.py:123-143 with io/DistributedDataloader.py:7-12):
every call runs its own time loop, so one GPU advances one small grid per
launch.  The classes here advance every shot of a batch by one time step
per launch (torchwi_amd.shots.ShotPlan over the C-ABI plan), keep all state
device-resident, and replace full virtual-source storage by checkpoints of
the wavefield every K steps plus recomputation when storage would not fit.

Each shot's result is computed exactly as a single ``Time2d`` call computes
it (same per-cell arithmetic); a batch gradient is the sum of the per-shot
gradients in shot order.

* ``Time2dBatch``   -- S shots on one (nx, ny) model, receivers on every
                       interior column (the reference's acquisition).
* ``Time2dWindows`` -- S shots, each on its own aperture window cropped
                       from a large model, with its own receiver window.
* ``Time2dForwardBatch`` -- forward modelling only: S shots on one model in
                       a plan that stores no virtual sources or checkpoints
                       (the reference's forward_modeling.py shot loop).
"""
import torch

from .shots import GRAPH_STEPS, ShotPlan, auto_time_block, default_ckpt_interval, plan_bytes

__all__ = ["Time2dBatch", "Time2dWindows", "Time2dForwardBatch", "default_ckpt_interval"]


def _as_list(x):
    if isinstance(x, torch.Tensor):
        return [float(v) for v in x.detach().reshape(-1).cpu().tolist()]
    if isinstance(x, (int, float)):
        return [float(x)]
    return [float(v) for v in x]


def _wavelet(w):
    w = torch.as_tensor(w)
    if w.dtype != torch.float32:
        raise RuntimeError("w: expected scalar type Float but found %s" % w.dtype)
    return w.contiguous()


def _resolve_time_block(time_block, cells_per_step, ckpt_every):
    """Steps per launch: the given number, or for "auto" the blocked kernel
    on large plans (shots.auto_time_block) unless an explicit checkpoint
    interval is not a multiple of its block."""
    if time_block != "auto":
        return int(time_block)
    tb = auto_time_block(cells_per_step)
    if not isinstance(ckpt_every, str) and int(ckpt_every) > 0 and int(ckpt_every) % tb:
        return 1
    return tb


def _choose_ckpt(ckpt_every, need_full, nt, device, hbm_fraction, held=0, time_block=1):
    if ckpt_every != "auto":
        return int(ckpt_every)
    free, _ = torch.cuda.mem_get_info(device)
    if need_full <= hbm_fraction * (free + held):
        return 0
    return default_ckpt_interval(nt, time_block)


class _BatchOperator(torch.autograd.Function):

    @staticmethod
    def forward(ctx, vel, args):
        sx, sy, ry, m = args
        p = m.prepare(vel, sx, sy, ry)
        p.forward()
        m._fwd_generation = getattr(m, "_fwd_generation", 0) + 1
        ctx.m, ctx.p, ctx.generation = m, p, m._fwd_generation
        return p.frd_all()

    @staticmethod
    def backward(ctx, grad_output):
        m, p = ctx.m, ctx.p
        if m._fwd_generation != ctx.generation:
            # the plan's stored virtual sources / checkpoints now belong to a later forward
            raise RuntimeError("the module ran another forward after this one: its saved wavefield storage "
                               "was overwritten; call backward before the module's next forward")
        p.resid_all().copy_(grad_output)
        p.backward()
        return m.reduce_grad(p), None


class Time2dBatch(torch.nn.Module):
    """Many shots of one velocity model, marched together on one GPU (in
    lockstep within each of the plan's shot groups).

    ``forward(vel, sx, sy, ry)`` takes vel (nx, ny) and per-shot positions
    (sequences or 1-D tensors of length S) and returns the gathers
    (S, nt, nx); its backward gives the summed gradient (nx, ny).
    ``ckpt_every``: 0 = full virtual-source storage, K > 0 = checkpoint
    every K steps and recompute, "auto" = full storage when it fits in
    ``hbm_fraction`` of free device memory, else K = ceil(sqrt(nt)).
    ``time_block``: time steps per kernel launch; "auto" picks the
    temporally blocked kernel (3 steps) for plans of at least
    shots.AUTO_BLOCK_CELLS cells per step and one step per launch below.
    ``streams``: shot groups on separate streams (shots.ShotPlan; 0 automatic).
    ``grad_group``: shots per gradient group of the one-step adjoint
    (shots.ShotPlan; values > 1 are refused for now); 1 keeps one gradient
    per shot, whose sum in shot order is the returned gradient.
    The returned gathers are the plan's buffer: the next forward overwrites
    them (as the reference's module buffer, TimeFunction.py:128).
    """

    def __init__(self, nx, ny, h, w, dt, order, device="cuda", ckpt_every="auto",
                 hbm_fraction=0.8, tile_rows=0, graph_steps=GRAPH_STEPS, time_block="auto",
                 tile_cols=0, grad_group=1, streams=0):
        super().__init__()
        self.streams = int(streams)
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise NotImplementedError("Time2dBatch runs on the GPU only")
        self.graph_steps = graph_steps
        self.time_block = time_block
        if order not in (2, 4, 8):
            raise ValueError("Wrong order: use 2, 4, and 8")
        self.nx, self.ny, self.order, self.ne = nx, ny, order, order // 2
        self.h, self.dt = float(h), float(dt)
        self.w = _wavelet(w)
        self.nt = int(self.w.numel())
        self.ckpt_every = ckpt_every
        self.hbm_fraction = hbm_fraction
        self.tile_rows = tile_rows
        self.tile_cols = tile_cols
        self.grad_group = int(grad_group)
        self._plan = None
        self._key = None

    def _shots(self, sx, sy, ry):
        return [dict(model=0, sx=a, sy=b, ry=c, rx0=0, nrec=self.nx, w=self.w) for a, b, c in zip(sx, sy, ry)]

    def prepare(self, vel, sx, sy, ry):
        sx, sy, ry = _as_list(sx), _as_list(sy), _as_list(ry)
        if not (len(sx) == len(sy) == len(ry)) or len(sx) == 0:
            raise ValueError("sx, sy, ry must have the same positive length")
        key = (tuple(sx), tuple(sy), tuple(ry))
        if self._key != key:
            shots = self._shots(sx, sy, ry)
            models = [(self.nx, self.ny, self.order)]
            held = plan_bytes(models, self._plan.shots, self.nt, self._plan.K) if self._plan else 0
            tb = _resolve_time_block(self.time_block, len(shots) * (self.nx + self.order) * (self.ny + self.order),
                                     self.ckpt_every)
            K = _choose_ckpt(self.ckpt_every, plan_bytes(models, shots, self.nt, 0), self.nt,
                             self.device, self.hbm_fraction, held, tb)
            self._plan = None
            torch.cuda.empty_cache()
            self._plan = ShotPlan(self.order, self.nt, self.h, self.dt, [(self.nx, self.ny)], shots, K,
                                  self.device, tile_rows=self.tile_rows, graph_steps=self.graph_steps,
                                  time_block=tb, tile_cols=self.tile_cols,
                                  grad_group=self.grad_group if tb == 1 else 1, streams=self.streams)
            self._key = key
        self._plan.models[0].load(vel)
        return self._plan

    @property
    def plan(self):
        return self._plan

    def reduce_grad(self, p):
        """Sum the per-shot (or per-group) padded gradients in shot order,
        crop to the model and transpose to (nx, ny) (TimeFunction.py:141-145)."""
        ne = self.ne
        tot = p.grad_of(0).clone()
        for s in range(p.grad_group, p.S, p.grad_group):
            tot.add_(p.grad_of(s))
        return tot[ne:ne + self.ny, ne:ne + self.nx].transpose(0, 1)

    def forward(self, vel, sx, sy, ry):
        return _BatchOperator.apply(vel, (sx, sy, ry, self))

    @torch.no_grad()
    def misfit_gradient(self, vel, sx, sy, ry, d_obs):
        """One FWI gradient: forward all shots, residual frd - d_obs, L2 loss
        0.5*||frd - d_obs||^2 (inv_torch.py:134), adjoint, summed gradient.
        Returns (loss, grad (nx, ny))."""
        p = self.prepare(vel, sx, sy, ry)
        p.forward()
        self._fwd_generation = getattr(self, "_fwd_generation", 0) + 1
        r = p.resid_all()
        torch.sub(p.frd_all(), d_obs, out=r)
        rv = r.view(-1)
        loss = 0.5 * torch.dot(rv, rv)
        p.backward()
        return loss, self.reduce_grad(p)

    @torch.no_grad()
    def modeling(self, vel, sx, sy, ry):
        p = self.prepare(vel, sx, sy, ry)
        p.forward()
        self._fwd_generation = getattr(self, "_fwd_generation", 0) + 1
        return p.frd_all()


class Time2dWindows(torch.nn.Module):
    """Shots on per-shot aperture windows of one large model, in one plan.

    windows: list of dicts {x0, x1, sx, sy, ry, rx0, nrec} in cells of the
    full model: the shot's sub-model is columns [x0, x1) of vel (all depths),
    edge-replicated like any Time2d model; sx, sy, ry are physical positions
    inside the window (relative to its left edge); receivers are window
    columns [rx0, rx0+nrec).  A shot's results equal Time2d on its cropped
    sub-model.  Gradients of all windows are summed into the full model's
    (nx, ny) gradient in shot order.
    """

    def __init__(self, nx, ny, h, w, dt, order, windows, device="cuda", ckpt_every="auto",
                 hbm_fraction=0.8, tile_rows=0, graph_steps=GRAPH_STEPS, time_block="auto",
                 tile_cols=0, streams=0):
        super().__init__()
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise NotImplementedError("Time2dWindows runs on the GPU only")
        self.nx, self.ny, self.order, self.ne = nx, ny, order, order // 2
        self.h, self.dt = float(h), float(dt)
        self.w = _wavelet(w)
        self.nt = int(self.w.numel())
        self.windows = [dict(q) for q in windows]
        for q in self.windows:
            if not (0 <= q["x0"] < q["x1"] <= nx):
                raise ValueError("window [%d, %d) outside [0, %d)" % (q["x0"], q["x1"], nx))
        models = [(q["x1"] - q["x0"], ny) for q in self.windows]
        shots = [dict(model=i, sx=q["sx"], sy=q["sy"], ry=q["ry"], rx0=q["rx0"], nrec=q["nrec"],
                      w=q.get("w", self.w)) for i, q in enumerate(self.windows)]
        time_block = _resolve_time_block(time_block, sum((a + order) * (b + order) for a, b in models), ckpt_every)
        K = _choose_ckpt(ckpt_every, plan_bytes([(a, b, order) for a, b in models], shots, self.nt, 0),
                         self.nt, self.device, hbm_fraction, time_block=time_block)
        self.plan = ShotPlan(order, self.nt, h, dt, models, shots, K, self.device, tile_rows=tile_rows,
                             graph_steps=graph_steps, time_block=time_block, tile_cols=tile_cols,
                             streams=streams)

    def load(self, vel):
        for i, q in enumerate(self.windows):
            self.plan.models[i].load(vel[q["x0"]:q["x1"], :])

    @torch.no_grad()
    def modeling(self, vel):
        self.load(vel)
        self.plan.forward()
        return [self.plan.frd_of(s) for s in range(self.plan.S)]

    @torch.no_grad()
    def misfit_gradient(self, vel, d_obs):
        """d_obs: list of (nt, nrec_s) gathers.  Returns (loss, grad (nx, ny))."""
        p = self.plan
        self.load(vel)
        p.forward()
        loss = torch.zeros((), dtype=torch.float32, device=self.device)
        for s in range(p.S):
            r = p.resid_of(s)
            torch.sub(p.frd_of(s), d_obs[s], out=r)
            rv = r.view(-1)
            loss += 0.5 * torch.dot(rv, rv)
        p.backward()
        ne = self.ne
        grad = torch.zeros(self.ny, self.nx, dtype=torch.float32, device=self.device)
        for s, q in enumerate(self.windows):
            nxw = q["x1"] - q["x0"]
            grad[:, q["x0"]:q["x1"]] += p.grad_of(s)[ne:ne + self.ny, ne:ne + nxw]
        return loss, grad.transpose(0, 1)


class Time2dForwardBatch(torch.nn.Module):
    """Forward modelling of many shots of one velocity model on one GPU.

    The batched replacement of the reference's modelling loop, one
    ``Time2dForward`` call per shot under Horovod
    (examples/time/forward_modeling.py:44-52): ``forward(vel, sx, sy, ry)``
    takes vel (nx, ny) and per-shot positions (length S) and returns the
    gathers (S, nt, nx), each shot's equal to ``Time2dForward``'s for it bit
    for bit.  The plan is forward-only (TwPlanSpec.forward_only): no virtual
    sources, no checkpoints, no adjoint ring, so its memory is the wavefield
    rings and the gathers (about 5 fields per shot).  There is no gradient
    (the reference's forward-only operator has none either,
    TimeFunction.py:24-29).  ``time_block`` "auto" picks the temporally
    blocked kernel from shots.AUTO_BLOCK_CELLS_FORWARD cells per step.  The
    returned gathers are the plan's buffer: the next forward overwrites them.
    """

    def __init__(self, nx, ny, h, w, dt, order, device="cuda", time_block="auto", graph_steps=GRAPH_STEPS,
                 tile_rows=0, tile_cols=0, streams=0):
        super().__init__()
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise NotImplementedError("Time2dForwardBatch runs on the GPU only")
        if order not in (2, 4, 8):
            raise ValueError("Wrong order: use 2, 4, and 8")
        self.nx, self.ny, self.order = nx, ny, order
        self.h, self.dt = float(h), float(dt)
        self.w = _wavelet(w)
        self.nt = int(self.w.numel())
        self.time_block, self.graph_steps = time_block, graph_steps
        self.tile_rows, self.tile_cols, self.streams = tile_rows, tile_cols, int(streams)
        self._plan = None
        self._key = None

    @property
    def plan(self):
        return self._plan

    @torch.no_grad()
    def forward(self, vel, sx, sy, ry):
        sx, sy, ry = _as_list(sx), _as_list(sy), _as_list(ry)
        if not (len(sx) == len(sy) == len(ry)) or len(sx) == 0:
            raise ValueError("sx, sy, ry must have the same positive length")
        key = (tuple(sx), tuple(sy), tuple(ry))
        if self._key != key:
            shots = [dict(model=0, sx=a, sy=b, ry=c, rx0=0, nrec=self.nx, w=self.w) for a, b, c in zip(sx, sy, ry)]
            cells = len(shots) * (self.nx + self.order) * (self.ny + self.order)
            tb = (auto_time_block(cells, forward_only=True) if self.time_block == "auto"
                  else int(self.time_block))
            self._plan = None
            torch.cuda.empty_cache()
            self._plan = ShotPlan(self.order, self.nt, self.h, self.dt, [(self.nx, self.ny)], shots, 0, self.device,
                                  tile_rows=self.tile_rows, graph_steps=self.graph_steps, time_block=tb,
                                  forward_only=True, tile_cols=self.tile_cols, streams=self.streams)
            self._key = key
        self._plan.models[0].load(vel)
        self._plan.forward()
        return self._plan.frd_all()
