"""Device-resident shot plans: the host side of the batched C-ABI plan.

A ``ShotPlan`` holds every device buffer of a set of shots that march in
lockstep (one launch per time step, or per 3 with the blocked kernel, for
all of them, or for each of two shot groups on two streams) and the C plan
built over them.  Shots may live on different grids: each shot names one of the
plan's models (a padded velocity grid with its derived arrays), so a plan
can carry many shots of one model (the reference's shot loop,
examples/time/inv_torch.py:123-143) or one cropped sub-model per shot
(aperture windows).  All per-shot buffers of one kind share one arena
tensor, so the C plan sees plain pointers and the caching allocator sees a
handful of large allocations.
"""
import ctypes
import math

import torch

from . import _lib

ALIGN = 32  # row stride alignment in floats (128 B)
GRAPH_STEPS = 100  # time steps per captured hipGraph block (0: direct launches)
TIME_BLOCK = 1     # time steps per launch of the step kernel (2..4: temporal blocking)
# Cells per time step (all shots of a plan) from which the temporally blocked
# kernel with 3 steps per launch is the default.  Below, the three wavefields
# of all shots (12 B per cell) largely fit MI355X's 256 MB Infinity Cache and
# the one-step kernel streams from it; above, the blocked kernel, which reads
# and writes the fields once per 3 steps, wins.  Measured with both kernels'
# shots in two groups on two streams, nt = 600 (Gcell-updates/s, one step per
# launch / blocked, checkpointed and full-storage gradients): 32 shots of
# 1032 x 392, 13M cells: 214 / 194 checkpointed; 3 shots of 3608 x 1208, 13M:
# 205 / 201 and 285 / 287; 5 shots of 3008 x 1008, 15M: 215 / 233 and
# 290 / 331; 8 shots of 2408 x 808, 16M: 224 / 231 and 292 / 329; 40 shots of
# 1032 x 392, 16M: 211 / 209 and 282 / 297; 48 shots, 19M: 185 / 214 and
# 265 / 305; 8 shots of 3008 x 1008, 24M: 178 / 263; 64 shots of 1032 x 392,
# 26M: 161 / 224.
AUTO_BLOCK_CELLS = 15_000_000
# The same for forward-only modelling plans, which store no virtual sources:
# the one-step forward moves 16 B per cell-update and keeps the cache longer.
# Measured (two shot groups, nt = 600, Gcell-updates/s, one step per launch /
# blocked): 32 shots of 1032 x 392, 13M cells: 491 / 325; 8 shots of
# 2408 x 808, 16M: 551 / 415; 48 shots of 1032 x 392, 19M: 410 / 367;
# 8 shots of 3008 x 1008, 24M: 363 / 499; 64 shots of 1032 x 392, 26M:
# 345 / 398; 12 shots of 3608 x 1208, 52M: 379 / 586; 64 shots of
# 3008 x 1008, 194M: 351 / 638.
AUTO_BLOCK_CELLS_FORWARD = 22_000_000


def auto_time_block(cells_per_step, forward_only=False):
    """Steps per launch for a plan advancing cells_per_step cells per step
    (gradient plans, or forward-only modelling plans)."""
    return 3 if cells_per_step >= (AUTO_BLOCK_CELLS_FORWARD if forward_only else AUTO_BLOCK_CELLS) else 1
RING = 4           # fields per wavefield ring (3 rotate for one-step launches; 2 pairs for blocks)


def round_up(n, m):
    return (n + m - 1) // m * m


def default_ckpt_interval(nt, multiple=1):
    """About sqrt(nt) steps per checkpoint segment, rounded up to a multiple
    of the kernel's time block."""
    k = max(2, int(math.ceil(math.sqrt(nt))))
    m = max(1, int(multiple))
    return (k + m - 1) // m * m


class ModelGrid:
    """A padded model of an (nx, ny) interior: vel and the derived m, c3."""

    def __init__(self, nx, ny, order, device):
        self.nx, self.ny, self.order, self.ne = nx, ny, order, order // 2
        self.dimx, self.dimy = nx + order, ny + order
        self.ld = round_up(self.dimx, ALIGN)
        z = lambda: torch.zeros(self.dimy, self.ld, dtype=torch.float32, device=device)  # noqa: E731
        self.vel, self.vm, self.vc3 = z(), z(), z()

    @property
    def fstride(self):
        return self.dimy * self.ld

    def load(self, vel):
        """Edge-replicate vel (nx, ny), depth fastest, into the padded model,
        as TimeOperator.pad_vel does (TimeOperator.py:55-58)."""
        if tuple(vel.shape) != (self.nx, self.ny):
            raise ValueError("vel must be (nx, ny) = (%d, %d), got %s" % (self.nx, self.ny, tuple(vel.shape)))
        ne = self.ne
        with torch.no_grad():
            vp = torch.nn.functional.pad(
                vel.detach().to(self.vel.device, torch.float32).transpose(0, 1).reshape(1, 1, self.ny, self.nx),
                (ne, ne, ne, ne), mode="replicate").reshape(self.dimy, self.dimx)
            self.vel[:, :self.dimx].copy_(vp)


def plan_bytes(models, shots, nt, ckpt_every, forward_only=False):
    """Device bytes of a ShotPlan (models: list of (nx, ny, order);
    shots: list of dicts with 'model' and 'nrec')."""
    K = int(ckpt_every)
    if forward_only:
        tot = 0
        fs = []
        for nx, ny, order in models:
            f = (ny + order) * round_up(nx + order, ALIGN)
            fs.append(f)
            tot += 3 * f
        for q in shots:
            tot += fs[q["model"]] * RING + nt * q["nrec"]
        return 4 * tot
    nseg = 0 if K == 0 else (nt + K - 1) // K
    slots = nt if K == 0 else K
    fs = []
    tot = 0
    for nx, ny, order in models:
        f = (ny + order) * round_up(nx + order, ALIGN)
        fs.append(f)
        tot += 3 * f
    for q in shots:
        f = fs[q["model"]]
        tot += f * (RING + (RING if K else 0) + slots + 2 * nseg + 1) + 2 * nt * q["nrec"]
    return 4 * tot


class ShotPlan:
    """Buffers and C plan of shots marching in lockstep.

    models: list of (nx, ny) interiors (all of one FD order).
    shots:  list of dicts {model, sx, sy, ry, rx0, nrec, w}; sx, sy, ry are
            physical positions within the shot's model, [rx0, rx0+nrec) its
            receiver columns (interior coordinates), w its (nt,) float32
            wavelet (a tensor; shots may share one).
    ckpt_every: 0 = full virtual-source storage, K > 0 = checkpoints.
    graph_steps: time steps per replayed hipGraph block (0: one direct launch
            per step).
    time_block: time steps advanced per launch by the temporally blocked
            kernel (2..4; 1: one step per launch).  ckpt_every must then be
            a multiple of it.
    tile_rows, tile_cols: tile shape of the one-step kernel (64x32, 64x64,
            128x16, 128x32; 0 automatic) or, with time_block > 1, the row-chunk
            height of the blocked kernel.
    grad_group: 1: one gradient per shot.  > 1 (the imaging condition of up
            to this many consecutive shots of one model accumulated into one
            gradient, the group's first shot's grad_of) is refused by the
            library for now: its adjoint kernel faulted on MI355X for a cause
            not yet found.
    streams: 1 all shots on the caller's stream, 2 two shot groups on two
            streams whose launches overlap (each group marching through its
            steps and checkpoint segments independently), 0 automatic (two
            whenever there are two shots).  The results are the same.
    static_workspace: put the plan's descriptors in the library's static
            per-device workspace (tw_plan_create_static: no device allocation
            by the library; one such plan per device at a time, graph_steps
            must be 0) instead of a workspace tensor.
    """

    def __init__(self, order, nt, h, dt, models, shots, ckpt_every, device, tile_rows=0,
                 graph_steps=GRAPH_STEPS, time_block=TIME_BLOCK, forward_only=False, tile_cols=0,
                 grad_group=1, static_workspace=False, streams=0):
        self.order, self.nt, self.h, self.dt = order, int(nt), float(h), float(dt)
        self.device = torch.device(device)
        self.K = int(ckpt_every)
        self.nseg = 0 if self.K == 0 else (self.nt + self.K - 1) // self.K
        self.models = [ModelGrid(nx, ny, order, self.device) for nx, ny in models]
        self.shots = [dict(q) for q in shots]
        S = self.S = len(self.shots)
        if S == 0:
            raise ValueError("a shot plan needs at least one shot")
        dev = self.device
        fs = [self.models[q["model"]].fstride for q in self.shots]
        nrec = [int(q["nrec"]) for q in self.shots]
        slots = self.nt if self.K == 0 else self.K
        # arena offsets (floats); every field stride is a multiple of 32 floats
        self.foff = [0] * S
        self.roff = [0] * S
        acc = racc = 0
        for s in range(S):
            self.foff[s], self.roff[s] = acc, racc
            acc += fs[s]
            racc += self.nt * nrec[s]
        nf = acc
        self.forward_only = bool(forward_only)
        if self.forward_only and self.K:
            raise ValueError("forward_only plans take ckpt_every = 0")
        # Every buffer of the plan is a 2 MiB-aligned slice of one device
        # allocation, so the plan's working set is one contiguous virtual
        # range.  Measured on MI355X against one allocation per buffer: the
        # checkpointed gradient of 64 shots of 3008 x 1008 (215 GB) took
        # 12.05 s instead of 13.1 s, the full-storage gradient of 32 shots of
        # 1032 x 392 381 ms instead of 390 ms.
        sizes = [("u", RING * nf, True), ("frd", racc, True)]
        if not self.forward_only:
            if self.K:
                sizes.append(("lam", RING * nf, True))
            sizes += [("virt", slots * nf, False), ("grad", nf, True), ("resid", racc, True)]
            if self.K:
                sizes.append(("ckpt", 2 * self.nseg * nf, False))
        aln = (2 << 20) // 4
        total = sum(round_up(n, aln) for _, n, _ in sizes)
        self.arena = torch.empty(total, dtype=torch.float32, device=dev)
        self.lam = self.virt = self.ckpt = self.grad = self.resid = None
        off = 0
        for name, n, zero in sizes:
            t = self.arena[off:off + n]
            if zero:
                t.zero_()
            setattr(self, name, t)
            off += round_up(n, aln)
        if not self.forward_only and not self.K:
            self.lam = self.u
        self._wavelets = []
        spec = _lib.TwPlanSpec(order=order, nt=self.nt, h=self.h, dt=self.dt, ckpt_every=self.K,
                               tile_rows=int(tile_rows), graph_steps=int(graph_steps),
                               time_block=int(time_block), forward_only=int(self.forward_only),
                               tile_cols=int(tile_cols), grad_group=int(grad_group), streams=int(streams))
        self.graph_steps = int(graph_steps)
        self.grad_group = max(1, int(grad_group))
        self.time_block = int(time_block)
        cs = (_lib.TwShotSpec * S)()
        wcache = {}
        for s, q in enumerate(self.shots):
            m = self.models[q["model"]]
            f, o = fs[s], self.foff[s]
            w = q["w"]
            key = id(w)
            if key not in wcache:
                wd = torch.as_tensor(w)
                if wd.dtype != torch.float32:
                    raise RuntimeError("w: expected scalar type Float but found %s" % wd.dtype)
                if wd.numel() < self.nt:
                    raise RuntimeError("w has %d samples, needs nt=%d" % (wd.numel(), self.nt))
                wcache[key] = wd.to(dev).contiguous()
                self._wavelets.append(wcache[key])
            c = cs[s]
            c.dimx, c.dimy, c.ld = m.dimx, m.dimy, m.ld
            c.sx, c.sy, c.ry = float(q["sx"]), float(q["sy"]), float(q["ry"])
            c.rx0, c.nrec = int(q.get("rx0", 0)), nrec[s]
            c.vel, c.vm, c.vc3 = m.vel.data_ptr(), m.vm.data_ptr(), m.vc3.data_ptr()
            c.w = wcache[key].data_ptr()
            # ring field k of shot s lives at arena offset RING*o + k*f
            c.u = self.u.data_ptr() + 4 * RING * o
            c.frd = self.frd.data_ptr() + 4 * self.roff[s]
            if not self.forward_only:
                c.lam = self.lam.data_ptr() + 4 * RING * o
                c.resid = self.resid.data_ptr() + 4 * self.roff[s]
                c.virt = self.virt.data_ptr() + 4 * slots * o
                c.ckpt = (self.ckpt.data_ptr() + 4 * 2 * self.nseg * o) if self.K else None
                c.grad = self.grad.data_ptr() + 4 * o
        self._fs, self._nrec = fs, nrec
        L = _lib.lib()
        handle = ctypes.c_void_p()
        self.workspace = None
        if static_workspace:
            rc = L.tw_plan_create_static(ctypes.byref(spec), cs, S, _lib.stream_of(self.u), ctypes.byref(handle))
            _lib.check(rc, "tw_plan_create_static")
        else:
            nbytes = L.tw_plan_workspace_bytes(ctypes.byref(spec), cs, S)
            if nbytes == 0:
                raise RuntimeError("torchwi_amd: invalid shot plan: %s" % _lib.last_error())
            self.workspace = torch.empty(nbytes, dtype=torch.uint8, device=dev)
            rc = L.tw_plan_create(ctypes.byref(spec), cs, S, _lib.ptr(self.workspace), nbytes,
                                  _lib.stream_of(self.workspace), ctypes.byref(handle))
            _lib.check(rc, "tw_plan_create")
        self.plan = handle

    # ---- views -------------------------------------------------------------
    def frd_of(self, s):
        return self.frd[self.roff[s]:self.roff[s] + self.nt * self._nrec[s]].view(self.nt, self._nrec[s])

    def resid_of(self, s):
        return self.resid[self.roff[s]:self.roff[s] + self.nt * self._nrec[s]].view(self.nt, self._nrec[s])

    def grad_of(self, s):
        """Padded-grid gradient of shot s, (dimy, ld) (with grad_group > 1:
        of the group shot s leads, or zeros)."""
        m = self.models[self.shots[s]["model"]]
        return self.grad[self.foff[s]:self.foff[s] + self._fs[s]].view(m.dimy, m.ld)

    def uniform(self):
        return len(set(self._nrec)) == 1

    def frd_all(self):
        """(S, nt, nrec) view when every shot has the same receiver count."""
        return self.frd.view(self.S, self.nt, self._nrec[0])

    def resid_all(self):
        return self.resid.view(self.S, self.nt, self._nrec[0])

    # ---- execution ---------------------------------------------------------
    def __del__(self):
        try:
            if getattr(self, "plan", None):
                _lib.lib().tw_plan_destroy(self.plan)
                self.plan = None
        except Exception:
            pass

    def forward(self):
        _lib.check(_lib.lib().tw_plan_forward(self.plan, _lib.stream_of(self.u)), "tw_plan_forward")

    def backward(self, accumulate=False):
        _lib.check(_lib.lib().tw_plan_backward(self.plan, int(bool(accumulate)),
                                               _lib.stream_of(self.u)), "tw_plan_backward")

    @property
    def ngroups(self):
        """Shot groups the C plan runs on separate streams (1 or 2)."""
        n = _lib.lib().tw_plan_shot_groups(self.plan)
        _lib.check(min(n, 0), "tw_plan_shot_groups")
        return n

    def set_timing(self, enable):
        _lib.check(_lib.lib().tw_plan_set_timing(self.plan, int(bool(enable))), "tw_plan_set_timing")

    def timing(self, reset=False):
        """{kind: (total ms, launches)} of HIP-event-timed step kernels."""
        n = len(_lib.KINDS)
        ms = (ctypes.c_double * n)()
        cnt = (ctypes.c_longlong * n)()
        _lib.check(_lib.lib().tw_plan_timing(self.plan, int(bool(reset)), ms, cnt), "tw_plan_timing")
        return {k: (ms[i], cnt[i]) for i, k in enumerate(_lib.KINDS)}
