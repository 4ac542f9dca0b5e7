"""In-GPU shot batching for the time-domain 2D propagator.

The reference computes one shot per call and spreads shots over processes
(examples/time/inv_torch.py:123-143 with io/DistributedDataloader.py:7-12):
every call runs its own time loop, so one GPU advances one small grid per
launch.  ``Time2dBatch`` instead advances every shot of a batch by one time
step per launch (the C-ABI plan in include/torchwi_hip.h), keeps all state
device-resident, and replaces full virtual-source storage by checkpoints of
the wavefield every K steps plus recomputation when storage would not fit.

Each shot's result is computed exactly as a single ``Time2d`` call computes
it (same kernels, same per-cell arithmetic); the batch gradient is the sum
of the per-shot gradients in shot order.
"""
import ctypes
import math

import torch

from . import _lib

_ALIGN = 32  # row stride alignment in floats (128 B)


def round_up(n, m):
    return (n + m - 1) // m * m


def default_ckpt_interval(nt):
    """About sqrt(nt) steps per checkpoint segment."""
    return max(2, int(math.ceil(math.sqrt(nt))))


def _as_list(x, name):
    if isinstance(x, torch.Tensor):
        return [float(v) for v in x.detach().reshape(-1).cpu().tolist()]
    if isinstance(x, (int, float)):
        return [float(x)]
    return [float(v) for v in x]


class BatchState:
    """Device buffers and the C plan for one set of shots."""

    def __init__(self, geom, sx, sy, ry, ckpt_every, device, rx0=None, nrec=None, tile_rows=0):
        g = geom
        S = len(sx)
        self.S = S
        self.K = int(ckpt_every)
        self.nt = g.nt
        fshape = (g.dimy, g.ld)
        f = g.dimy * g.ld
        self.fstride = f
        dev = device
        zeros = lambda *s: torch.zeros(*s, dtype=torch.float32, device=dev)  # noqa: E731
        self.vel = zeros(fshape)
        self.vm = zeros(fshape)
        self.vc3 = zeros(fshape)
        self.u = zeros(S, 3, f)
        self.lam = zeros(S, 3, f) if self.K > 0 else self.u
        nrec = g.nx if nrec is None else nrec
        rx0 = [0] * S if rx0 is None else rx0
        self.nrec = nrec
        self.frd = zeros(S, g.nt, nrec)
        self.resid = zeros(S, g.nt, nrec)
        nslots = g.nt if self.K == 0 else self.K
        self.nseg = 0 if self.K == 0 else (g.nt + self.K - 1) // self.K
        self.virt = torch.empty(S, nslots, f, dtype=torch.float32, device=dev)
        self.ckpt = torch.empty(S, 2 * self.nseg, f, dtype=torch.float32, device=dev) if self.K else None
        self.grad = zeros(S, f)
        self.w = g.w_dev

        spec = _lib.TwPlanSpec(order=g.order, nt=g.nt, h=g.h, dt=g.dt, ckpt_every=self.K,
                               tile_rows=int(tile_rows))
        shots = (_lib.TwShotSpec * S)()
        for s in range(S):
            q = shots[s]
            q.dimx, q.dimy, q.ld = g.dimx, g.dimy, g.ld
            q.sx, q.sy, q.ry = sx[s], sy[s], ry[s]
            q.rx0, q.nrec = rx0[s], nrec
            q.vel = self.vel.data_ptr()
            q.vm = self.vm.data_ptr()
            q.vc3 = self.vc3.data_ptr()
            q.w = self.w.data_ptr()
            q.u = self.u[s].data_ptr()
            q.lam = self.lam[s].data_ptr()
            q.frd = self.frd[s].data_ptr()
            q.resid = self.resid[s].data_ptr()
            q.virt = self.virt[s].data_ptr()
            q.ckpt = self.ckpt[s].data_ptr() if self.K else None
            q.grad = self.grad[s].data_ptr()
        L = _lib.lib()
        nbytes = L.tw_plan_workspace_bytes(ctypes.byref(spec), shots, S)
        if nbytes == 0:
            raise RuntimeError("torchwi_amd: invalid shot batch: %s" % _lib.last_error())
        self.workspace = torch.empty(nbytes, dtype=torch.uint8, device=dev)
        handle = ctypes.c_void_p()
        rc = L.tw_plan_create(ctypes.byref(spec), shots, S, _lib.ptr(self.workspace), nbytes,
                              _lib.stream_of(self.workspace), ctypes.byref(handle))
        _lib.check(rc, "tw_plan_create")
        self.plan = handle

    def __del__(self):
        try:
            if getattr(self, "plan", None):
                _lib.lib().tw_plan_destroy(self.plan)
                self.plan = None
        except Exception:
            pass

    def forward(self):
        _lib.check(_lib.lib().tw_plan_forward(self.plan, _lib.stream_of(self.u)), "tw_plan_forward")

    def set_timing(self, enable):
        _lib.check(_lib.lib().tw_plan_set_timing(self.plan, int(bool(enable))), "tw_plan_set_timing")

    def timing(self, reset=False):
        """{kind: (total ms, launches)} of HIP-event-timed step kernels."""
        n = len(_lib.KINDS)
        ms = (ctypes.c_double * n)()
        cnt = (ctypes.c_longlong * n)()
        _lib.check(_lib.lib().tw_plan_timing(self.plan, int(bool(reset)), ms, cnt), "tw_plan_timing")
        return {k: (ms[i], cnt[i]) for i, k in enumerate(_lib.KINDS)}

    def backward(self, accumulate=False):
        _lib.check(_lib.lib().tw_plan_backward(self.plan, int(bool(accumulate)),
                                               _lib.stream_of(self.u)), "tw_plan_backward")


class _Geom:
    def __init__(self, nx, ny, h, w, dt, order, device):
        if order not in (2, 4, 8):
            raise ValueError("Wrong order: use 2, 4, and 8")
        self.nx, self.ny, self.order, self.ne = nx, ny, order, order // 2
        self.dimx, self.dimy = nx + order, ny + order
        self.ld = round_up(self.dimx, _ALIGN)
        self.h, self.dt = float(h), float(dt)
        w = torch.as_tensor(w)
        if w.dtype != torch.float32:
            raise RuntimeError("w: expected scalar type Float but found %s" % w.dtype)
        self.nt = int(w.numel())
        self.w_dev = w.to(device).contiguous()


def state_bytes(geom, S, ckpt_every):
    """Device bytes a BatchState of S shots needs."""
    f = geom.dimy * geom.ld
    K = int(ckpt_every)
    nseg = 0 if K == 0 else (geom.nt + K - 1) // K
    slots = geom.nt if K == 0 else K
    fields = S * (3 + (3 if K else 0) + slots + 2 * nseg + 1) + 3
    return 4 * (fields * f + 2 * S * geom.nt * geom.nx)


class _BatchOperator(torch.autograd.Function):

    @staticmethod
    def forward(ctx, vel, args):
        sx, sy, ry, m = args
        st = m.prepare(vel, sx, sy, ry)
        st.forward()
        ctx.m, ctx.st = m, st
        return st.frd

    @staticmethod
    def backward(ctx, grad_output):
        m, st = ctx.m, ctx.st
        st.resid.copy_(grad_output)
        st.backward()
        return m.reduce_grad(st), None


class Time2dBatch(torch.nn.Module):
    """Many shots of one velocity model, marched in lockstep on one GPU.

    ``forward(vel, sx, sy, ry)`` takes vel (nx, ny) and per-shot positions
    (sequences or 1-D tensors of length S) and returns the gathers
    (S, nt, nx); its backward gives the summed gradient (nx, ny).
    ``ckpt_every``: 0 = full virtual-source storage, K > 0 = checkpoint
    every K steps and recompute, "auto" = full storage when it fits in
    ``hbm_fraction`` of free device memory, else K = ceil(sqrt(nt)).
    """

    def __init__(self, nx, ny, h, w, dt, order, device="cuda", ckpt_every="auto",
                 hbm_fraction=0.8, tile_rows=0):
        super().__init__()
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise NotImplementedError("Time2dBatch runs on the GPU only")
        self.g = _Geom(nx, ny, h, w, dt, order, self.device)
        self.nt, self.nx, self.ny, self.order = self.g.nt, nx, ny, order
        self.ckpt_every = ckpt_every
        self.hbm_fraction = hbm_fraction
        self.tile_rows = tile_rows
        self._state = None
        self._key = None

    def choose_ckpt(self, S):
        if self.ckpt_every != "auto":
            return int(self.ckpt_every)
        free, _ = torch.cuda.mem_get_info(self.device)
        if self._state is not None:
            free += state_bytes(self.g, self._state.S, self._state.K)
        if state_bytes(self.g, S, 0) <= self.hbm_fraction * free:
            return 0
        return default_ckpt_interval(self.nt)

    def prepare(self, vel, sx, sy, ry):
        sx, sy, ry = _as_list(sx, "sx"), _as_list(sy, "sy"), _as_list(ry, "ry")
        if not (len(sx) == len(sy) == len(ry)) or len(sx) == 0:
            raise ValueError("sx, sy, ry must have the same positive length")
        key = (tuple(sx), tuple(sy), tuple(ry))
        if self._key != key:
            K = self.choose_ckpt(len(sx))
            self._state = None
            torch.cuda.empty_cache()
            self._state = BatchState(self.g, sx, sy, ry, K, self.device, tile_rows=self.tile_rows)
            self._key = key
        st = self._state
        self.load_model(vel, st)
        return st

    def load_model(self, vel, st):
        g = self.g
        if tuple(vel.shape) != (g.nx, g.ny):
            raise ValueError("vel must be (nx, ny) = (%d, %d), got %s" % (g.nx, g.ny, tuple(vel.shape)))
        with torch.no_grad():
            vp = torch.nn.functional.pad(
                vel.detach().to(self.device, torch.float32).transpose(0, 1).reshape(1, 1, g.ny, g.nx),
                (g.ne, g.ne, g.ne, g.ne), mode="replicate").reshape(g.dimy, g.dimx)
            st.vel[:, :g.dimx].copy_(vp)

    def reduce_grad(self, st):
        """Sum the per-shot padded gradients in shot order, crop to the model
        and transpose to (nx, ny) (TimeFunction.py:141-145)."""
        g = self.g
        tot = st.grad[0].clone()
        for s in range(1, st.S):
            tot.add_(st.grad[s])
        tot = tot.view(g.dimy, g.ld)[g.ne:g.ne + g.ny, g.ne:g.ne + g.nx]
        return tot.transpose(0, 1)

    def forward(self, vel, sx, sy, ry):
        return _BatchOperator.apply(vel, (sx, sy, ry, self))

    @torch.no_grad()
    def misfit_gradient(self, vel, sx, sy, ry, d_obs):
        """One FWI gradient: forward all shots, residual frd - d_obs, L2 loss
        0.5*||frd - d_obs||^2, adjoint, summed gradient.  Returns (loss,
        grad (nx, ny))."""
        st = self.prepare(vel, sx, sy, ry)
        st.forward()
        torch.sub(st.frd, d_obs, out=st.resid)
        loss = 0.5 * torch.sum(st.resid * st.resid)
        st.backward()
        return loss, self.reduce_grad(st)

    @torch.no_grad()
    def modeling(self, vel, sx, sy, ry):
        st = self.prepare(vel, sx, sy, ry)
        st.forward()
        return st.frd
