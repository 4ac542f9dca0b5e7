"""nn.Module front end of the time-domain 2D acoustic propagator.

Same classes, constructor signatures, buffers and ``forward(vel, sx, sy, ry)``
as torchwi/operator/TimeOperator.py.  One deliberate difference: the
reference rounds nx, ny up to a multiple of 16 on CUDA (TimeOperator.py:36-38)
because its kernels assume whole 16x16 tiles, which moves the Keys boundary
and changes the physics; the gfx950 kernels handle any grid, so the grid is
always the caller's (the reference CPU semantics).
"""
import torch

from . import TimeFunction


class Time2dForward(torch.nn.Module):
    def __init__(self, nx, ny, h, w, dt, order, device):
        """vel and grad are (nx, ny), y fastest; the padded model vpad is
        (dimy, dimx), x fastest.  Reference: TimeOperator.py:19-53."""
        super().__init__()
        self.device = device
        self.h = h
        self.w = w.to("cpu")  # wavelet kept on the host, as the reference does
        self.nt = len(w)
        self.dt = dt
        self.order = order
        self.ne = order // 2
        self.nx_org, self.ny_org = nx, ny
        self.nx = self.nx_org
        self.ny = self.ny_org
        self.dimx = self.nx + self.order
        self.dimy = self.ny + self.order
        self.dimxy = self.dimx * self.dimy

        self.time_modeling = TimeFunction.get_operator(self.device, forward_only=True)

        self.u1 = torch.zeros(self.dimxy, device=self.device)
        self.u2 = torch.zeros(self.dimxy, device=self.device)
        self.u3 = torch.zeros(self.dimxy, device=self.device)
        self.frd = torch.zeros((self.nt * self.nx), device=self.device)

    def pad_vel(self, vel):
        """Edge-replicate vel (nx, ny) into vpad (dimy*dimx,).  Reference:
        TimeOperator.py:55-58."""
        pad_shape = (self.ne, self.nx - self.nx_org + self.ne, self.ne, self.ny - self.ny_org + self.ne)
        with torch.no_grad():
            self.vpad = torch.nn.ReplicationPad2d(pad_shape)(
                vel.transpose(0, 1).reshape(1, 1, self.ny_org, self.nx_org)).reshape(-1).contiguous()

    def forward(self, vel, sx, sy, ry):
        self.pad_vel(vel)
        return self.time_modeling(vel, (sx, sy, ry, self))


class Time2d(Time2dForward):
    def __init__(self, nx, ny, h, w, dt, order, device, prop="ext", exa=False):
        """Reference: TimeOperator.py:69-82.  Full storage keeps nt*dimx*dimy
        virtual-source samples; exa=True keeps one amplitude and one int16 step
        per cell instead."""
        super(Time2d, self).__init__(nx, ny, h, w, dt, order, device)
        self.time_modeling = TimeFunction.get_operator(self.device, prop, exa)
        if exa:
            self.exa = torch.zeros(self.dimxy, device=self.device)
            self.iexa = torch.zeros(self.dimxy, dtype=torch.int16, device=self.device)
        else:
            self.virt = torch.zeros((self.nt * self.dimxy), device=self.device)
