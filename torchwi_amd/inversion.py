"""Full-waveform inversion on one GPU: the caller side of the propagator.

The reference's inversion script (examples/time/inv_torch.py:16-143) runs
one shot per process call and sums gradients over Horovod ranks; an epoch
is: zero the gradient, for every shot model its gather, form
0.5 * ||frd - d_obs||^2 and back-propagate, then normalise the summed
gradient by the infinity norm of the first epoch's gradient and take an
optimiser step (Adam, Nesterov, momentum or plain gradient descent;
inv_torch.py:42-56, 75-81).  ``TimeInversion`` runs the same epoch with every
shot of the model in one batched, device-resident misfit-gradient call
(``Time2dBatch``): the loss, the residual, the gradient through the
parameterisation, the normalisation and the optimiser step stay on the GPU,
and nothing is copied to the host unless the caller asks for a number.

The summed gradient is the per-shot gradients added in shot order, as the
reference's autograd accumulation adds them, so one epoch moves the
parameter exactly as the reference's per-shot loop over the drop-in
``Time2d`` module does.
"""
import torch

from .batch import Time2dBatch
from .parameter import SlothParameter, SlownessParameter, VelocityParameter

__all__ = ["TimeInversion"]

_PARAMETERS = {"velocity": VelocityParameter, "sloth": SlothParameter, "slowness": SlownessParameter}


class TimeInversion(torch.nn.Module):
    """One model, S shots, one batched gradient per epoch.

    vinit (nx, ny) initial velocity; sx, sy, ry per-shot positions (length S);
    w (nt,) float32 wavelet; vmin, vmax clipping bounds (both or neither);
    parameter: "velocity" | "sloth" | "slowness"; optimizer: "adam" | "nag" |
    "momentum" | "gd" with lr and momentum as in inv_torch.py:42-56;
    grad_norm: "first" (divide by the first epoch's infinity norm) | "none".
    Remaining keyword arguments go to ``Time2dBatch`` (ckpt_every,
    time_block, graph_steps, ...).
    """

    def __init__(self, vinit, sx, sy, ry, h, w, dt, order, device="cuda", vmin=None, vmax=None,
                 parameter="velocity", optimizer="adam", lr=0.02, momentum=0.9, grad_norm="first",
                 **batch_kw):
        super().__init__()
        if grad_norm not in ("first", "none"):
            raise ValueError("grad_norm must be 'first' or 'none'")
        if parameter not in _PARAMETERS:
            raise ValueError("parameter must be one of %s" % sorted(_PARAMETERS))
        nx, ny = vinit.shape
        self.sx, self.sy, self.ry = sx, sy, ry
        self.grad_norm_mode = grad_norm
        # a copy: the parameterisations keep the tensor they are given as their
        # parameter (Param.py:8), and the caller's initial model must survive
        v0 = torch.as_tensor(vinit, dtype=torch.float32).to(device).clone()
        self.velgen = _PARAMETERS[parameter](v0, vmin, vmax)
        self.modeling = Time2dBatch(nx, ny, h, w, dt, order, device, **batch_kw)
        self.gnorm0 = None
        params = self.velgen.parameters()
        if optimizer == "adam":
            self.optimizer = torch.optim.Adam(params, lr=lr)
        elif optimizer == "nag":
            self.optimizer = torch.optim.SGD(params, lr=lr, momentum=momentum, nesterov=True)
        elif optimizer == "momentum":
            self.optimizer = torch.optim.SGD(params, lr=lr, momentum=momentum)
        elif optimizer == "gd":
            self.optimizer = torch.optim.SGD(params, lr=lr)
        else:
            raise ValueError("optimizer must be adam, nag, momentum or gd")

    def velocity(self):
        return self.velgen()

    def gradient(self):
        return self.velgen.gradient()

    def forward(self):
        """Gathers (S, nt, nx) of the current model (autograd-enabled)."""
        return self.modeling(self.velgen(), self.sx, self.sy, self.ry)

    def norm_grad(self):
        if self.grad_norm_mode == "none":
            return
        if self.gnorm0 is None:
            self.gnorm0 = self.velgen.gradient().norm(float("inf"))
        self.velgen.normalize_gradient(self.gnorm0)

    def epoch(self, d_obs):
        """One inversion epoch against observed gathers d_obs (S, nt, nx).
        Returns the loss 0.5 * ||frd - d_obs||^2 as a device tensor."""
        self.optimizer.zero_grad()
        frd = self.forward()
        loss = 0.5 * torch.sum((frd - d_obs) ** 2)
        loss.backward()
        self.norm_grad()
        self.optimizer.step()
        return loss.detach()
