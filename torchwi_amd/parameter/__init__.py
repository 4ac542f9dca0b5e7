"""Model parameterisations for inversion (the reference's torchwi.parameter).

A parameter module owns the optimised tensor ``par`` (a torch Parameter in
the parameterisation's own units) and returns the velocity from it; the
velocity then feeds the propagator, and autograd carries the gradient back
through the change of variables.  Names, constructor arguments and methods
follow torchwi/parameter/Param.py:3-71:

* ``VelocityParameter``  par = v
* ``SlothParameter``     par = 1 / v**2, v = sqrt(1 / par)
* ``SlownessParameter``  par = 1 / v,    v = 1 / par

With both ``vmin`` and ``vmax`` given (truthy, as in the reference), every
forward call first clamps ``par`` in place to the parameter-space image of
[vmin, vmax] (Param.py:9-11, 34-41).
"""
import torch

__all__ = ["VelocityParameter", "SlothParameter", "SlownessParameter"]


class _Parameterisation(torch.nn.Module):
    """Common part of the parameterisations (Param.py:3-44)."""

    def __init__(self, vel, vmin=None, vmax=None):
        super().__init__()
        self.vmin, self.vmax = vmin, vmax
        self.par = torch.nn.Parameter(self.vel_to_par(vel))
        self.pmin = self.pmax = None
        if vmin and vmax:
            a, b = self.vel_to_par(vmin), self.vel_to_par(vmax)
            self.pmin, self.pmax = min(a, b), max(a, b)

    # change of variables (identity here; overridden below)
    def vel_to_par(self, vel):
        return vel

    def par_to_vel(self, par):
        return par

    def vel(self):
        return self.par_to_vel(self.par)

    # gradient helpers used by the inversion loop (inv_torch.py:75-81)
    def gradient(self):
        return self.par.grad

    def gradient_times(self, multiplier):
        self.par.grad *= multiplier

    def grad_norm(self):
        return self.par.grad.norm(float("inf"))

    def normalize_gradient(self, gnorm0):
        self.par.grad /= gnorm0

    def pclip(self):
        if self.vmin and self.vmax:
            with torch.no_grad():
                self.par[:] = torch.clamp(self.par, self.pmin, self.pmax)

    def forward(self):
        self.pclip()
        return self.vel()

    def report(self):
        return "Parameter: %s" % type(self).__name__


class VelocityParameter(_Parameterisation):
    """par = v (Param.py:47-49)."""


class SlothParameter(_Parameterisation):
    """par = 1 / v**2 (Param.py:52-60)."""

    def vel_to_par(self, vel):
        return 1.0 / vel ** 2

    def par_to_vel(self, par):
        return torch.sqrt(1.0 / par)


class SlownessParameter(_Parameterisation):
    """par = 1 / v (Param.py:63-71)."""

    def vel_to_par(self, vel):
        return 1.0 / vel

    def par_to_vel(self, par):
        return 1.0 / par
