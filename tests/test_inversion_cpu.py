"""Parameterisations (torchwi_amd.parameter) and the inversion driver's host
logic, on the CPU.  The parameter classes follow the reference's
torchwi/parameter/Param.py:3-71; the checks are the closed forms of each
change of variables and of its derivative."""
import pytest
import torch

from torchwi_amd.parameter import SlothParameter, SlownessParameter, VelocityParameter


def test_velocity_parameter_clips_in_place_and_returns_the_parameter():
    v = torch.tensor([[1.0, 2.0], [6.0, 3.0]])
    p = VelocityParameter(v, 1.5, 5.5)
    assert (p.pmin, p.pmax) == (1.5, 5.5)
    out = p()
    assert torch.equal(out, torch.tensor([[1.5, 2.0], [5.5, 3.0]]))
    assert torch.equal(p.par.data, out) and out is p.par
    assert p.report() == "Parameter: VelocityParameter"
    q = VelocityParameter(v)            # no bounds: no clipping
    assert torch.equal(q(), v)
    r = VelocityParameter(v, 0.0, 5.5)  # falsy bound: no clipping, as in the reference
    assert torch.equal(r(), v)


@pytest.mark.parametrize("cls,to_par,dv_dpar", [
    (SlothParameter, lambda v: 1.0 / v ** 2, lambda p: -0.5 * p ** -1.5),
    (SlownessParameter, lambda v: 1.0 / v, lambda p: -1.0 / p ** 2),
])
def test_reciprocal_parameterisations(cls, to_par, dv_dpar):
    v = torch.tensor([1.0, 2.0, 4.0, 8.0], dtype=torch.float64)
    p = cls(v, 1.5, 5.5)
    assert torch.allclose(p.par.data, to_par(v))
    lo, hi = sorted([to_par(torch.tensor(1.5, dtype=torch.float64)).item(),
                     to_par(torch.tensor(5.5, dtype=torch.float64)).item()])
    assert p.pmin == pytest.approx(lo) and p.pmax == pytest.approx(hi)
    vel = p()                           # clamps par to [pmin, pmax], i.e. v to [1.5, 5.5]
    assert torch.allclose(vel, torch.tensor([1.5, 2.0, 4.0, 5.5], dtype=torch.float64))
    vel.sum().backward()                # chain rule through the change of variables
    assert torch.allclose(p.gradient(), dv_dpar(p.par.data))


def test_gradient_helpers():
    p = VelocityParameter(torch.ones(3))
    p.par.grad = torch.tensor([2.0, -8.0, 4.0])
    assert p.grad_norm().item() == 8.0
    p.gradient_times(0.5)
    assert torch.equal(p.gradient(), torch.tensor([1.0, -4.0, 2.0]))
    p.normalize_gradient(p.grad_norm())
    assert torch.equal(p.gradient(), torch.tensor([0.25, -1.0, 0.5]))


def test_inversion_refuses_bad_choices_and_the_cpu():
    from torchwi_amd.inversion import TimeInversion
    v = torch.full((20, 10), 2000.0)
    w = torch.zeros(30)
    args = (v, [50.0], [15.0], [15.0], 10.0, w, 1e-3, 8)
    with pytest.raises(NotImplementedError):
        TimeInversion(*args, device="cpu")
    with pytest.raises(ValueError):
        TimeInversion(*args, device="cpu", grad_norm="max")
    with pytest.raises(ValueError):
        TimeInversion(*args, device="cpu", parameter="impedance")
