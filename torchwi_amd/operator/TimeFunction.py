"""Backend selection and the autograd Functions of the time-domain operator.

Mirrors torchwi/operator/TimeFunction.py: ``get_operator(device, prop, exa,
forward_only)`` returns ``Function.apply`` taking ``(vel, (sx, sy, ry, module))``
and reading/writing the module's buffers.  The GPU backends are the gfx950
modules in torchwi_amd.propagator.  There is no CPU path in this package: a
CPU device raises NotImplementedError instead of falling back.

As in the reference, the virtual sources (or EXA arrays) saved for backward
are the module's own buffers (TimeFunction.py:128-130), so a second forward
on the same module overwrites what the first one's backward needs.  The
reference then silently returns a wrong gradient; here every forward stamps
the module with a generation number and a backward whose forward is no
longer the module's latest raises instead.
"""
import torch


def _is_gpu(device):
    return torch.device(device).type == "cuda"


def _get_propagator(device, prop="ext", exa=False, forward_only=False):
    """Reference: TimeFunction.py:3-29."""
    if not _is_gpu(device):
        raise NotImplementedError(
            "torchwi_amd runs the time-domain propagator on the GPU only (device=%r)" % (device,))
    if prop not in ("ext", "numba"):
        raise NotImplementedError("Check prop=[ext|numba]")
    if forward_only:
        from torchwi_amd.propagator import td2d_forward_cuda as p
    elif exa:
        from torchwi_amd.propagator import td2d_exa_cuda as p
    else:
        from torchwi_amd.propagator import td2d_cuda as p
    return p


def _crop_grad(grad, m):
    """Padded-grid gradient (dimy*dimx) -> (nx_org, ny_org), TimeFunction.py:141-145."""
    g = grad.view(m.dimy, m.dimx)[m.ne:m.ne + m.ny_org, m.ne:m.ne + m.nx_org]
    return g.transpose(0, 1)


def _stamp(ctx, m):
    m._fwd_generation = getattr(m, "_fwd_generation", 0) + 1
    ctx.generation = m._fwd_generation


def _check_stamp(ctx, m):
    if getattr(m, "_fwd_generation", 0) != ctx.generation:
        raise RuntimeError("the module ran another forward after this one: its saved wavefield storage "
                           "was overwritten; call backward before the module's next forward")


def get_operator(device, prop="ext", exa=False, forward_only=False):
    """Reference: TimeFunction.py:32-148."""
    _propagator = _get_propagator(device, prop, exa, forward_only)

    if forward_only:

        class TimeForwardOperator(torch.autograd.Function):

            @staticmethod
            def forward(ctx, vel, args):
                sx, sy, ry, m = args
                _propagator.forward(m.frd, m.u1, m.u2, m.u3, m.vpad, m.w,
                                    m.order, m.dimx, m.dimy, m.nt, m.h, m.dt,
                                    sx.item(), sy.item(), ry.item())
                return m.frd.view(m.nt, m.nx)[:, :m.nx_org]

            @staticmethod
            def backward(ctx, grad_output):
                return None, None

        return TimeForwardOperator.apply

    if exa:

        class TimeEXAOperator(torch.autograd.Function):

            @staticmethod
            def forward(ctx, vel, args):
                sx, sy, ry, m = args
                _propagator.forward(m.frd, m.exa, m.iexa, m.u1, m.u2, m.u3, m.vpad, m.w,
                                    m.order, m.dimx, m.dimy, m.nt, m.h, m.dt,
                                    sx.item(), sy.item(), ry.item())
                ctx.model = m
                _stamp(ctx, m)
                ctx.save_for_backward(m.exa, m.iexa, ry)
                return m.frd.view(m.nt, m.nx)[:, :m.nx_org]

            @staticmethod
            def backward(ctx, grad_output):
                exa, iexa, ry = ctx.saved_tensors
                m = ctx.model
                _check_stamp(ctx, m)
                grad = _propagator.backward(exa, iexa, m.u1, m.u2, m.u3, m.vpad, grad_output,
                                            m.order, m.dimx, m.dimy, m.nt, m.h, m.dt, ry.item())
                return _crop_grad(grad, m), None

        return TimeEXAOperator.apply

    class TimeOperator(torch.autograd.Function):

        @staticmethod
        def forward(ctx, vel, args):
            sx, sy, ry, m = args
            _propagator.forward(m.frd, m.virt, m.u1, m.u2, m.u3, m.vpad, m.w,
                                m.order, m.dimx, m.dimy, m.nt, m.h, m.dt,
                                sx.item(), sy.item(), ry.item())
            ctx.model = m
            _stamp(ctx, m)
            ctx.save_for_backward(m.virt, ry)
            return m.frd.view(m.nt, m.nx)[:, :m.nx_org]

        @staticmethod
        def backward(ctx, grad_output):
            virt, ry = ctx.saved_tensors
            m = ctx.model
            _check_stamp(ctx, m)
            grad = _propagator.backward(virt, m.u1, m.u2, m.u3, m.vpad, grad_output,
                                        m.order, m.dimx, m.dimy, m.nt, m.h, m.dt, ry.item())
            return _crop_grad(grad, m), None

    return TimeOperator.apply
