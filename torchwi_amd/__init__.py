"""torchwi_amd -- MI355X-native time-domain 2D acoustic propagator and its
adjoint gradient, behind TorchWI's operator surface.

    import torchwi_amd as torchwi
    model = torchwi.Time2d(nx, ny, h, w, dt, order, 'cuda')
    frd = model(vel, sx, sy, ry)            # (nt, nx) shot gather
    loss = 0.5 * torch.sum((frd - true) ** 2)
    loss.backward()                          # vel.grad (nx, ny)

``Time2dBatch`` runs many shots on one GPU in one plan (the replacement for
the reference's per-process shot loop), with wavefield checkpointing;
``Time2dForwardBatch`` models many shots without storing anything for an
adjoint.
"""
from torchwi_amd.operator.TimeOperator import Time2d, Time2dForward  # noqa: F401
from torchwi_amd.batch import Time2dBatch, Time2dForwardBatch, Time2dWindows  # noqa: F401
from torchwi_amd.inversion import TimeInversion  # noqa: F401

__all__ = ["Time2d", "Time2dForward", "Time2dBatch", "Time2dForwardBatch", "Time2dWindows", "TimeInversion"]
