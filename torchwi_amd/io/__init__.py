"""On-disk formats on either side of the propagator: RSF files (velocity
models, wavelets, shot tables) and per-shot gather files, with the
reference's time-domain datasets (torchwi/io/__init__.py)."""
from . import rsf
from .dataset import TimeDataset, TimeForwardDataset, read_gathers, write_gather

__all__ = ["rsf", "TimeDataset", "TimeForwardDataset", "read_gathers", "write_gather"]
