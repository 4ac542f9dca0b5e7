"""Time-domain shot datasets and shot-gather files.

``TimeDataset`` / ``TimeForwardDataset`` follow the reference's datasets
(torchwi/io/Dataset.py:6-66): items are ``(sx, sy, data[, index])`` and
``(sx, sy[, index])``; observed data are read per shot from ``ftrue % index``
as raw float32 ('bin'/'binary') or ``.npy`` ('npy').  The reference's 'su'
reader depends on an external package (pkrh) and is not provided; asking for
it raises.  Deviation: a ``shape`` given to ``TimeDataset`` is applied (the
reference assigns an undefined name there, Dataset.py:43).

``write_gather`` / ``read_gathers`` are the per-shot gather files the
reference's forward-modelling example writes (``seismo.%04d``: the (nt, nx)
float32 gather, ``tofile``; examples/time/forward_modeling.py:52), stacked
for the batched engine.
"""
import numpy as np
import torch


class TimeDataset(torch.utils.data.Dataset):
    def __init__(self, sxy, ftrue, with_index=False, data_type='bin', shape=None, transpose=False):
        super().__init__()
        self.ftrue = ftrue
        self.with_index = with_index
        self.sxy = sxy
        self.data_type = data_type
        self.shape = shape
        self.transpose = transpose

    def __len__(self):
        return len(self.sxy)

    def __getitem__(self, index):
        truedata = self.load_data(index)
        sx, sy = self.sxy[index]
        if self.with_index:
            return sx, sy, truedata, index
        return sx, sy, truedata

    def load_data(self, index):
        if self.data_type in ('bin', 'binary'):
            truedata = self.load_bin(index)
        elif self.data_type == 'npy':
            truedata = self.load_npy(index)
        elif self.data_type == 'su':
            raise NotImplementedError("SU gathers need the external 'pkrh' reader, which is not part of this package")
        else:
            raise ValueError("unknown data_type %r (bin, binary, npy)" % (self.data_type,))
        if self.transpose:
            truedata = truedata.T
        return torch.from_numpy(np.ascontiguousarray(truedata))

    def load_bin(self, index):
        truedata = np.fromfile(self.ftrue % index, dtype=np.float32)
        if self.shape:
            truedata.shape = self.shape
        return truedata

    def load_npy(self, index):
        return np.load(self.ftrue % index, allow_pickle=False)


class TimeForwardDataset(torch.utils.data.Dataset):
    def __init__(self, sxy, with_index=True):
        super().__init__()
        self.with_index = with_index
        self.sxy = sxy
        self.nshot = len(sxy)

    def __len__(self):
        return self.nshot

    def __getitem__(self, index):
        sx, sy = self.sxy[index]
        if self.with_index:
            return sx, sy, index
        return sx, sy


def write_gather(pattern, index, gather):
    """Write one shot's (nt, nx) gather as raw float32 to ``pattern % index``."""
    g = gather.detach().to('cpu').numpy() if isinstance(gather, torch.Tensor) else np.asarray(gather)
    np.ascontiguousarray(g, dtype=np.float32).tofile(pattern % index)


def read_gathers(pattern, indices, nt, nrec, device=None):
    """Stack the raw float32 gathers ``pattern % i`` (each nt x nrec) of the
    given shots into one (S, nt, nrec) tensor, on ``device`` if given."""
    out = np.empty((len(indices), nt, nrec), np.float32)
    for s, i in enumerate(indices):
        g = np.fromfile(pattern % i, dtype=np.float32)
        if g.size != nt * nrec:
            raise ValueError("%s holds %d values, expected nt*nrec = %d" % (pattern % i, g.size, nt * nrec))
        out[s] = g.reshape(nt, nrec)
    t = torch.from_numpy(out)
    return t.to(device) if device is not None else t
