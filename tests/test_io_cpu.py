"""RSF files and shot-gather files (CPU).

The header text a writer emits is RSF's canonical form (the one the
reference's writer produces too, torchwi/io/rsf.py:369-393): ``in``,
``esize``, ``data_format`` on lines of their own (string values quoted),
then one line per axis ``n# o# d#`` with ``label#`` / ``unit#`` only when
set.  A reader takes the last ``key=value`` of each key anywhere in the text
(rsf.py:313-324), which is how Madagascar headers accumulate history; the
data file is read n1 fastest.
"""
import os

import numpy as np
import pytest
import torch

from torchwi_amd.io import TimeDataset, TimeForwardDataset, read_gathers, rsf, write_gather


def test_tofile_header_text_and_round_trip(tmp_path):
    vel = np.arange(6 * 4, dtype=np.float32).reshape(6, 4) * 0.5 + 1500.0   # (nx, ny): n1 = ny fastest
    f = str(tmp_path / "vel.rsf")
    rsf.tofile(f, vel, n1=4, d1=10.0, label1="depth", unit1="m", n2=6, d2=12.5, o2=-25.0)
    head = open(f).read()
    want = ('in="%s@"\nesize=4\ndata_format="native_float"\n'
            'n1=4 o1=0.0 d1=10.0 label1="depth" unit1="m"\n'
            'n2=6 o2=-25.0 d2=12.5\n') % os.path.abspath(f)
    assert head == want
    raw = np.fromfile(f + "@", dtype="<f4")
    assert np.array_equal(raw, vel.ravel())
    ny, nx, h, data = rsf.fromfile(f, "n1 n2 d1 data")
    assert (ny, nx, h) == (4, 6, 10.0)
    assert data.dtype == np.float32 and data.shape == (6, 4) and np.array_equal(data, vel)
    with rsf.input(f) as r:
        assert r.shape == (6, 4) and r.dimension == 2 and r.size() == 24 and r.size(1) == 4
        assert r["label1"] == "depth" and r["o2"] == -25.0 and r.esize == 4
        tr = list(r.traces())
    assert len(tr) == 6 and all(np.array_equal(t, vel[i]) for i, t in enumerate(tr))


def test_header_parsing_last_key_wins_spaces_and_quotes(tmp_path):
    data = np.linspace(0, 1, 12, dtype=np.float32)
    d = tmp_path / "trace.rsf@"
    data.tofile(str(d))
    text = ("sfspike n1=5 d1=0.004 o1=0\n"              # an earlier program's history
            "\tn1=12  d1 = 0.002 o1=-0.01 label1='Time' unit1=\"s\"\n"
            "\tlabel2=\"two words\" n2=1\n"
            "\tin=\"%s\"\n\tdata_format=native_float esize=4\n") % d
    f = tmp_path / "trace.rsf"
    f.write_text(text)
    n1, d1, o1, lab, lab2, x = rsf.fromfile(str(f), ["n1", "d1", "o1", "label1", "label2", "data"])
    assert (n1, d1, o1, lab, lab2) == (12, 0.002, -0.01, "Time", "two words")
    assert x.shape == (1, 12) and np.array_equal(x[0], data)


@pytest.mark.parametrize("typ,dt", [("int", np.int32), ("complex", np.complex64)])
def test_other_types(tmp_path, typ, dt):
    a = (np.arange(10) - 3).astype(dt) * (1 + (1j if typ == "complex" else 0))
    f = str(tmp_path / "a.rsf")
    rsf.tofile(f, a, n1=10, data_format="native_" + typ)
    assert 'data_format="native_%s"' % typ in open(f).read()
    assert ('esize=%d' % (8 if typ == "complex" else 4)) in open(f).read()
    got = rsf.fromfile(f, "data")
    assert got.dtype == dt and np.array_equal(got, a)


def test_ascii_form(tmp_path):
    a = np.array([1.5, -2.25, 3.0], np.float32)
    f = str(tmp_path / "a.rsf")
    rsf.tofile(f, a, n1=3, data_format="ascii_float")
    assert "esize=0" in open(f).read()
    assert np.array_equal(rsf.fromfile(f, "data"), a)


def test_errors_raise(tmp_path):
    f = tmp_path / "noin.rsf"
    f.write_text("n1=3 d1=1\n")
    with pytest.raises(rsf.RsfError, match="in="):
        rsf.fromfile(str(f), "n1")
    f.write_text('n1=3 in="%s"\n' % (tmp_path / "missing@"))
    with pytest.raises(rsf.RsfError, match="does not exist"):
        rsf.fromfile(str(f), "n1")
    with pytest.raises(rsf.RsfError, match="unknown header key"):
        rsf.output(str(tmp_path / "x.rsf"), bogus=3)
    with pytest.raises(rsf.RsfError, match="data_format"):
        rsf.output(str(tmp_path / "x.rsf"), data_format="native_char")
    f.write_text('n1=x in="%s"\n' % f)
    with pytest.raises(rsf.RsfError, match="bad value"):
        rsf.fromfile(str(f), "n1")
    # a short data file is an error, not a silently truncated array
    rsf.tofile(str(tmp_path / "s.rsf"), np.zeros(4, np.float32), n1=4)
    txt = (tmp_path / "s.rsf").read_text().replace("n1=4", "n1=5")
    (tmp_path / "s.rsf").write_text(txt)
    with pytest.raises(rsf.RsfError, match="holds 4 of 5"):
        rsf.fromfile(str(tmp_path / "s.rsf"), "data")


def test_shot_gathers_and_datasets(tmp_path):
    nt, nx, S = 7, 5, 3
    rng = np.random.default_rng(3)
    g = rng.standard_normal((S, nt, nx)).astype(np.float32)
    pat = str(tmp_path / "seismo.%04d")
    for s in range(S):
        write_gather(pat, s, torch.from_numpy(g[s]))
    assert os.path.getsize(pat % 2) == nt * nx * 4
    assert torch.equal(read_gathers(pat, [2, 0], nt, nx), torch.from_numpy(g[[2, 0]]))
    with pytest.raises(ValueError):
        read_gathers(pat, [0], nt + 1, nx)
    sxy = [(10.0 * s, 20.0) for s in range(S)]
    ds = TimeDataset(sxy, pat, with_index=True, shape=(nt, nx))
    sx, sy, d, i = ds[1]
    assert (sx, sy, i) == (10.0, 20.0, 1) and torch.equal(d, torch.from_numpy(g[1]))
    assert TimeDataset(sxy, pat, transpose=True, shape=(nt, nx))[0][2].shape == (nx, nt)
    np.save(str(tmp_path / "obs0.npy"), g[0])
    assert torch.equal(TimeDataset(sxy, str(tmp_path / "obs%d.npy"), data_type="npy")[0][2], torch.from_numpy(g[0]))
    with pytest.raises(NotImplementedError):
        TimeDataset(sxy, pat, data_type="su")[0]
    fw = TimeForwardDataset(sxy)
    assert len(fw) == S and fw[2] == (20.0, 20.0, 2)
