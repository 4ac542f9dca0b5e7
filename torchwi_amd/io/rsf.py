"""RSF files (Madagascar's regularly sampled format): a text header plus the
data file it points to.

The format, as this module reads and writes it:

* The header is free text made of ``key=value`` tokens.  Tokens may sit
  anywhere, among the history lines earlier programs appended, and
  whitespace may surround ``=``.  A value is a double- or single-quoted
  string (which may hold blanks) or a run of non-blank characters.  When a
  key occurs more than once, its last occurrence is the one in force.
* ``in`` names the data file.  ``data_format`` is ``<form>_<type>``, with form
  ``native`` (raw machine-order binary) or ``ascii`` (a text table), and type
  ``float`` (4-byte), ``int`` (4-byte) or ``complex`` (2 x 4-byte).  ``esize``
  is the bytes per element (0 for ascii).
* Axis k (k = 1 is the fastest) is described by ``nk`` (samples), ``ok``
  (origin), ``dk`` (sampling), ``labelk`` and ``unitk``.  The data array has
  shape ``(n_N, ..., n2, n1)``, N being the highest axis with ``nN`` set.

A header written here is the canonical form that RSF tools and the
reference's writer emit: ``in``, ``esize`` and ``data_format`` on lines of
their own (strings quoted), then one line per axis ``nk= ok= dk=``, with
``labelk`` / ``unitk`` only when set.  Errors raise ``RsfError``.

Interface (the names TorchWI's examples call, torchwi/io/rsf.py):
``fromfile(path, keys)``, ``tofile(path, data, **keys)``, ``input(path)`` (a
reader: ``r[key]``, ``r.shape``, ``r.read()``, ``r.traces()``) and
``output(path, **keys)`` (a writer: ``w.write(data)``).
"""
import os
import re

import numpy as np

MAX_AXES = 9
_TYPES = {"float": np.dtype(np.float32), "int": np.dtype(np.int32), "complex": np.dtype(np.complex64)}
_FORMS = ("native", "ascii")
_AXIS_FIELDS = ("n", "o", "d", "label", "unit")
# a key, '=', then a quoted string or a run of non-blank characters
_TOKEN = re.compile(r"""([A-Za-z_][A-Za-z0-9_]*)\s*=\s*("[^"]*"|'[^']*'|\S+)""")
_AXIS_KEY = re.compile(r"^(n|o|d|label|unit)([1-9])$")


class RsfError(ValueError):
    """A malformed header, a missing data file or an unsupported key/format."""


def parse_header(text):
    """All key=value pairs of a header text as strings (quotes removed), the
    last occurrence of each key winning."""
    keys = {}
    for key, val in _TOKEN.findall(text):
        if len(val) >= 2 and val[0] == val[-1] and val[0] in "\"'":
            val = val[1:-1]
        keys[key] = val
    return keys


def _split_format(fmt):
    form, _, typ = str(fmt).lower().partition("_")
    if form not in _FORMS or typ not in _TYPES:
        raise RsfError("unsupported data_format %r (use native|ascii _ float|int|complex)" % (fmt,))
    return form, typ


def _axis_value(field, text):
    """Typed value of an axis field given as text."""
    if field == "n":
        return int(text)
    if field in ("o", "d"):
        return float(text)
    return str(text)


class Header:
    """A parsed header: the data file, its format and the axes."""

    def __init__(self, infile="", data_format="native_float", axes=None):
        self.infile = infile
        self.form, self.type = _split_format(data_format)
        # axes[k-1] = dict(n, o, d, label, unit) of axis k
        self.axes = [dict(n=0, o=0.0, d=0.0, label="", unit="") for _ in range(MAX_AXES)]
        for k, ax in enumerate(axes or []):
            self.axes[k].update(ax)

    @classmethod
    def from_text(cls, text):
        kv = parse_header(text)
        h = cls(kv.get("in", ""), kv.get("data_format", "native_float"))
        for key, val in kv.items():
            m = _AXIS_KEY.match(key)
            if m:
                try:
                    h.axes[int(m.group(2)) - 1][m.group(1)] = _axis_value(m.group(1), val)
                except ValueError:
                    raise RsfError("bad value %s=%s" % (key, val)) from None
        return h

    @property
    def data_format(self):
        return "%s_%s" % (self.form, self.type)

    @property
    def dtype(self):
        return _TYPES[self.type]

    @property
    def esize(self):
        return 0 if self.form == "ascii" else self.dtype.itemsize

    @property
    def ndim(self):
        return max((k + 1 for k, ax in enumerate(self.axes) if ax["n"] > 0), default=0)

    @property
    def shape(self):
        """(n_N, ..., n1): slowest axis first."""
        return tuple(self.axes[k]["n"] for k in range(self.ndim - 1, -1, -1))

    def value(self, key):
        if key == "in":
            return self.infile
        if key == "data_format":
            return self.data_format
        if key == "esize":
            return self.esize
        if key in ("form", "type"):
            return getattr(self, key)
        m = _AXIS_KEY.match(key)
        if not m:
            raise RsfError("unknown header key %r" % (key,))
        return self.axes[int(m.group(2)) - 1][m.group(1)]

    def text(self):
        """Canonical header text."""
        lines = ['in="%s"' % self.infile, "esize=%d" % self.esize, 'data_format="%s"' % self.data_format]
        for k in range(self.ndim):
            ax = self.axes[k]
            line = "n%d=%d o%d=%r d%d=%r" % (k + 1, ax["n"], k + 1, float(ax["o"]), k + 1, float(ax["d"]))
            for f in ("label", "unit"):
                if ax[f]:
                    line += ' %s%d="%s"' % (f, k + 1, ax[f])
            lines.append(line)
        return "\n".join(lines) + "\n"


class Reader:
    """An RSF file open for reading (``input``)."""

    def __init__(self, path):
        self.path = path
        with open(path) as f:
            self.header = Header.from_text(f.read())
        if not self.header.infile:
            raise RsfError("%s: the header has no in= (data file) entry" % path)
        if not os.path.exists(self.header.infile):
            raise RsfError("%s: data file %s does not exist" % (path, self.header.infile))
        self._pos = 0      # elements consumed by read(count) / traces()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        return False

    def close(self):
        pass

    def __getitem__(self, key):
        return self.data if key == "data" else self.header.value(key)

    @property
    def shape(self):
        return self.header.shape

    @property
    def dimension(self):
        return self.header.ndim

    @property
    def esize(self):
        return self.header.esize

    @property
    def dtype(self):
        return self.header.dtype

    def size(self, skip=0):
        """Elements in the fastest axes after dropping the `skip` slowest."""
        return int(np.prod(self.shape[skip:], dtype=np.int64))

    def _all(self):
        h = self.header
        if h.form == "native":
            a = np.fromfile(h.infile, dtype=h.dtype, count=self.size())
            if a.size != self.size():
                raise RsfError("%s: data file holds %d of %d elements" % (self.path, a.size, self.size()))
            return a.reshape(self.shape)
        a = np.loadtxt(h.infile, dtype=h.dtype, ndmin=1)
        return a.reshape(self.shape) if a.size == self.size() else a

    @property
    def data(self):
        return self._all()

    def read(self, count=None):
        """The whole array (count=None), or the next `count` elements of a
        native file in storage order."""
        if count is None:
            return self._all()
        h = self.header
        if h.form != "native":
            raise RsfError("partial reads need a native file")
        a = np.fromfile(h.infile, dtype=h.dtype, count=count, offset=self._pos * h.dtype.itemsize)
        self._pos += a.size
        return a

    def traces(self):
        """Consecutive traces (n1 samples each), first to last."""
        if self.header.form != "native":
            yield from np.atleast_2d(self._all())
            return
        n1 = max(1, self.header.axes[0]["n"])
        self._pos = 0
        while True:
            t = self.read(n1)
            if t.size == 0:
                return
            yield t


class Writer:
    """An RSF file open for writing (``output``): the header is written
    with the data."""

    def __init__(self, path, data_format="native_float", **keys):
        self.path = path
        infile = keys.pop("in", None) or path + "@"
        keys.pop("esize", None)   # derived from data_format
        self.header = Header(os.path.abspath(infile), data_format)
        for key, val in keys.items():
            m = _AXIS_KEY.match(key)
            if not m:
                raise RsfError("unknown header key %r" % (key,))
            self.header.axes[int(m.group(2)) - 1][m.group(1)] = _axis_value(m.group(1), val)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        return False

    def close(self):
        pass

    def write(self, data, fmt=None):
        h = self.header
        a = np.asarray(data)
        if h.ndim == 0:   # no axes given: one axis of all the samples
            h.axes[0]["n"] = int(a.size)
        a = a.astype(h.dtype, copy=False)
        if h.form == "native":
            np.ascontiguousarray(a).tofile(h.infile)
        else:
            if h.type == "complex":
                raise RsfError("ascii complex data are not supported")
            n1 = max(1, h.axes[0]["n"])
            table = a.reshape(-1, n1) if a.ndim > 1 else a.reshape(-1)
            np.savetxt(h.infile, table, fmt or ("%d" if h.type == "int" else "%.18e"))
        with open(self.path, "w") as f:
            f.write(h.text())


def input(filename):  # noqa: A001 -- the reference module's name for it
    """Open an RSF file for reading."""
    return Reader(filename)


def output(filename, data_format="native_float", **keys):
    """An RSF writer for `filename` with header keys (n1=, d1=, label1= ...)."""
    return Writer(filename, data_format, **keys)


def fromfile(filename, keys):
    """Header values and/or the array ("data") of an RSF file.  `keys` is a
    list or a blank-separated string; one key gives its value, several a
    list in the same order."""
    if isinstance(keys, str):
        keys = keys.split()
    r = Reader(filename)
    vals = [r[k] for k in keys]
    return vals[0] if len(vals) == 1 else vals


def tofile(filename, data, **keys):
    """Write `data` with header keys to `filename` (data in `filename@`)."""
    Writer(filename, **keys).write(data)
