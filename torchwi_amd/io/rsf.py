"""RSF (Madagascar) files: a text header and a binary or ASCII data file.

The interface follows the reference's reader/writer (torchwi/io/rsf.py): the
module functions ``input``, ``output``, ``fromfile``, ``tofile`` and the
class ``RSF`` with ``get``/``gets``/``read``/``traces``/``write``, the same
header keys, the same header text on write, and the same error behaviour (a
message on stderr and exit status 1, i.e. ``SystemExit``).  Differences: it
runs on NumPy 2 (the reference's ``np.product``, rsf.py:224, is gone there),
and ``__setitem__`` works (rsf.py:163-165 refers to an undefined name).

Header grammar (rsf.py:313-324, 369-382): ``key=value`` pairs anywhere in
the text, the LAST occurrence of a key wins, values may be quoted with ' or
".  Integer keys: esize, n1..n9; float keys: o1..o9, d1..d9.  The data file
holds n1 values fastest; ``read(read_all=True)`` returns shape
(n_dim, ..., n2, n1).  ``data_format`` is ``<form>_<type>`` with form
native|ascii and type float|int|complex (float32, int32, complex64).
"""
import os
import re
import sys
import tempfile

import numpy as np

MAXDIM = 9
RSFKEYS_READ = (['in', 'esize', 'data_format', 'type', 'form', '_in', 'infile'] +
                [k + str(i) for i in range(1, MAXDIM + 1) for k in ('n', 'o', 'd', 'label', 'unit')])
RSFKEYS_WRITE = RSFKEYS_READ[:3]
KEY_INT = ['esize'] + ['n%d' % i for i in range(1, MAXDIM + 1)]
KEY_FLOAT = [k + str(i) for i in range(1, MAXDIM + 1) for k in ('o', 'd')]
_AXKEYS = ['ax%d' % i for i in range(1, MAXDIM + 1)] + ['axes']
_DTYPES = {'float': np.float32, 'int': np.int32, 'complex': np.complex64}


def warn(msg):
    sys.stderr.write(msg + "\n")


def errexit(msg, errno=1):
    """Report and exit with status errno, as the reference does (rsf.py:400)."""
    warn(msg)
    sys.exit(errno)


def quote(s):
    return '"' + s + '"'


def unquote(s):
    if s.startswith("'") or s.startswith('"'):
        return s[1:-1]
    return s


def _check_key(key):
    if key not in RSFKEYS_READ + _AXKEYS:
        print("Wrong keyword:", key)
        sys.exit(1)


def _set_type(key, val, quote_str=False):
    if quote_str:
        return val if (key in KEY_INT or key in KEY_FLOAT) else quote(val)
    if key in KEY_INT:
        return int(val)
    if key in KEY_FLOAT:
        return float(val)
    return unquote(val)


def fromstring(key, text, un_quote=True):
    """Value of the last ``key=value`` in text: (value, True), or (0, False)."""
    vals = re.findall(re.escape(key) + r"\s*=\s*(\S+)", text)
    if not vals:
        return 0, False
    val = vals[-1]
    for q in ("'", '"'):
        if val.startswith(q) and not val.endswith(q):
            full = re.findall(re.escape(key) + "=" + q + "(.+?)" + q, text)
            if full:
                val = full[-1]
    if un_quote:
        val = unquote(val)
    return val, True


def file2str(filename=''):
    if filename and filename.lower() != 'stdin':
        with open(filename, 'r') as fh:
            return fh.read()
    return sys.stdin.read()


def str2file(string, filename=''):
    if filename and filename.lower() != 'stdout':
        with open(filename, 'w') as fh:
            fh.write(string)
    else:
        sys.stdout.write(string)


class Axis:
    def __init__(self, n=0, o=0., d=0., label='', unit=''):
        self.n = int(n)
        self.o = float(o)
        self.d = float(d)
        self.label = unquote(label)
        self.unit = unquote(unit)

    def to_str(self, idim):
        s = 'n%d=%d o%d=%s d%d=%s' % (idim, self.n, idim, self.o, idim, self.d)
        if self.label:
            s += ' label%d="%s"' % (idim, self.label)
        if self.unit:
            s += ' unit%d="%s"' % (idim, self.unit)
        return s

    def __repr__(self):
        return "Axis(n=%d, o=%s, d=%s, label=%r, unit=%r)" % (self.n, self.o, self.d, self.label, self.unit)


class RSF:
    def __init__(self, mode=None, readonly=True, data_format=None):
        self._in = ""
        self._open = False
        self.f = None
        self.filename = ''
        self.wrote = False
        self._data_format = data_format or 'native_float'
        self._mode = mode or self._guess_mode(readonly)
        self.axes = [Axis() for _ in range(MAXDIM)]

    def _guess_mode(self, readonly):
        if readonly:
            return 'r' if self.form == 'ascii' else 'rb'
        return 'wb'

    def open_in(self, tmpfile=False, abspath=True):
        if tmpfile:
            self.f = tempfile.NamedTemporaryFile(mode=self._mode, suffix=".rsf@", dir="./", delete=False)
            self.set_in(self.f.name, abspath)
        else:
            self.f = open(self._in, self._mode)
        self._open = True

    def __enter__(self):
        return self

    def __exit__(self, exception_type, exception_value, traceback):
        self.close()

    # ---- header access -----------------------------------------------------
    def get(self, key):
        if key == 'in':
            return self._in
        if key in ('esize', 'data_format', 'form', 'type'):
            return getattr(self, key)
        if key == 'axes':
            return self.axes
        if key.startswith('ax'):
            return self.axes[int(key[-1]) - 1]
        if key == 'data':
            return self.read(read_all=True)
        if key == 'header':
            return self.header_dict()
        name, idim = key[:-1], int(key[-1])
        return getattr(self.axes[idim - 1], name)

    def __getitem__(self, key):
        return self.get(key)

    def __setitem__(self, key, value):
        self.set(**{key: value})

    def header_dict(self, contain_in=False):
        dic = {}
        if contain_in:
            dic['in'] = self._in
        dic.update(self.axes_dict())
        dic['data_format'] = self.data_format
        dic['esize'] = self.esize
        return dic

    def gets(self, keys):
        if len(keys) == 1:
            return self.get(keys[0])
        return [self.get(k) for k in keys]

    def axes_dict(self):
        return {"ax%d" % i: self.get("ax%d" % i) for i in range(1, self.dimension + 1)}

    def set_in(self, _in, abspath=True):
        self._in = os.path.abspath(_in) if abspath else _in

    def set(self, **kwargs):
        for key, val in kwargs.items():
            if key in ('in', 'infile', '_in'):
                self._in = val
            elif key == 'esize':
                self._esize = val
            elif key in ('data_format', 'form', 'type'):
                setattr(self, key, val)
            elif key == 'axes':
                for idim, ax in enumerate(val):
                    self.axes[idim] = ax
            elif key.startswith('ax'):
                self.axes[int(key[-1]) - 1] = val
            else:
                name, idim = key[:-1], int(key[-1])
                setattr(self.axes[idim - 1], name, val)

    def size(self, istart=0):
        """Product of shape[istart:], shape = (n_dim, ..., n2, n1): all values
        for 0, the values of one slice of the slowest axis for 1 (the
        reference's code, rsf.py:216-224; its docstring counts the other way)."""
        return int(np.prod(self.shape[istart:], dtype=np.int64))

    @property
    def dimension(self):
        dim = 0
        for idim in range(1, MAXDIM + 1):
            if self.get("n%d" % idim) > 0:
                dim = idim
        return dim

    @property
    def shape(self):
        return tuple(self.get("n%d" % idim) for idim in range(self.dimension, 0, -1))

    @property
    def dtype(self):
        if self.type not in _DTYPES:
            print("Wrong type:", self.type)
            sys.exit(1)
        return _DTYPES[self.type]

    @property
    def data_format(self):
        return self._data_format

    @data_format.setter
    def data_format(self, val):
        self._data_format = val
        if self.type.lower() not in ('int', 'float', 'complex'):
            print('Wrong type:', self.type)
            sys.exit(1)
        if self.form.lower() not in ('native', 'ascii'):
            print('Wrong format:', self.form)
            sys.exit(1)

    @property
    def type(self):
        return self._data_format.split('_')[1].lower()

    @type.setter
    def type(self, val):
        self._data_format = "%s_%s" % (self._data_format.split('_')[0], val)

    @property
    def form(self):
        return self._data_format.split('_')[0].lower()

    @form.setter
    def form(self, val):
        self._data_format = "%s_%s" % (val, self._data_format.split('_')[1])

    @property
    def esize(self):
        if self.form == 'ascii':
            return 0
        return 8 if self.type == 'complex' else 4

    # ---- data --------------------------------------------------------------
    def close(self):
        if self._open:
            self.f.close()
            self._open = False

    def read(self, read_all=False):
        """All values (read_all, shaped and closing the file) or the next
        trace of n1 values; native or ASCII."""
        if self.form == "native":
            count = self.size() if read_all else self.get("n1")
            data = np.fromfile(self.f, dtype=self.dtype, count=count)
            if read_all:
                data.shape = self.shape
                self.close()
        else:
            # the writer prints every type with a float format ('%.18e'), so
            # integers are read as doubles and converted
            data = np.loadtxt(self.f, dtype=np.float64 if self.type == 'int' else self.dtype).astype(self.dtype)
            if data.shape == ():
                data.shape = (1,)
        return data

    def traces(self):
        if self.form == 'native':
            trc = self.read()
            while len(trc) > 0:
                yield trc
                trc = self.read()
        else:
            yield self.read()

    def write(self, data, close=False, fmt='%.18e'):
        self.write_rsf()
        if self.form == "native":
            np.asarray(data).astype(self.dtype).tofile(self.f)
        elif fmt:
            np.savetxt(self.f, data, fmt)
        else:
            np.savetxt(self.f, data)
        if close:
            self.close()

    def read_rsf(self, filename):
        header = file2str(filename)
        for key in RSFKEYS_READ:
            val, found = fromstring(key, header)
            if found:
                self.set(**{key: _set_type(key, val)})
        infile = self.get('in')
        if not infile:
            errexit("Error: Cannot find 'in' keyword from rsf header")
        if not os.path.exists(infile):
            errexit("Error: File not exists - %s" % infile)

    def write_rsf(self):
        if self.wrote:
            return
        s = ""
        for key in RSFKEYS_WRITE:
            s += "%s=%s\n" % (key, _set_type(key, self.get(key), quote_str=True))
        for idim in range(self.dimension):
            s += self.axes[idim].to_str(idim + 1) + "\n"
        str2file(s, self.filename)
        self.wrote = True


def input(filename='', mode=None, data_format=None):  # noqa: A001 -- the reference's name
    r = RSF(mode=mode, readonly=True, data_format=data_format)
    r.filename = filename
    r.read_rsf(filename)
    r.open_in()
    return r


def output(filename='', mode=None, data_format=None, _in=None, open_in=True, abspath=True, **kwargs):
    r = RSF(mode=mode, readonly=False, data_format=data_format)
    r.filename = filename
    for key in kwargs:
        _check_key(key)
    r.set(**kwargs)
    r.wrote = False
    if filename and filename.lower() != 'stdout':
        r.set_in(_in if _in else filename + '@', abspath)
        if open_in:
            r.open_in()
    elif _in:
        r.set_in(_in, abspath)
        if open_in:
            r.open_in()
    elif open_in:
        r.open_in(tmpfile=True, abspath=abspath)
    return r


def fromfile(filename, key):
    """Header values and/or data of an RSF file: ``fromfile(f, "n1 d1 data")``."""
    with input(filename) as f:
        if isinstance(key, str):
            key = key.split()
        return f.gets(key)


def tofile(filename, data, **kwargs):
    """Write data with the header keys given (n1=..., d1=..., ...)."""
    with output(filename, **kwargs) as f:
        f.write(data)
