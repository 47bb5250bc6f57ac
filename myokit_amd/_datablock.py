"""
Time x space views of a tissue log: ``DataBlock1d`` and ``DataBlock2d``
(``myokit/_datablock.py:87-786, 787-1631``), for the analyses that follow a
tissue run (traces, conduction velocity, frames).

A block holds a time vector of ``n`` entries, 0d series of length ``n`` and
1d series as ``(n, w)`` arrays (2d: ``(n, h, w)``, y first, as the
reference). Blocks are made from a DataLog with per-cell keys
(``from_log``) or, without per-cell keys, from the dense output of
``SimulationOpenCL.run_block`` (``from_run_block``).
"""
from collections import OrderedDict

import numpy as np

from ._datalog import DataLog, split_key


_README = {1: """DataBlock1d file written by DataBlock1d.save

header_block1d.txt: nt (points in time), nx (width), "dtype" ("d" or "f"),
the quoted names of the 0d series, a line "1", the quoted names of the 1d
series.
data.bin: little-endian values of the given type: the nt times, every 0d
series, every 1d series as an (nt, nx) array in C order.
""", 2: """DataBlock2d file written by DataBlock2d.save

header_block2d.txt: nt (points in time), ny (height), nx (width), "dtype"
("d" or "f"), the quoted names of the 0d series, a line "2", the quoted
names of the 2d series.
data.bin: little-endian values of the given type: the nt times, every 0d
series, every 2d series as an (nt, ny, nx) array in C order.
"""}


def _block_save(filename, d, dims, time, zero, multi):
    """Writes the zip format of ``DataBlock1d/2d.save``
    (``myokit/_datablock.py:631-704, 1423-1500``), in double precision."""
    import os
    import zipfile
    head = [str(len(time))] + [str(x) for x in dims] + ['"d"']
    head += ['"' + k + '"' for k in zero] + [str(d)]
    head += ['"' + k + '"' for k in multi]
    dt = np.dtype('<f8')
    body = [np.ascontiguousarray(time, dtype=dt).tobytes()]
    body += [np.ascontiguousarray(v, dtype=dt).tobytes()
             for v in zero.values()]
    body += [np.ascontiguousarray(v, dtype=dt).tobytes()
             for v in multi.values()]
    with zipfile.ZipFile(os.path.expanduser(filename), 'w',
                         zipfile.ZIP_DEFLATED) as f:
        f.writestr('header_block' + str(d) + 'd.txt',
                   '\n'.join(head).encode('utf8'))
        f.writestr('data.bin', b''.join(body))
        f.writestr('readme.txt', _README[d].encode('utf8'))


def _block_load(filename, d, accept_1d=False):
    """Reads the zip format of ``DataBlock1d/2d.load``
    (``myokit/_datablock.py:460-621, 1233-1413``). Returns ``(d, dims, time,
    zero, multi)`` with ``dims`` = (nx,) or (ny, nx)."""
    import os
    import zipfile
    from ._err import DataBlockReadError
    kind = 'DataBlock' + str(d) + 'd'
    try:
        with zipfile.ZipFile(os.path.expanduser(filename), 'r') as f:
            names = f.namelist()
            if len(names) < 3:
                raise DataBlockReadError(
                    'Invalid ' + kind + ' file format: Not enough files in'
                    ' zip.')
            hname = 'header_block' + str(d) + 'd.txt'
            if hname not in names and accept_1d and \
                    'header_block1d.txt' in names:
                d, hname = 1, 'header_block1d.txt'
            if hname not in names:
                raise DataBlockReadError(
                    'Invalid ' + kind + ' file format: Header not found.')
            if 'data.bin' not in names:
                raise DataBlockReadError(
                    'Invalid ' + kind + ' file format: Data not found.')
            head = f.read(hname).decode('utf8').splitlines()
            body = f.read('data.bin')
    except zipfile.BadZipfile:
        raise DataBlockReadError('Unable to read ' + kind + ': Bad zip file.')
    it = iter(head)
    nt = int(next(it))
    dims = tuple(int(next(it)) for _ in range(d))
    code = next(it)[1:-1]
    if code not in ('d', 'f'):
        raise DataBlockReadError('Unable to read ' + kind
                                 + ': Unrecognized data type "' + code + '".')
    dt = np.dtype('<f8' if code == 'd' else '<f4')
    names0 = []
    name = next(it)
    while name != str(d):
        names0.append(name[1:-1])
        name = next(it)
    names1 = [x[1:-1] for x in it]
    n0 = nt
    n1 = nt * int(np.prod(dims))
    pos = [0]

    def take(count, what):
        end = pos[0] + count * dt.itemsize
        if end > len(body):
            raise DataBlockReadError(
                'Unable to read ' + kind + ': Header indicates more ' + what
                + ' data than found in the body.')
        a = np.frombuffer(body, dtype=dt, count=count, offset=pos[0])
        pos[0] = end
        return a.astype(dt.newbyteorder('='))
    time = take(n0, 'time')
    zero = OrderedDict((k, take(n0, '0d')) for k in names0)
    multi = OrderedDict((k, take(n1, str(d) + 'd').reshape((nt,) + dims))
                        for k in names1)
    return d, dims, time, zero, multi


def _cells_of(log):
    """Splits the log's keys into 0d names and per-cell names, the latter
    mapped to ``{index tuple: key}``."""
    zero, multi = [], OrderedDict()
    for key in log.keys():
        idx, name = split_key(key)
        if not idx:
            zero.append(name)
        else:
            t = tuple(int(x) for x in idx.split('.')[:-1])
            multi.setdefault(name, {})[t] = key
    return zero, multi


class DataBlock1d:
    """Time series of 1d arrays of width ``w`` (one value per cell)."""

    def __init__(self, w, time, copy=True):
        w = int(w)
        if w < 1:
            raise ValueError('Minimum w is 1.')
        self._nx = w
        time = np.array(time) if copy else np.asarray(time)
        if len(time.shape) != 1:
            raise ValueError('Time must be a sequence.')
        if np.any(np.diff(time) < 0):
            raise ValueError('Time must be non-decreasing.')
        self._time = time
        self._nt = len(time)
        self._0d = {}
        self._1d = {}

    @staticmethod
    def from_log(log):
        """Block from a DataLog with a time key, 0d and 1d columns
        (``DataBlock1d.from_log``, ``_datablock.py:266-330``)."""
        log.validate()
        time = log.time_key()
        if time is None:
            raise ValueError('No time variable set in data log.')
        zero, multi = _cells_of(log)
        if time not in zero:
            raise ValueError('Time variable must be 0-dimensional.')
        nx = None
        for name, cells in multi.items():
            dims = {len(t) for t in cells}
            if dims != {1}:
                raise ValueError(
                    'The given simulation log should only contain 0d or 1d'
                    ' variables. Found <' + name + '> with d = '
                    + str(max(dims)) + '.')
            size = max(t[0] for t in cells) + 1
            if nx is None:
                nx = size
            elif size != nx:
                raise ValueError('The given simulation log contains 1d'
                                 ' data sets of different sizes.')
        block = DataBlock1d(nx or 1, log[time], copy=True)
        nt = block._nt
        for name in zero:
            if name != time:
                block.set0d(name, log[name])
        for name, cells in multi.items():
            data = np.zeros((nt, nx))
            for (x,), key in cells.items():
                data[:, x] = log[key]
            block.set1d(name, data, copy=False)
        return block

    @staticmethod
    def from_run_block(out):
        """Block from ``SimulationOpenCL.run_block`` output of a 1D run."""
        t = np.asarray(out['time'])
        block = None
        for name, v in out.items():
            if name in ('time', 'pace'):
                continue
            v = np.asarray(v)
            if v.ndim != 2:
                raise ValueError('Expected (rows, nx) arrays.')
            if block is None:
                block = DataBlock1d(v.shape[1], t)
            block.set1d(name, v)
        if block is None:
            block = DataBlock1d(1, t)
        block.set0d('pace', np.asarray(out['pace']))
        return block

    def block2d(self):
        """The same data as a DataBlock2d of height 1."""
        b = DataBlock2d(self._nx, 1, self._time)
        for k, v in self._0d.items():
            b.set0d(k, v)
        for k, v in self._1d.items():
            b.set2d(k, v.reshape((self._nt, 1, self._nx)))
        return b

    def cv(self, name, threshold=-30, length=0.01, time_multiplier=1e-3,
           border=None):
        """Conduction velocity (cm/s) along the cable, from the first
        activation of each cell away from the ends (``DataBlock1d.cv``,
        ``_datablock.py:152-254``). As there, a cell's activation is the
        first log index after the first one where V is above -30 mV and
        rising, linearly interpolated to ``threshold``; the cells used are
        the run of consecutive activated cells from the first one, and the
        velocity is the least-squares slope of position against time."""
        if border is None:
            border = min(50, self._nx // 3)
        else:
            border = int(border)
            if border < 0:
                raise ValueError('The argument `border` cannot be negative.')
            elif border >= self._nx // 2:
                raise ValueError(
                    'The argument `border` must be less than half the number'
                    ' of cells.')
        V = self._1d[name]
        i1 = i2 = None
        have = False
        t = []
        for i in range(border, self._nx - border):
            v = V[:, i]
            idx = np.nonzero((v[1:] > -30) & (v[1:] - v[:-1] > 0))[0]
            if len(idx) == 0 or idx[0] == 0:
                if have:
                    i2 = i - 1
                    break
            else:
                if have:
                    i2 = i
                else:
                    i1 = i
                    have = True
                k = 1 + idx[0]
                v0, v1 = v[k - 1], v[k]
                t0, t1 = self._time[k - 1], self._time[k]
                t.append(t0 + (threshold - v0) * (t1 - t0) / (v1 - v0))
        if not have:
            return 0
        if np.all(np.asarray(t) == t[0]):
            return 0
        t = np.asarray(t) * time_multiplier
        x = np.arange(i1, 1 + i2, dtype=float) * length
        A = np.vstack([t, np.ones(len(t))]).T
        return np.linalg.lstsq(A, x, rcond=None)[0][0]

    @staticmethod
    def load(filename, progress=None, msg='Loading DataBlock1d'):
        """Reads a file written by :meth:`save` or by myokit's
        ``DataBlock1d.save``."""
        _, (nx,), time, zero, multi = _block_load(filename, 1)
        block = DataBlock1d(nx, time, copy=False)
        for k, v in zero.items():
            block.set0d(k, v, copy=False)
        for k, v in multi.items():
            block.set1d(k, v, copy=False)
        return block

    def save(self, filename):
        """Writes the block in myokit's DataBlock1d zip format."""
        _block_save(filename, 1, (self._nx,), self._time, self._0d,
                    self._1d)

    def get0d(self, name):
        return self._0d[name]

    def get1d(self, name):
        return self._1d[name]

    def keys0d(self):
        return self._0d.keys()

    def keys1d(self):
        return self._1d.keys()

    def len0d(self):
        return len(self._0d)

    def len1d(self):
        return len(self._1d)

    def remove0d(self, name):
        del self._0d[name]

    def remove1d(self, name):
        del self._1d[name]

    def set0d(self, name, data, copy=True):
        data = np.array(data, copy=True) if copy else np.asarray(data)
        if data.shape != (self._nt,):
            raise ValueError('Data must have shape (n,) = (' + str(self._nt)
                             + ',).')
        self._0d[str(name)] = data

    def set1d(self, name, data, copy=True):
        data = np.array(data, copy=True) if copy else np.asarray(data)
        if data.shape != (self._nt, self._nx):
            raise ValueError('Data must have shape (n, w) = ('
                             + str(self._nt) + ', ' + str(self._nx) + ').')
        self._1d[str(name)] = data

    def shape(self):
        """``(n, w)``."""
        return (self._nt, self._nx)

    def time(self):
        return self._time

    def to_log(self, copy=True):
        """A DataLog with keys ``x.name`` for the 1d series."""
        c = np.copy if copy else (lambda a: a)
        d = DataLog()
        d['time'] = c(self._time)
        d.set_time_key('time')
        for k, v in self._0d.items():
            d[k] = c(v)
        for k, v in self._1d.items():
            for x in range(self._nx):
                d[str(x) + '.' + k] = c(v[:, x])
        return d

    def trace(self, variable, x):
        """The time series of cell ``x``."""
        return self._1d[variable][:, x]


class DataBlock2d:
    """Time series of 2d arrays of shape ``(h, w)`` (y first)."""

    def __init__(self, w, h, time, copy=True):
        w, h = int(w), int(h)
        if w < 1 or h < 1:
            raise ValueError('Minimum w and h is 1.')
        self._nx, self._ny = w, h
        time = np.array(time) if copy else np.asarray(time)
        if len(time.shape) != 1:
            raise ValueError('Time must be a sequence.')
        if np.any(np.diff(time) < 0):
            raise ValueError('Time must be non-decreasing.')
        self._time = time
        self._nt = len(time)
        self._0d = {}
        self._2d = {}

    @staticmethod
    def from_log(log):
        """Block from a DataLog with a time key and 2d per-cell columns
        ``x.y.name`` (``DataBlock2d.from_log``, ``_datablock.py:1069-1128``)."""
        log.validate()
        time = log.time_key()
        if time is None:
            raise ValueError('No time variable set in data log.')
        zero, multi = _cells_of(log)
        if time not in zero:
            raise ValueError('Time variable must be 0-dimensional.')
        size = None
        for name, cells in multi.items():
            dims = {len(t) for t in cells}
            if dims != {2}:
                raise ValueError(
                    'The given simulation log should only contain 0d or 2d'
                    ' variables. Found <' + name + '> with d = '
                    + str(max(dims)) + '.')
            s = (max(t[0] for t in cells) + 1, max(t[1] for t in cells) + 1)
            if size is None:
                size = s
            elif s != size:
                raise ValueError('The given simulation log contains 2d'
                                 ' data sets of different sizes.')
        nx, ny = size or (1, 1)
        block = DataBlock2d(nx, ny, log[time], copy=True)
        for name in zero:
            if name != time:
                block.set0d(name, log[name])
        for name, cells in multi.items():
            data = np.zeros((block._nt, ny, nx))
            for (x, y), key in cells.items():
                data[:, y, x] = log[key]
            block.set2d(name, data, copy=False)
        return block

    @staticmethod
    def from_run_block(out):
        """Block from ``SimulationOpenCL.run_block`` output of a 2D run
        (arrays of shape ``(rows, ny, nx)``), without copying them."""
        t = np.asarray(out['time'])
        block = None
        for name, v in out.items():
            if name in ('time', 'pace'):
                continue
            v = np.asarray(v)
            if v.ndim != 3:
                raise ValueError('Expected (rows, ny, nx) arrays.')
            if block is None:
                block = DataBlock2d(v.shape[2], v.shape[1], t)
            block.set2d(name, v, copy=False)
        if block is None:
            block = DataBlock2d(1, 1, t)
        block.set0d('pace', np.asarray(out['pace']))
        return block

    @staticmethod
    def load(filename, progress=None, msg='Loading DataBlock2d'):
        """Reads a file written by :meth:`save` or by myokit's
        ``DataBlock2d.save``; a DataBlock1d file becomes a block of height
        1."""
        d, dims, time, zero, multi = _block_load(filename, 2, accept_1d=True)
        ny, nx = dims if d == 2 else (1, dims[0])
        block = DataBlock2d(nx, ny, time, copy=False)
        for k, v in zero.items():
            block.set0d(k, v, copy=False)
        for k, v in multi.items():
            block.set2d(k, v.reshape((len(time), ny, nx)), copy=False)
        return block

    def save(self, filename):
        """Writes the block in myokit's DataBlock2d zip format."""
        _block_save(filename, 2, (self._ny, self._nx), self._time, self._0d,
                    self._2d)

    def get0d(self, name):
        return self._0d[name]

    def get2d(self, name):
        return self._2d[name]

    def items0d(self):
        return self._0d.items()

    def items2d(self):
        return self._2d.items()

    def is_square(self):
        return self._nx == self._ny

    def keys0d(self):
        return self._0d.keys()

    def keys2d(self):
        return self._2d.keys()

    def len0d(self):
        return len(self._0d)

    def len2d(self):
        return len(self._2d)

    def remove0d(self, name):
        del self._0d[name]

    def remove2d(self, name):
        del self._2d[name]

    def set0d(self, name, data, copy=True):
        data = np.array(data, copy=True) if copy else np.asarray(data)
        if data.shape != (self._nt,):
            raise ValueError('Data must have shape (n,) = (' + str(self._nt)
                             + ',).')
        self._0d[str(name)] = data

    def set2d(self, name, data, copy=True):
        data = np.array(data, copy=True) if copy else np.asarray(data)
        if data.shape != (self._nt, self._ny, self._nx):
            raise ValueError('Data must have shape (n, h, w) = ('
                             + str(self._nt) + ', ' + str(self._ny) + ', '
                             + str(self._nx) + ').')
        self._2d[str(name)] = data

    def shape(self):
        """``(n, h, w)``."""
        return (self._nt, self._ny, self._nx)

    def time(self):
        return self._time

    def to_log(self, copy=True):
        """A DataLog with keys ``x.y.name`` for the 2d series."""
        c = np.copy if copy else (lambda a: a)
        d = DataLog()
        d['time'] = c(self._time)
        d.set_time_key('time')
        for k, v in self._0d.items():
            d[k] = c(v)
        for k, v in self._2d.items():
            for y in range(self._ny):
                for x in range(self._nx):
                    d[str(x) + '.' + str(y) + '.' + k] = c(v[:, y, x])
        return d

    def trace(self, variable, x, y):
        """The time series of cell ``(x, y)``."""
        return self._2d[variable][:, y, x]
