"""
Simulation logs.

``DataLog`` is a dict of ``key -> column`` like ``myokit.DataLog``
(``myokit/_datalog.py:42``) with the same key grammar: ``"engine.time"`` for
global variables, ``"3.membrane.V"`` (1D) or ``"3.1.membrane.V"`` (2D, x
first) for per-cell ones. Columns produced by the engine are numpy arrays
(float32 in single precision, float64 in double, as the reference's
``array('f')`` / ``array('d')``), which avoids the reference's per-value Python
appends at large cell counts.

``prepare_log`` follows ``myokit.prepare_log`` (``_datalog.py:1783-2193``):
flags, lists of names, and existing logs, with the same checks.
"""
import os as _os
import re
from collections import OrderedDict

import numpy as np

from . import _expressions as ex

LOG_NONE = 0
LOG_STATE = 1
LOG_BOUND = 2
LOG_INTER = 4
LOG_DERIV = 8
LOG_ALL = LOG_STATE + LOG_INTER + LOG_BOUND + LOG_DERIV

SINGLE_PRECISION = 32
DOUBLE_PRECISION = 64

_ID_NAME = re.compile(r'(\d+\.)+')

_README = """Log file written by DataLog.save

structure.txt: number of fields, length of each field, type code ('d' for
64-bit or 'f' for 32-bit floats), the name of the time field (or an empty
line), then one field name per line.
data.bin: the fields in the order of structure.txt, each as a little-endian
array of the given type and length.
"""


def split_key(key):
    """Splits ``"1.2.membrane.V"`` into ``("1.2.", "membrane.V")``."""
    m = _ID_NAME.match(key, 0)
    if m:
        return key[:m.end()], key[m.end():]
    return '', key


def dimco(*dims):
    """Yields all index tuples of a grid, first dimension changing fastest:
    ``dimco(2, 3)`` gives (0,0), (1,0), (0,1), (1,1), (0,2), (1,2)."""
    if len(dims) == 0:
        return
    grids = np.indices(tuple(reversed(dims))).reshape(len(dims), -1)
    for idx in grids.T:
        yield tuple(int(i) for i in reversed(idx))


def float_str(number, precision=DOUBLE_PRECISION):
    """A number as text the way ``myokit.float.str`` writes it
    (``myokit/float.py:64-103``): Python's repr when shorter than 10
    characters, else full precision with a sign column
    (``'{:< 1.17e}'``, single precision ``'{:< 1.9e}'``)."""
    if isinstance(number, str):
        return number
    s = str(number)
    if len(s) < 10:
        return s
    if precision == SINGLE_PRECISION:
        return '{:< 1.9e}'.format(float(number))
    return '{:< 1.17e}'.format(float(number))


class DataLog(OrderedDict):
    """An ordered dict of logged columns, with an optional time key."""

    def __init__(self, other=None, time=None):
        super().__init__()
        self._time = None
        self.cmeta = _MetaDict()
        if other is not None:
            for k, v in other.items():
                self[k] = v
            if isinstance(other, DataLog):
                self._time = other._time
        if time is not None:
            self.set_time_key(time)

    def time_key(self):
        return self._time

    def set_time_key(self, key):
        if key is not None and key not in self:
            raise ValueError('Time key not found in log: ' + str(key))
        self._time = key

    def time(self):
        if self._time is None:
            raise ValueError('No time variable set in log.')
        return self[self._time]

    def length(self):
        for v in self.values():
            return len(v)
        return 0

    def __len__(self):
        return OrderedDict.__len__(self)

    def has_nan(self):
        """True if any column has a NaN as its final logged value
        (``DataLog.has_nan``, ``myokit/_datalog.py:366-374``)."""
        for v in self.values():
            if len(v) > 0 and np.isnan(v[-1]):
                return True
        return False

    def npview(self):
        """Returns a DataLog with every column as a numpy array."""
        d = DataLog()
        for k, v in self.items():
            d[k] = np.asarray(v)
        d._time = self._time
        return d

    def validate(self):
        n = None
        for k, v in self.items():
            if n is None:
                n = len(v)
            elif len(v) != n:
                raise ValueError(
                    'All columns in a DataLog must have the same length.')

    def variable_info(self):
        """Maps each variable name to the sorted list of cell indices it is
        logged for (``()`` for global variables)."""
        info = OrderedDict()
        for k in self.keys():
            idx, name = split_key(k)
            t = tuple(int(x) for x in idx.split('.')[:-1]) if idx else ()
            info.setdefault(name, []).append(t)
        return info

    def extend_column(self, key, values):
        """Appends ``values`` to a column, whatever the column's type."""
        cur = self[key]
        if isinstance(cur, np.ndarray):
            OrderedDict.__setitem__(
                self, key, np.concatenate([cur, np.asarray(
                    values, dtype=cur.dtype)]))
        elif hasattr(cur, 'extend'):
            cur.extend(values)
        else:
            for x in values:
                cur.append(x)

    def save(self, filename, precision=DOUBLE_PRECISION):
        """Writes the log to a zip file in myokit's binary DataLog format
        (``DataLog.save``, ``myokit/_datalog.py:1051-1126``):
        ``structure.txt`` holds the number of columns, their length, the
        type code (``d`` or ``f``), the time key and the column names, one
        per line; ``data.bin`` the columns one after another as
        little-endian arrays of that type."""
        import zipfile
        self.validate()
        code = 'd' if precision == DOUBLE_PRECISION else 'f'
        dt = np.dtype('<f8' if code == 'd' else '<f4')
        head = [str(len(self)), str(self.length()), code, self._time or '']
        body = []
        for k, v in self.items():
            head.append(k)
            body.append(np.ascontiguousarray(v, dtype=dt).tobytes())
        with zipfile.ZipFile(_os.path.expanduser(filename), 'w',
                             zipfile.ZIP_DEFLATED) as f:
            f.writestr('data.bin', b''.join(body))
            f.writestr('structure.txt', '\n'.join(head).encode('utf8'))
            f.writestr('readme.txt', _README.encode('utf8'))

    @staticmethod
    def load(filename, progress=None, msg='Loading DataLog'):
        """Reads a log written by :meth:`save` (or by myokit's
        ``DataLog.save``, ``myokit/_datalog.py:525-668``). Columns are numpy
        arrays of the stored type."""
        import zipfile
        from ._err import DataLogReadError
        try:
            with zipfile.ZipFile(_os.path.expanduser(filename), 'r') as f:
                try:
                    head = f.read('structure.txt').decode('utf8')
                    body = f.read('data.bin')
                except KeyError:
                    raise DataLogReadError('Invalid log file format.')
        except zipfile.BadZipfile:
            raise DataLogReadError('Unable to read log: bad zip file.')
        lines = head.splitlines()
        if len(lines) < 4:
            raise DataLogReadError('Invalid log file format.')
        n, size, code, time = int(lines[0]), int(lines[1]), lines[2], lines[3]
        fields = lines[4:]
        if len(fields) != n:
            raise DataLogReadError('Invalid number of fields specified.')
        if size < 0:
            raise DataLogReadError('Invalid data size: ' + str(size) + '.')
        if code not in ('d', 'f'):
            raise DataLogReadError('Invalid data type: "' + code + '".')
        dt = np.dtype('<f8' if code == 'd' else '<f4')
        if n * size * dt.itemsize > len(body):
            raise DataLogReadError(
                'Header indicates larger data size than found in body.')
        log = DataLog()
        for k, name in enumerate(fields):
            a = np.frombuffer(body, dtype=dt, count=size,
                              offset=k * size * dt.itemsize)
            log[name] = a.astype(dt.newbyteorder('='))
        if time:
            log._time = time
        return log

    def trim(self, n):
        """Keeps the first ``n`` entries of every column."""
        for k in list(self.keys()):
            v = self[k]
            if isinstance(v, np.ndarray):
                OrderedDict.__setitem__(self, k, v[:n])
            else:
                del v[n:]


class _MetaDict(dict):
    def __missing__(self, key):
        d = self[key] = {}
        return d


def prepare_log(log, model, dims=None, global_vars=None,
                if_empty=LOG_NONE, allowed_classes=LOG_ALL,
                precision=DOUBLE_PRECISION):
    """Returns a DataLog for a simulation, built from a flag, a sequence of
    names, or an existing log (checked for compatibility)."""
    dtype = np.float64 if precision == DOUBLE_PRECISION else np.float32
    dims = tuple(dims) if dims else ()
    _combos = []

    def dcombos():
        # Cell prefixes ('x.' or 'x.y.', first index fastest), built only
        # when a name is expanded over every cell
        if not _combos:
            if dims:
                _combos.extend(
                    '.'.join(str(i) for i in x) + '.' for x in dimco(*dims))
            else:
                _combos.append('')
        return _combos

    def valid_index(kdims):
        # 'x.' or 'x.y.' with every index in range
        parts = kdims[:-1].split('.')
        if len(parts) != len(dims):
            return False
        try:
            idx = [int(x) for x in parts]
        except ValueError:
            return False
        return all(0 <= i < n and str(i) == p
                   for i, n, p in zip(idx, dims, parts))
    global_vars = list(global_vars or [])
    for name in global_vars:
        try:
            v = model.get(name)
        except KeyError:
            raise ValueError(
                'Unknown variable specified in global_vars <' + name + '>.')
        if v.is_state():
            raise ValueError('State cannot be global variable.')

    def check_class(var):
        if var.is_constant():
            raise ValueError(
                f'This log does not support constants, got <{var}>.')
        elif var.is_state():
            if not LOG_STATE & allowed_classes:
                raise ValueError(
                    f'This log does not support state variables, got <{var}>.')
        elif var.is_bound():
            if not LOG_BOUND & allowed_classes:
                raise ValueError(
                    f'This log does not support bound variables, got <{var}>.')
        elif not LOG_INTER & allowed_classes:
            raise ValueError('This log does not support intermediary'
                             f' variables, got <{var}>.')

    def new():
        return np.zeros(0, dtype=dtype)

    if log is None:
        if if_empty & allowed_classes == if_empty:
            log = if_empty
        else:
            raise ValueError(
                'The option given for `if_empty` should be an allowed class.')

    if type(log) == int:
        flag = log
        log = DataLog()
        if flag == LOG_ALL:
            flag = allowed_classes
        if LOG_STATE & flag:
            if not LOG_STATE & allowed_classes:
                raise ValueError('This log does not support state variables.')
            for s in model.states():
                for c in dcombos():
                    log[c + s.qname()] = new()
            flag -= LOG_STATE
        if LOG_BOUND & flag:
            if not LOG_BOUND & allowed_classes:
                raise ValueError('This log does not support bound variables.')
            for label, var in model.bindings():
                name = var.qname()
                if name in global_vars:
                    log[name] = new()
                else:
                    for c in dcombos():
                        log[c + name] = new()
            flag -= LOG_BOUND
        if LOG_INTER & flag:
            if not LOG_INTER & allowed_classes:
                raise ValueError(
                    'This log does not support intermediary variables.')
            for var in model.variables(inter=True, deep=True):
                name = var.qname()
                if name in global_vars:
                    log[name] = new()
                else:
                    for c in dcombos():
                        log[c + name] = new()
            flag -= LOG_INTER
        if LOG_DERIV & flag:
            if not LOG_DERIV & allowed_classes:
                raise ValueError('This log does not support time-derivatives.')
            for var in model.states():
                for c in dcombos():
                    log['dot(' + c + var.qname() + ')'] = new()
            flag -= LOG_DERIV
        if flag != 0:
            raise ValueError('One or more unknown flags given as log.')
        t = model.time()
        if t is not None and t.qname() in log:
            log.set_time_key(t.qname())
        return log

    if isinstance(log, dict):
        if not isinstance(log, DataLog):
            log = DataLog(log)
        t = model.time()
        if t is not None and t.qname() in log:
            log.set_time_key(t.qname())
        log.validate()
        for key in list(log.keys()):
            deriv = key[0:4] == 'dot(' and key[-1:] == ')'
            k = key[4:-1] if deriv else key
            kdims, kname = split_key(k)
            try:
                var = model.get(kname)
            except KeyError:
                raise ValueError(f'Unknown variable <{kname}> in log.')
            if deriv:
                if not LOG_DERIV & allowed_classes:
                    raise ValueError(
                        f'This log does not support derivatives, got <{key}>.')
                if not var.is_state():
                    raise ValueError('Cannot log time derivative of non-state'
                                     f' <{var.qname()}>.')
            else:
                check_class(var)
            if kdims:
                if kname in global_vars:
                    raise ValueError('Cannot specify a cell index for global'
                                     f' logging variable <{kname}>.')
                if not valid_index(kdims):
                    raise ValueError(f'Invalid index <{kdims}> in log.')
            elif dims:
                if kname not in global_vars:
                    raise ValueError('DataLog contains non-indexed entry for'
                                     f' cell-specific variable <{kname}>.')
        for v in log.values():
            if not (isinstance(v, np.ndarray) or (
                    hasattr(v, 'append') and callable(v.append))):
                raise ValueError('Logging dict must map qnames to objects'
                                 ' that support the append() method.')
        return log

    try:
        if len(log) > 0:
            log[0]
    except Exception:
        raise ValueError('Argument `log` has unexpected type. Expecting None,'
                         ' integer flag, sequence of names, dict or DataLog.')
    if isinstance(log, str):
        raise ValueError('String passed in as `log` argument, should be list'
                         ' or other sequence containing strings.')
    lst = log
    log = DataLog()
    checked = {}
    from ._model import Variable
    for key in lst:
        if isinstance(key, Variable):
            key = key.qname()
        elif isinstance(key, ex.LhsExpression):
            key = str(key)
        deriv = key[0:4] == 'dot(' and key[-1:] == ')'
        if deriv:
            key = key[4:-1]
        kdims, kname = split_key(key)
        try:
            var = checked[kname]
        except KeyError:
            try:
                var = model.get(kname)
            except KeyError:
                raise ValueError(f'Unknown variable <{kname}> in list.')
            if deriv:
                if not LOG_DERIV & allowed_classes:
                    raise ValueError(
                        f'This log does not support derivatives, got <{key}>.')
                if not var.is_state():
                    raise ValueError('Cannot log time derivative of non-state'
                                     f' <{var.qname()}>.')
            else:
                check_class(var)
            checked[kname] = var
        if kdims:
            if kname in global_vars:
                raise ValueError('Cannot specify a cell index for global'
                                 f' logging variable <{kname}>.')
            if not valid_index(kdims):
                raise ValueError(f'Invalid index <{kdims}> in list.')
            k = kdims + kname
            log['dot(' + k + ')' if deriv else k] = new()
        elif kname in global_vars:
            log['dot(' + kname + ')' if deriv else kname] = new()
        else:
            for c in dcombos():
                k = c + kname
                log['dot(' + k + ')' if deriv else k] = new()
    t = model.time()
    if t is not None and t.qname() in log:
        log.set_time_key(t.qname())
    return log
