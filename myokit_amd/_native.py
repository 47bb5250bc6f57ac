"""
ctypes binding to ``libmyokit_amd.so`` (C ABI: ``include/myokit_amd.h``).

The library is built in-tree by ``myokit_amd/_build.py``. There is no
fallback: if the library is missing, or no HIP device is present when a
simulation is created, an error is raised.
"""
import atexit
import ctypes
import os
import threading
import weakref

from ._err import DeviceError

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                        'libmyokit_amd.so')

MKH_E_NAN = -5
MKH_E_PACING_SIMULTANEOUS = -50
MKH_E_PACING_NEGATIVE_TIME = -40

COUPLING_NONE = 0
COUPLING_STENCIL = 1
COUPLING_HETERO = 2
COUPLING_GRAPH = 3

DRIVER_CELLS = 0
DRIVER_BLOCK = 1
DRIVER_GRID = 2
DRIVER_SHEET = 3
DRIVER_CHAIN = 4

# Every entry point declared in include/myokit_amd.h
_c = ctypes
_vp = _c.c_void_p
_ll = _c.c_longlong
_d = _c.c_double
_i = _c.c_int
_pd = _c.POINTER(_c.c_double)
_pll = _c.POINTER(_c.c_longlong)
_pi = _c.POINTER(_c.c_int)


class SimConfig(ctypes.Structure):
    _fields_ = [
        ('precision', _i), ('n_state', _i), ('n_field', _i),
        ('n_quant', _i), ('i_vm', _i), ('coupling', _i), ('driver', _i),
        ('block_threads', _i), ('nx', _ll), ('ny', _ll), ('device', _i),
        ('sheet_cpt', _i), ('chain_cpt', _i), ('adj_cpt', _i),
    ]


SIGNATURES = {
    'mkh_version': (_c.c_char_p, []),
    'mkh_last_error': (_c.c_char_p, []),
    'mkh_device_count': (_i, [_pi]),
    'mkh_device_info': (_i, [_i, _c.c_char_p, _i, _pll, _pi]),
    'mkh_program_build': (_i, [_c.c_char_p, _c.c_char_p, _c.c_char_p, _i,
                               _c.POINTER(_vp), _c.c_char_p, _i]),
    'mkh_program_free': (_i, [_vp]),
    'mkh_sim_create': (_i, [_vp, _c.POINTER(SimConfig), _c.POINTER(_vp)]),
    'mkh_sim_destroy': (_i, [_vp]),
    'mkh_sim_set_state': (_i, [_vp, _pd, _ll]),
    'mkh_sim_get_state': (_i, [_vp, _pd, _ll]),
    'mkh_sim_set_fields': (_i, [_vp, _pd, _ll]),
    'mkh_sim_set_conductance': (_i, [_vp, _d, _d]),
    'mkh_sim_set_conductance_field': (_i, [_vp, _pd, _ll, _pd, _ll]),
    'mkh_sim_set_connections': (_i, [_vp, _pll, _pll, _pd, _ll]),
    'mkh_sim_set_domains': (_i, [_vp, _pll, _ll]),
    'mkh_sim_set_pacing_all': (_i, [_vp]),
    'mkh_sim_set_pacing_none': (_i, [_vp]),
    'mkh_sim_set_pacing_rect': (_i, [_vp, _ll, _ll, _ll, _ll]),
    'mkh_sim_set_pacing_list': (_i, [_vp, _pll, _ll]),
    'mkh_sim_set_logging': (_i, [_vp, _pi, _ll, _i, _i]),
    'mkh_sim_set_nan_cell': (_i, [_vp, _ll]),
    'mkh_sim_init': (_i, [_vp, _d, _d, _d, _pd, _ll, _d, _pll]),
    'mkh_sim_set_log': (_i, [_vp, _vp, _pd, _pd]),
    'mkh_sim_step': (_i, [_vp, _pd, _pi]),
    'mkh_sim_clean': (_i, [_vp]),
    'mkh_sim_set_option': (_i, [_vp, _c.c_char_p, _ll]),
    'mkh_sim_get_timing': (_i, [_vp, _pd, _pll, _pll]),
    'mkh_sim_reset_timing': (_i, [_vp]),
    'mkh_sim_log_rows': (_i, [_vp, _pll]),
    'mkh_sim_driver': (_i, [_vp, _pi]),
    'mkh_adj_run': (_i, [_vp, _d, _d, _d, _pd, _ll, _pll, _ll, _pd, _ll, _i,
                         _pd, _pd, _pd]),
    'mkh_adj_timing': (_i, [_vp, _pd]),
    'mkh_adj_reverse': (_i, [_vp, _pi, _pll]),
    'mkh_adj_set_cell_order': (_i, [_vp, _pll, _ll]),
    'mkh_sim_copy_state': (_i, [_vp, _vp]),
    'mkh_ft_create': (_i, [_vp, _vp, _ll, _ll, _d,
                           ctypes.POINTER(ctypes.c_void_p)]),
    'mkh_ft_step': (_i, [_vp, _pd, ctypes.POINTER(ctypes.c_int)]),
    'mkh_ft_destroy': (_i, [_vp]),
    'mkh_host_alloc': (_i, [_ll, ctypes.POINTER(ctypes.c_void_p)]),
    'mkh_host_free': (_i, [_vp]),
    'mkh_adj_trace': (_i, [_vp, _d, _d, _d, _pd, _ll, _pll, _ll, _pd, _ll]),
    'mkh_adj_trace_static': (_i, [_vp, _d, _d, _d, _pd, _ll, _pll, _ll, _pd,
                                  _ll]),
    'mkh_pacing_levels': (_i, [_pd, _ll, _d, _pd, _ll, _pd, _pd]),
    'mkh_plan': (_i, [_d, _d, _d, _pd, _ll, _d, _pd, _pd, _pd, _pi, _ll,
                      _pll]),
}

_lib = None

# Set once the library's handles have been released at interpreter exit:
# after that, finalizers that run late (engine objects collected during
# interpreter teardown, page-locked blocks of arrays still referenced) make
# no further calls into the library or the HIP runtime.
closed = False

# Live device handles, released by _shutdown: a weak reference to each
# owner (an object with a handle ``h``) and the function that destroys the
# handle (e.g. mkh_sim_destroy), keyed by id of the owner.
_handles = {}
_programs = []


def track(owner, destroy):
    """Registers ``owner.h`` to be released by ``destroy(h)`` at interpreter
    exit, unless the owner is collected (and calls ``untrack``) first."""
    _handles[id(owner)] = (weakref.ref(owner), destroy)


def untrack(owner):
    _handles.pop(id(owner), None)


def track_program(p):
    _programs.append(p)


def _shutdown():
    """Releases every engine, program and pooled page-locked block while the
    HIP runtime (and any profiler attached to it) is still up. Registered
    with atexit when the library is loaded, so it runs before the C
    runtime's exit handlers."""
    global closed
    if _lib is None or closed:
        return
    for ref, destroy in list(_handles.values()):
        owner = ref()
        h = getattr(owner, 'h', None)
        if h:
            destroy(h)
            owner.h = None
    _handles.clear()
    for p in _programs:
        _lib.mkh_program_free(p)
    del _programs[:]
    pinned.release_free()
    closed = True


def lib():
    """Loads the library (once) and declares every signature."""
    global _lib
    if _lib is None:
        if not os.path.isfile(LIB_PATH):
            raise ImportError(
                'libmyokit_amd.so not found at ' + LIB_PATH + '. Build it'
                ' with `python -m myokit_amd._build` (or'
                ' __graft_entry__.build()).')
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
        atexit.register(_shutdown)
    if closed:
        raise DeviceError('The device library was released at exit.')
    return _lib


def last_error():
    return lib().mkh_last_error().decode('utf-8', 'replace')


def check(code, what=''):
    """Raises DeviceError for a non-zero return code."""
    if code != 0:
        raise DeviceError((what + ': ' if what else '') + last_error()
                          + ' (code ' + str(code) + ')')
    return code


def device_count():
    n = ctypes.c_int(0)
    lib().mkh_device_count(ctypes.byref(n))
    return n.value


def device_info(device=0):
    name = ctypes.create_string_buffer(256)
    mem = ctypes.c_longlong(0)
    cu = ctypes.c_int(0)
    check(lib().mkh_device_info(device, name, 256, ctypes.byref(mem),
                                ctypes.byref(cu)), 'mkh_device_info')
    return {'name': name.value.decode(), 'memory': mem.value,
            'compute_units': cu.value}


def ptr(array, ctype):
    """Pointer to a contiguous numpy array's data."""
    return array.ctypes.data_as(ctypes.POINTER(ctype))


class _PinnedPool:
    """Page-locked host blocks (``mkh_host_alloc``) for log outputs, so that
    log rows reach the host while later steps compute. An array from
    :meth:`empty` gives its block back to the pool when the last view of it
    is gone; the pool keeps at most ``keep`` bytes of free blocks."""

    GRANULE = 1 << 20

    def __init__(self, keep=2 << 30):
        self.keep = keep
        self._free = []         # (size, address)
        self._kept = 0
        self._lock = threading.Lock()

    def empty(self, shape, dtype):
        import numpy as np
        dtype = np.dtype(dtype)
        count = 1
        for d in shape:
            count *= int(d)
        nbytes = count * dtype.itemsize
        if nbytes == 0:
            return np.empty(shape, dtype)
        size = -(-nbytes // self.GRANULE) * self.GRANULE
        addr = None
        with self._lock:
            # smallest free block that fits, if not more than 1.5x the need
            best = None
            for k, (sz, a) in enumerate(self._free):
                if size <= sz <= size * 3 // 2 and (
                        best is None or sz < self._free[best][0]):
                    best = k
            if best is not None:
                size, addr = self._free.pop(best)
                self._kept -= size
        if addr is None:
            p = ctypes.c_void_p()
            check(lib().mkh_host_alloc(size, ctypes.byref(p)),
                  'mkh_host_alloc')
            addr = p.value
        buf = (ctypes.c_char * size).from_address(addr)
        weakref.finalize(buf, self._give_back, addr, size)
        return np.frombuffer(buf, dtype=dtype, count=count).reshape(shape)

    def _give_back(self, addr, size):
        if closed:
            return      # released with the process
        with self._lock:
            if self._kept + size <= self.keep:
                self._free.append((size, addr))
                self._kept += size
                return
        if _lib is not None:
            _lib.mkh_host_free(addr)

    def release_free(self):
        """Frees the pooled blocks that no array uses."""
        with self._lock:
            blocks, self._free, self._kept = self._free, [], 0
        for size, addr in blocks:
            _lib.mkh_host_free(addr)


pinned = _PinnedPool()
