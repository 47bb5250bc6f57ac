"""
The C ABI rejects a simulation whose configuration does not describe the
kernel text it is created from: the module exports the layout it was
generated for (precision, states, fields, cells per thread of the sheet,
chain and reverse-sheet tiles) as ``mk_layout``, and ``mkh_sim_create``
compares ``mkh_sim_config`` with it. A caller of the reference has no such
pair to keep in step (openclsim.c builds the kernel from the same
arguments it runs), so a mismatch here is an error, never a launch with
the wrong tile shape.
"""
import ctypes

import pytest

import myokit_amd as myokit
from myokit_amd import _kernelgen, _native
from myokit_amd.simulation import _program

pytestmark = pytest.mark.gpu


def _config(s, **kw):
    cfg = _native.SimConfig()
    cfg.precision = 32
    cfg.n_state = s._nstate
    cfg.n_field = 0
    cfg.n_quant = 1
    cfg.i_vm = s._vm.index()
    cfg.coupling = 1
    cfg.driver = -1
    cfg.block_threads = 256
    cfg.nx, cfg.ny = 256, 128
    cfg.device = 0
    cfg.sheet_cpt = _kernelgen.sheet_cpt(32)
    cfg.chain_cpt = _kernelgen.chain_cpt(32)
    cfg.adj_cpt = _kernelgen.adjoint_sheet_cpt(32)
    for k, v in kw.items():
        setattr(cfg, k, v)
    return cfg


def _create(prog, cfg):
    h = ctypes.c_void_p()
    r = _native.lib().mkh_sim_create(prog, ctypes.byref(cfg), ctypes.byref(h))
    if r == 0:
        _native.lib().mkh_sim_destroy(h)
        return None
    return _native.last_error()


def test_config_must_match_the_kernel_layout():
    s = myokit.SimulationOpenCL(myokit.load_model('example'), None,
                                ncells=(256, 128), rl=True)
    src, cm = s._kernel_source([s._model.get('membrane.V')], adjoint=True)
    prog = _program(src, _kernelgen.compile_options(32), 0)
    assert _create(prog, _config(s)) is None
    # a zero turns a persistent driver off and is always accepted
    assert _create(prog, _config(s, sheet_cpt=0, chain_cpt=0,
                                 adj_cpt=0)) is None
    for key, value in (('sheet_cpt', 8), ('chain_cpt', 4), ('adj_cpt', 4),
                       ('n_state', s._nstate + 1), ('n_field', 2),
                       ('precision', 64)):
        err = _create(prog, _config(s, **{key: value}))
        print(key, value, '->', err, flush=True)
        assert err is not None and key in err and 'does not match' in err
