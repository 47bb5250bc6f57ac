"""
CPU tests of the C-ABI library and of the HIP kernel generator: the library
loads and exports every symbol include/myokit_amd.h declares; generated
kernels compile for gfx950 with hiprtc (no device needed); the writer's
literals and operator forms are exact.
"""
import ctypes
import math
import os
import re

import numpy as np
import pytest

import myokit_amd as myokit
from myokit_amd import _kernelgen, _native
from myokit_amd import _expressions as ex
from myokit_amd.simulation import _prepare_cell_model

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    with open(os.path.join(ROOT, 'include', 'myokit_amd.h')) as f:
        text = f.read()
    return sorted(set(re.findall(r'\b(mkh_\w+)\s*\(', text)))


def test_library_exports_every_declared_symbol():
    L = _native.lib()
    names = declared_symbols()
    assert len(names) >= 28
    for n in names:
        assert hasattr(L, n), n
        assert n in _native.SIGNATURES, n
    assert L.mkh_version().decode().startswith('myokit_amd')


def test_no_device_is_an_error_not_a_fallback():
    if _native.device_count() > 0:
        pytest.skip('a device is present')
    s = myokit.SimulationOpenCL(myokit.load_model('example'), ncells=4)
    with pytest.raises(myokit.DeviceError):
        s.run(1)


@pytest.mark.parametrize('precision,coupling,rl', [
    (32, 1, True), (64, 1, False), (32, 2, False), (64, 3, True),
    (32, 0, False)])
def test_generated_kernel_compiles_for_gfx950(precision, coupling, rl):
    cm = _prepare_cell_model(myokit.load_model('example'), rl=rl,
                             diffusion=coupling != 0,
                             fields=['ina.gNa', 'ik.IK.gK'],
                             quantities=['membrane.V', 'ina.INa'])
    src = _kernelgen.generate_kernel(cm, precision, coupling)
    prog = ctypes.c_void_p()
    log = ctypes.create_string_buffer(1 << 16)
    r = _native.lib().mkh_program_build(
        src.encode(), _kernelgen.compile_options(precision).encode(), None,
        -1, ctypes.byref(prog), log, 1 << 16)
    assert r == 0, _native.last_error()
    _native.lib().mkh_program_free(prog)


def test_literals_are_exact():
    for v in (0.1, -84.5286, 1e-300, 3.1e5, 2.0 ** -40):
        s = _kernelgen.real_literal(v, 64)
        assert float.fromhex(s.strip('()')) == v
        s = _kernelgen.real_literal(v, 32)
        assert float.fromhex(s.strip('()')[:-1]) == float(np.float32(v))


def test_writer_forms():
    w = _kernelgen.HipExpressionWriter(64)
    m = myokit.Model()
    c = m.add_component('c')
    a = c.add_variable('a')
    b = c.add_variable('b')
    a.set_rhs(1)
    b.set_rhs(2)
    na, nb = ex.Name(a), ex.Name(b)
    w.names = {na: 'A', nb: 'B'}
    assert w.ex(ex.Quotient(na, nb)) == '__builtin_floor((A / B))'
    assert w.ex(ex.Remainder(na, nb)) == '(A - B * __builtin_floor(A / B))'
    assert w.ex(ex.Power(na, ex.Number(3))) == 'mk_powi(A, 3)'
    assert w.ex(ex.Power(na, nb)) == 'mk_pow(A, B)'
    assert w.ex(ex.Log(na, nb)) == '(mk_log(A) / mk_log(B))'
    assert w.ex(ex.If(ex.Less(na, nb), na, nb)) == '((A < B) ? A : B)'
    pw = ex.Piecewise(ex.Less(na, nb), na, ex.More(na, nb), nb,
                      ex.Number(0))
    assert w.ex(pw) == ('((A < B) ? A : ((A > B) ? B : '
                        + _kernelgen.real_literal(0, 64) + '))')
    # Single precision division by a variable uses rcp, by a literal the
    # reciprocal literal
    w32 = _kernelgen.HipExpressionWriter(32)
    w32.names = w.names
    assert w32.ex(ex.Divide(na, nb)) == '(A * __builtin_amdgcn_rcpf(B))'
    assert w32.ex(ex.Divide(na, ex.Number(4))) == '(A * 0x1.0000000000000p-2f)'


def test_affine_exponent_folding():
    # exp(k*x + m) becomes exp2(fma(x, k*log2e, m*log2e)) in single
    # precision; check the folded constants reproduce the argument
    cm = _prepare_cell_model(myokit.load_model('example'), rl=False,
                             diffusion=True, fields=[], quantities=[])
    m = cm.model
    w = _kernelgen.HipExpressionWriter(32)
    w.constants = cm.constants
    V = m.get('membrane.V')
    w.names = {ex.Name(V): 'V'}
    for e in [m.get('ina.a').rhs(), m.get('ik.x.alpha').rhs(),
              m.get('ina.h.alpha').rhs(), m.get('ik1.g.beta').rhs()]:
        for node in e.walk():
            if isinstance(node, ex.Exp):
                res = w.affine(node[0])
                if res is None:
                    continue
                name, k, c = res
                assert name == 'V'
                for v in (-80.0, -20.0, 30.0):
                    want = node[0].eval({V: v})
                    assert k * v + c == pytest.approx(want, rel=1e-12,
                                                      abs=1e-12)
    # Not affine: division by a variable
    a = ex.Divide(ex.Name(V), ex.Name(V))
    assert w.affine(a) is None


def test_constants_evaluated_on_host():
    cm = _prepare_cell_model(myokit.load_model('example'), rl=False,
                             diffusion=True, fields=['cell.K_o'],
                             quantities=['membrane.V'])
    names = {v.qname() for v in cm.per_cell_constants}
    # gK and E of ik depend on the K_o field: computed per cell
    assert 'ik.IK.gK' in names and 'ik.IK.E' in names
    cons = {v.qname(): x for v, x in cm.constants.items()}
    assert cons['cell.RTF'] == pytest.approx(8314 * 310 / 96500, rel=1e-15)
    assert cons['ina.ENa'] == pytest.approx(
        8314 * 310 / 96500 * math.log(140 / 10), rel=1e-15)
