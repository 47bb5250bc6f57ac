"""
CPU tests of the model layer: parser, expression evaluation and derivatives,
Rush-Larsen gate detection, bindings. Pinned to fixtures written by the
reference (tests/golden/make_golden.py) and to closed forms.
"""
import json
import math
import os

import numpy as np
import pytest

import myokit_amd as myokit
from myokit_amd.lib import hh

HERE = os.path.dirname(__file__)
GOLD = os.path.join(HERE, 'golden')
MODELS = {'lr1991': 'example',
          'br1977': os.path.join(HERE, 'data', 'beeler-1977-model.mmt')}


@pytest.mark.parametrize('name', ['lr1991', 'br1977'])
def test_derivatives_match_reference(name):
    # Reference Model.evaluate_derivatives at random states and inputs
    g = np.load(os.path.join(GOLD, 'derivatives.npz'))
    m = myokit.load_model(MODELS[name])
    for x, (pace, idiff), d in zip(g[name + '_states'], g[name + '_inputs'],
                                   g[name + '_derivs']):
        r = m.evaluate_derivatives(
            x, {'pace': pace, 'diffusion_current': idiff, 'time': 0})
        np.testing.assert_allclose(r, d, rtol=1e-12, atol=1e-300)


@pytest.mark.parametrize('name', ['lr1991', 'br1977'])
def test_rl_detection_matches_reference(name):
    with open(os.path.join(GOLD, 'rl_forms.json')) as f:
        gold = json.load(f)[name]
    m = myokit.load_model(MODELS[name])
    vm = m.label('membrane_potential')
    for s in m.states():
        r = hh.get_alpha_and_beta(s, vm)
        r = None if r is None else [r[0].qname(), r[1].qname()]
        assert r == gold['alpha_beta'][s.qname()]
    c = hh.convert_hh_states_to_inf_tau_form(m, vm)
    v = c.get(vm.qname())
    for s in c.states():
        r = hh.get_inf_and_tau(s, v)
        r = None if r is None else [r[0].qname(), r[1].qname()]
        assert r == gold['inf_tau'][s.qname()]


def test_rl_expression_closed_form():
    # tau = 1 / (alpha + beta), inf = alpha * tau; the RL update of a gate
    # with constant alpha, beta is the exact solution of the linear ODE
    m = myokit.load_model('example')
    c = hh.convert_hh_states_to_inf_tau_form(m)
    h = c.get('ina.h')
    inf, tau = hh.get_inf_and_tau(h)
    x = c.state()
    subst = {s: v for s, v in zip(c.states(), x)}
    a = c.get('ina.h.alpha').rhs().eval(subst)
    b = c.get('ina.h.beta').rhs().eval(subst)
    assert tau.rhs().eval(subst) == pytest.approx(1 / (a + b), rel=1e-15)
    assert inf.rhs().eval(subst) == pytest.approx(a / (a + b), rel=1e-15)
    e = hh.get_rl_expression(h, myokit.Name('dt_symbol'))
    assert e is not None


def test_state_order_bindings_labels():
    m = myokit.load_model('example')
    assert [s.qname() for s in m.states()] == [
        'membrane.V', 'ina.m', 'ina.h', 'ina.j', 'ica.d', 'ica.f', 'ik.x',
        'ica.Ca_i']
    assert m.label('membrane_potential').qname() == 'membrane.V'
    assert m.time().qname() == 'engine.time'
    assert m.binding('pace').qname() == 'engine.pace'
    assert m.binding('diffusion_current').qname() == 'membrane.i_diff'
    assert m.get('ik.IK.gK').is_constant()
    assert not m.get('ik.IK.gK').is_literal()
    assert m.get('ina.gNa').is_literal()
    assert m.get('ina.INa').is_intermediary()
    assert not m.has_interdependent_components()


def test_protocol_in_file():
    _, p, _ = myokit.load('example')
    e = p.events()
    assert len(e) == 1
    assert (e[0].level(), e[0].start(), e[0].duration(), e[0].period(),
            e[0].multiplier()) == (1.0, 50.0, 0.5, 1000.0, 0)


def test_clone_is_independent():
    m = myokit.load_model('example')
    c = m.clone()
    c.get('ina.gNa').set_rhs(1.0)
    assert m.get('ina.gNa').rhs().eval() == 16
    assert c.evaluate_derivatives() != m.evaluate_derivatives()


def test_parse_errors():
    with pytest.raises(myokit.ParseError):
        myokit.parse_model('[[model]]\nc.x = 1\n[c]\ndot(x) = 1 +\n')
    with pytest.raises(myokit.ParseError):
        myokit.parse_model('[[model]]\n[c]\nx = foo(3)\n')
    with pytest.raises(myokit.IntegrityError):
        myokit.parse_model('[[model]]\n[c]\nx = y\n')
    with pytest.raises(myokit.CyclicalDependencyError):
        myokit.parse_model('[[model]]\n[c]\nx = y\ny = x\n')


def test_syntax_features():
    src = '''[[model]]
name: test
desc: """
Multi-line
description
"""
c.x = 2
[c]
use d.k as kk
dot(x) = -kk * x + a \\
    + b
    a = 1 [mV]
        in [mV]
    b = (3 +
         4) * 2
    desc: nested
y = if(x > 1, 1, 0) + piecewise(x < 0, -1, x < 1, 0, 1) : A description
z = 7 // 2 + 7 % 2 + 2^3^2 + log(8, 2) + log10(100) + abs(-2)
    label zz
t = 0 bind time
[d]
k = 0.5
'''
    m = myokit.parse_model(src)
    assert m.meta['desc'] == 'Multi-line\ndescription'
    x = m.get('c.x')
    assert x.is_state()
    assert m.get('c.x.b').rhs().eval() == 14
    assert m.get('c.x.a').unit() == 'mV'
    assert m.get('c.y').meta['desc'] == 'A description'
    assert m.get('c.y').rhs().eval({x: 2}) == 2
    # 7 // 2 = 3, 7 % 2 = 1, 2^3^2 = 2^9 (right-associative), log2(8) = 3
    assert m.get('c.z').rhs().eval() == pytest.approx(3 + 1 + 512 + 3 + 2 + 2)
    assert m.label('zz') is m.get('c.z')
    assert m.evaluate_derivatives([2]) == [pytest.approx(-0.5 * 2 + 1 + 14)]


def test_quotient_remainder_semantics():
    # Rounding towards minus infinity (myokit.Quotient/Remainder)
    e = myokit.parse_expression('-7 // 3')
    assert e.eval() == -3
    e = myokit.parse_expression('5 // -3')
    assert e.eval() == -2
    e = myokit.parse_expression('-7 % 3')
    assert e.eval() == 2


def test_symbolic_derivative_against_finite_difference():
    # d/dV of every intermediary of LR1991, chain rule through the model,
    # against a central difference
    m = myokit.load_model('example')
    V = m.get('membrane.V')
    nV = myokit.Name(V)
    x = m.state()
    base = {s: v for s, v in zip(m.states(), x)}
    for var in ['ina.INa', 'ik.IK', 'ik1.IK1', 'ica.ICa', 'ikp.IKp',
                'ina.h.beta', 'ik.xi']:
        e = m.get(var).rhs()
        d = e.diff(nV).eval(base)
        h = 1e-6
        up = dict(base)
        dn = dict(base)
        up[V] = x[0] + h
        dn[V] = x[0] - h
        fd = (e.eval(up) - e.eval(dn)) / (2 * h)
        assert d == pytest.approx(fd, rel=1e-6, abs=1e-12), var


def test_diff_closed_forms():
    x = myokit.Name('x')
    m = myokit.Model()
    c = m.add_component('c')
    v = c.add_variable('x')
    v.set_rhs(0)
    n = myokit.Name(v)
    cases = [
        (myokit.Exp(myokit.Multiply(myokit.Number(2), n)),
         lambda t: 2 * math.exp(2 * t)),
        (myokit.Log(n), lambda t: 1 / t),
        (myokit.Power(n, myokit.Number(3)), lambda t: 3 * t * t),
        (myokit.Divide(myokit.Number(1), n), lambda t: -1 / t ** 2),
        (myokit.Sqrt(n), lambda t: 0.5 / math.sqrt(t)),
        (myokit.Sin(n), math.cos),
        (myokit.ATan(n), lambda t: 1 / (1 + t * t)),
        (myokit.Power(n, n), lambda t: t ** t * (math.log(t) + 1)),
    ]
    for e, f in cases:
        for t in (0.3, 1.7):
            assert e.local_diff(n).eval({n: t}) == pytest.approx(f(t),
                                                                 rel=1e-12)
    del x


def test_prepare_bindings_and_unused():
    from myokit_amd.simulation import _prepare_model
    m = myokit.load_model(os.path.join(HERE, 'data', 'lr-1991.mmt'))
    mm, vm, rl, bound = _prepare_model(m, rl=False, diffusion=True)
    assert sorted(bound.values()) == ['idiff', 'pace', 'time']
    # Unsupported bindings are dropped and keep their value
    assert mm.binding('realtime') is None
    assert mm.get('engine.realtime').is_literal()
    # Bound pace is zeroed (it starts at 1 in this file)
    assert mm.get('engine.pace').rhs().eval() == 0
