"""
Behaviour tests of SimulationOpenCL restated from the reference's own suite
(``myokit/tests/test_simulation_opencl.py``, class SimulationOpenCLTest,
with its fixtures: the Beeler-Reuter 1977 model and
``blocktrain(duration=2, offset=1, period=1000)`` on 1, 10 and 4 x 3
cells), run on the GPU: conductance fields in 1D and 2D, connections
against set_conductance in single precision, initial-value expressions, a
negative start time, protocol changes, run/pre/reset bookkeeping, progress
reporting and cancelling, and set_constant taking effect at the next run.
"""
import contextlib
import io
import os

import numpy as np
import pytest

import myokit_amd as myokit

pytestmark = pytest.mark.gpu

DATA = os.path.join(os.path.dirname(__file__), 'data')


def _model():
    return myokit.load_model(os.path.join(DATA, 'beeler-1977-model.mmt'))


def _protocol():
    return myokit.pacing.blocktrain(duration=2, offset=1, period=1000)


def _V(d, *index):
    return np.asarray(d['.'.join(str(i) for i in index) + '.membrane.V'])


class CancellingReporter(myokit.ProgressReporter):
    """Cancels after ``okays`` updates (the reference's test helper,
    ``myokit/tests/__init__.py:121-136``)."""

    def __init__(self, okays=0):
        self.okays = int(okays)

    def enter(self, msg=None):
        pass

    def exit(self):
        pass

    def update(self, f):
        self.okays -= 1
        return self.okays >= 0


def test_conductance_field_run_1d():
    # test_simulation_opencl.py:224-276
    s1 = myokit.SimulationOpenCL(_model(), _protocol(), ncells=10)
    s1.set_paced_cells(3)
    s1.set_conductance_field(np.ones(9))
    d = s1.run(15, log=['engine.time', 'membrane.V'], log_interval=0.1)
    for i in (0, 4, 7, 9):
        assert np.max(_V(d, i)) > 0
    s1.reset()
    s1.set_paced_cells(x=3, nx=3)
    x = np.ones(9)
    x[0] = 0
    x[-2:] = 0
    s1.set_conductance_field(x)
    d = s1.run(15, log=['engine.time', 'membrane.V'], log_interval=0.1)
    assert np.max(_V(d, 0)) < 0
    for i in range(1, 8):
        assert np.max(_V(d, i)) > 0, i
    assert np.max(_V(d, 8)) < 0
    assert np.max(_V(d, 9)) < 0


def test_conductance_field_run_2d():
    # test_simulation_opencl.py:278-322; shape (y=3, x=4)
    s2 = myokit.SimulationOpenCL(_model(), _protocol(), ncells=(4, 3))
    x = np.ones((3, 3))
    y = np.ones((2, 4))
    # Cell x=2, y=1 is not stimulated
    x[1, 1:3] = 0
    y[0:2, 2] = 0
    s2.set_paced_cells(2, 2)
    s2.set_conductance_field(x, y)
    d = s2.run(15, log=['engine.time', 'membrane.V'], log_interval=0.1)
    for iy in range(3):
        for ix in range(4):
            if (ix, iy) == (2, 1):
                assert np.max(_V(d, ix, iy)) < 0
            else:
                assert np.max(_V(d, ix, iy)) > 0, (ix, iy)


def test_connections_run_equals_conductance():
    # test_simulation_opencl.py:394-458: single precision, a chain of
    # connections against set_conductance(g, 0), within 1e-4 mV
    s1 = myokit.SimulationOpenCL(_model(), _protocol(), ncells=10)
    s1.set_paced_cells(3)
    g, t, dt = 5, 10, 0.1
    log = ['engine.time', 'membrane.V', 'membrane.i_diff']
    s1.set_conductance(g, 0)
    d1a = s1.run(t, log=log, log_interval=dt)
    s1.reset()
    s1.set_connections([(i, i + 1, g) for i in range(9)])
    d1b = s1.run(t, log=log, log_interval=dt)
    assert s1.engine_driver() == 'chain'
    for i in (0, 9):
        assert np.max(np.abs(_V(d1a, i) - _V(d1b, i))) < 1e-4


def test_initial_value_expressions():
    # test_simulation_opencl.py:660-683
    m = myokit.parse_model('''
        [[model]]
        c.x = 1 + sqrt(3)
        c.y = 1 / c.p
        c.z = 3

        [c]
        t = 0 bind time
        dot(x) = 1 label membrane_potential
        dot(y) = 2
        dot(z) = 3
        p = log(3)
        q = 0 bind diffusion_current
    ''')
    s = myokit.SimulationOpenCL(m, ncells=2)
    x = s.state()
    for k in range(3):
        assert isinstance(x[k], float)
    assert x[:3] == x[3:]
    assert x == m.initial_values(True) * 2
    assert x == s.default_state()
    # and the engine starts from those values (200 single-precision steps
    # of forward Euler: rounding of ~1e-7 per step)
    s.run(1.0, log=[])
    y = s.state()
    assert y[0] == pytest.approx(x[0] + 1, abs=1e-4)
    assert y[2] == pytest.approx(x[2] + 3, abs=1e-4)


def test_negative_time():
    # test_simulation_opencl.py:685-701
    s0 = myokit.SimulationOpenCL(_model(), _protocol(), ncells=1)
    p = myokit.pacing.blocktrain(duration=1, level=1, period=2)
    s0.set_protocol(p)
    s0.set_time(-10)
    assert s0.time() == -10
    d = s0.run(5, log=['engine.pace'])
    assert s0.time() == -5
    assert np.all(np.asarray(d['engine.pace']) == 0)
    d = s0.run(6, log=['engine.time', 'engine.pace'])
    assert s0.time() == 1
    pace = np.asarray(d['engine.pace'])
    assert np.all(pace[:-1] == 0)
    assert pace[-1] == 1


def test_protocol_changes():
    # test_simulation_opencl.py:825-859
    s0 = myokit.SimulationOpenCL(_model(), _protocol(), ncells=1)
    d0 = s0.run(3, log=['engine.pace'])
    assert np.max(d0['engine.pace']) == 1
    x = 0.0123
    s0.set_protocol(myokit.pacing.blocktrain(level=x, duration=1000,
                                             period=1000))
    s0.reset()
    d0 = s0.run(3, log=['engine.pace'])
    # the level as stored in single precision
    assert np.all(np.asarray(d0['engine.pace']) == np.float32(x))
    s0.reset()
    s0.set_protocol(None)
    d0 = s0.run(3, log=['engine.pace'])
    assert np.all(np.asarray(d0['engine.pace']) == 0)
    p = myokit.pacing.blocktrain(period=1000, duration=2, offset=3)
    s0.reset()
    s0.set_protocol(p)
    d0 = s0.run(3, log=['engine.pace'])
    assert np.max(d0['engine.pace']) == 0
    d0 = s0.run(3, log=['engine.pace'])
    assert np.max(d0['engine.pace']) == 1


def test_run_pre_reset_bookkeeping():
    # test_simulation_opencl.py:861-909
    s0 = myokit.SimulationOpenCL(_model(), _protocol(), ncells=1)
    x0 = s0.default_state()
    assert x0 == s0.state() and s0.time() == 0
    s0.run(1)
    x1 = s0.state()
    assert x0 == s0.default_state() and x0 != x1 and s0.time() == 1
    s0.reset()
    assert x0 == s0.state() and s0.time() == 0
    s0.pre(1)
    assert x1 == s0.default_state() == s0.state() and s0.time() == 0
    s0.run(1)
    assert x1 == s0.default_state() and x1 != s0.state()
    assert s0.time() == 1
    s0.reset()
    assert x1 == s0.state() and s0.time() == 0
    s0.set_time(10)
    s0.run(1)
    assert s0.time() == 11
    with pytest.raises(ValueError, match='negative'):
        s0.run(-1)


def test_run_progress_reporter():
    # test_simulation_opencl.py:911-925
    s0 = myokit.SimulationOpenCL(_model(), _protocol(), ncells=1)
    out = io.StringIO()
    with contextlib.redirect_stdout(out):
        s0.run(3, progress=myokit.ProgressPrinter())
    assert out.getvalue() != ''
    with pytest.raises(ValueError, match='progress'):
        s0.run(1, progress=True)
    with pytest.raises(myokit.SimulationCancelledError):
        s0.run(20, progress=CancellingReporter(0))


def test_set_constant_takes_effect_at_the_next_run():
    # test_simulation_opencl.py:927-950 (interface) and the effect the
    # reference's CVODES comparison relies on: a changed constant is used
    # by the next run, after runs with the old value (the kernel is
    # regenerated)
    s = myokit.SimulationOpenCL(_model(), _protocol(), ncells=1)
    s.set_step_size(0.001)
    a = _V(s.run(6.5, log=['engine.time', 'membrane.V'], log_interval=0.1),
           0)
    s.reset()
    s.set_constant('membrane.C', 1.5)
    b = _V(s.run(6.5, log=['engine.time', 'membrane.V'], log_interval=0.1),
           0)
    s.reset()
    s.set_constant('membrane.C', 1)
    c = _V(s.run(6.5, log=['engine.time', 'membrane.V'], log_interval=0.1),
           0)
    assert np.max(np.abs(a - b)) > 1
    np.testing.assert_array_equal(a, c)
    m2 = _model()
    v = m2['membrane'].add_variable('calculated_constant')
    v.set_rhs('3 * C')
    s2 = myokit.SimulationOpenCL(m2)
    s2.set_constant('membrane.C', 3)
    s2.set_constant(m2.get('membrane.C'), 2)
    for name in ('membrane.V', 'ina.INa', 'engine.time'):
        with pytest.raises(ValueError, match='not a literal'):
            s2.set_constant(name, 1)
    with pytest.raises(ValueError, match='not a literal'):
        s2.set_constant(v, 2)
