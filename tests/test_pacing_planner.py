"""
CPU tests of event-based pacing (Python PacingSystem and the library's C++
port) and of the time-step planner (the step / log rule of sim_step,
myokit/_sim/openclsim.c:1051-1162).

Pins: the reference's own known-answer sequences
(myokit/tests/test_pacing_system_c.py:21-120, restated below), sequences
recorded from the reference PacingSystem for random protocols
(tests/golden/pacing_py.json), and closed forms of the time grid.
"""
import ctypes
import json
import os

import numpy as np
import pytest

import myokit_amd as myokit
from myokit_amd import _native

GOLD = os.path.join(os.path.dirname(__file__), 'golden')


def c_pacing(protocol, times, t0=0.0):
    ev = np.ascontiguousarray(protocol.to_arrays().reshape(-1))
    if len(ev) == 0:
        ev = np.zeros(1)
    times = np.ascontiguousarray(times, dtype=np.float64)
    lev = np.zeros(len(times))
    nxt = np.zeros(len(times))
    r = _native.lib().mkh_pacing_levels(
        _native.ptr(ev, ctypes.c_double), len(protocol.events()), t0,
        _native.ptr(times, ctypes.c_double), len(times),
        _native.ptr(lev, ctypes.c_double), _native.ptr(nxt, ctypes.c_double))
    return r, lev, nxt


def plan(tmin, tmax, dt, protocol, log_interval, max_steps=10 ** 6):
    ev = np.ascontiguousarray(
        protocol.to_arrays().reshape(-1) if protocol else np.zeros(1))
    nev = len(protocol.events()) if protocol else 0
    t = np.zeros(max_steps)
    h = np.zeros(max_steps)
    p = np.zeros(max_steps)
    lg = np.zeros(max_steps, dtype=np.int32)
    n = ctypes.c_longlong(0)
    r = _native.lib().mkh_plan(
        tmin, tmax, dt, _native.ptr(ev, ctypes.c_double), nev, log_interval,
        _native.ptr(t, ctypes.c_double), _native.ptr(h, ctypes.c_double),
        _native.ptr(p, ctypes.c_double), _native.ptr(lg, ctypes.c_int),
        max_steps, ctypes.byref(n))
    assert r == 0
    k = n.value
    return t[:k], h[:k], p[:k], lg[:k].astype(bool)


def test_kat_event_at_t0():
    # test_pacing_system_c.py:21-53
    p = myokit.Protocol()
    p.schedule(2, 0, 1, 10, 0)
    times = [0, 0, 0, 0, 0.5, 1, 2, 10]
    r, lev, nxt = c_pacing(p, times)
    assert r == 0
    assert list(lev) == [2, 2, 2, 2, 2, 0, 0, 2]
    assert list(nxt) == [1, 1, 1, 1, 1, 10, 10, 11]
    s = myokit.PacingSystem(p)
    for t, l, n in zip(times, lev, nxt):
        assert s.advance(t) == l
        assert s.next_time() == n


def test_kat_event_at_t1_and_empty():
    # test_pacing_system_c.py:55-66
    p = myokit.Protocol()
    p.schedule(2, 1, 1, 10, 0)
    r, lev, nxt = c_pacing(p, [0])
    assert (lev[0], nxt[0]) == (0, 1)
    r, lev, nxt = c_pacing(myokit.Protocol(), [0])
    assert lev[0] == 0 and nxt[0] > 1e123


def test_kat_simultaneous_reoccurrence():
    # test_pacing_system_c.py:68-90: the recurring event clashes at 3000
    p = myokit.Protocol()
    p.schedule(1, 0, 1, 1000)
    p.schedule(1, 3000, 1)
    r, lev, nxt = c_pacing(p, [1, 1000, 1001])
    assert r == 0 and list(nxt) == [1000, 1001, 2000]
    r, lev, nxt = c_pacing(p, [1, 1000, 1001, 2000])
    assert r == _native.MKH_E_PACING_SIMULTANEOUS
    s = myokit.PacingSystem(p)
    for t in (1, 1000, 1001):
        s.advance(t)
    with pytest.raises(myokit.SimultaneousProtocolEventError):
        s.advance(2000)


def test_kat_negative_start_time():
    # test_pacing_system_c.py:92-110
    p = myokit.pacing.blocktrain(level=1, duration=1, period=2)
    r, lev, nxt = c_pacing(p, [-100], t0=-100)
    assert (lev[0], nxt[0]) == (0, 0)
    p = myokit.pacing.blocktrain(level=1, duration=1, period=2, offset=1)
    r, lev, nxt = c_pacing(p, [-100, 1], t0=-100)
    assert (lev[0], nxt[0]) == (0, 1)
    assert lev[1] == 1


def test_against_reference_sequences():
    with open(os.path.join(GOLD, 'pacing_py.json')) as f:
        cases = json.load(f)
    checked = 0
    for case in cases:
        p = myokit.Protocol()
        for row in case['events']:
            p.schedule(*row)
        seq = np.array(case['sequence']).reshape(-1, 3)
        s = myokit.PacingSystem(p)
        if case['error']:
            # Fails when advancing to the time after the recorded sequence
            r, lev, nxt = c_pacing(p, case['times'][:len(seq) + 1])
            assert r == _native.MKH_E_PACING_SIMULTANEOUS
            r, lev, nxt = c_pacing(p, case['times'][:len(seq)])
            assert r == 0
            np.testing.assert_array_equal(lev, seq[:, 2])
            continue
        r, lev, nxt = c_pacing(p, seq[:, 0])
        assert r == 0
        np.testing.assert_array_equal(lev, seq[:, 2])
        np.testing.assert_array_equal(nxt, seq[:, 1])
        for t, n, l in seq:
            assert s.advance(t) == l
            assert s.next_time() == n
        checked += 1
    assert checked > 10


def test_plan_uniform_grid_and_logs():
    t, h, p, lg = plan(0, 10, 0.01, None, 0.5)
    assert len(t) == 1000
    # dt = tmin + k*dt0 - t, so every step ends on the grid
    np.testing.assert_allclose(t + h, 0.01 * np.arange(1, 1001), atol=1e-12)
    np.testing.assert_allclose(t[lg], 0.5 * np.arange(20), atol=1e-9)


def test_plan_event_splits_step():
    # An event at 1.005 (off the 0.01 grid) forces two intermediary steps
    p = myokit.Protocol()
    p.schedule(1, 1.005, 0.5)
    t, h, pc, lg = plan(0, 2, 0.01, p, 1.0)
    k = np.nonzero(np.isclose(t, 1.005))[0]
    assert len(k) == 1
    assert pc[k[0]] == 1 and pc[k[0] - 1] == 0
    assert h[k[0] - 1] == pytest.approx(0.005)
    assert h[k[0]] == pytest.approx(0.005)
    # pacing ends at 1.505, also off the grid
    assert np.any(np.isclose(t, 1.505))
    assert len(t) == 200 + 2


def test_plan_tmax_and_half_open_log():
    # The run ends exactly at tmax; the last log point is before tmax
    t, h, p, lg = plan(3.0, 4.003, 0.01, None, 0.25)
    assert t[-1] + h[-1] == pytest.approx(4.003)
    assert t[lg][-1] == pytest.approx(4.0)
    np.testing.assert_allclose(t[lg], 3 + 0.25 * np.arange(5), atol=1e-9)


def test_plan_matches_oracle_log_times():
    # The library planner and the oracle's restated loop log the same times
    from oracle.cgen import OracleSimulation
    p = myokit.Protocol()
    p.schedule(1, 0.3, 0.7, 5.2, 3)
    o = OracleSimulation(myokit.load_model('example'), p, ncells=1,
                         diffusion=False)
    o.dt = 0.013
    r = o.run(20, log_vars=['membrane.V'], log_interval=0.37)
    t, h, pc, lg = plan(0, 20, 0.013, p, 0.37)
    np.testing.assert_array_equal(t[lg], r['time'])
    np.testing.assert_array_equal(pc[lg], r['pace'])
    assert r['steps'] == len(t)


def test_value_at_times():
    p = myokit.pacing.blocktrain(10, 2, offset=1)
    v = p.value_at_times([0, 1, 2.9, 3, 11, 12.5, 13])
    assert v == [0, 1, 1, 0, 1, 1, 0]
