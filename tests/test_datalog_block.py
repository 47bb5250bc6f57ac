"""
DataLog and DataBlock1d/2d: file formats (myokit's zip formats), views and
conduction velocity. The reference's own fixture ``cv1d.zip`` (a
DataBlock1d file from ``myokit/tests/data``) with its expected velocity
(``test_datablock.py:138-143``) pins the block reader and ``cv``.
"""
import os
import zipfile

import numpy as np
import pytest

import myokit_amd as myokit

DATA = os.path.join(os.path.dirname(__file__), 'data')


def test_cv_matches_reference_fixture():
    b = myokit.DataBlock1d.load(os.path.join(DATA, 'cv1d.zip'))
    assert b.shape() == (500, 64)
    assert list(b.keys0d()) == ['engine.pace']
    assert b.cv('membrane.V') == pytest.approx(5.95272837350686004e+01,
                                               rel=1e-12)
    with pytest.raises(ValueError):
        b.cv('membrane.V', border=-1)
    with pytest.raises(ValueError):
        b.cv('membrane.V', border=1000)


def test_cv_closed_form_and_no_upstroke():
    # Cell x jumps from -85 to +20 mV at t = 1 + x / v: every activation
    # time carries the same interpolation offset, so the least-squares
    # velocity is v (cells/ms; 0.01 cm per cell, ms -> s: x 10 cm/s)
    t = np.linspace(0, 50, 5001)
    nx, v = 40, 2.0
    V = np.where(t[:, None] >= np.arange(nx)[None, :] / v + 1.0, 20.0, -85.0)
    b = myokit.DataBlock1d(nx, t)
    b.set1d('V', V)
    cv = b.cv('V', border=5)
    assert cv == pytest.approx(v * 0.01 / 1e-3, rel=1e-3)
    b.set1d('flat', np.full((len(t), nx), -85.0))
    assert b.cv('flat') == 0


def test_block1d_save_load_round_trip(tmp_path):
    b = myokit.DataBlock1d.load(os.path.join(DATA, 'cv1d.zip'))
    p = str(tmp_path / 'b.zip')
    b.save(p)
    c = myokit.DataBlock1d.load(p)
    np.testing.assert_array_equal(c.time(), b.time())
    np.testing.assert_array_equal(c.get1d('membrane.V'), b.get1d('membrane.V'))
    np.testing.assert_array_equal(c.get0d('engine.pace'),
                                  b.get0d('engine.pace'))
    with zipfile.ZipFile(p) as z:
        head = z.read('header_block1d.txt').decode().splitlines()
    assert head == ['500', '64', '"d"', '"engine.pace"', '1', '"membrane.V"']
    # A 1d file read as a 2d block of height 1
    d = myokit.DataBlock2d.load(p)
    assert d.shape() == (500, 1, 64)
    np.testing.assert_array_equal(d.get2d('membrane.V')[:, 0, :],
                                  b.get1d('membrane.V'))


def test_block2d_log_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    t = np.arange(6.0)
    V = rng.normal(size=(6, 3, 4))
    d = myokit.DataLog()
    d['engine.time'] = t
    d.set_time_key('engine.time')
    d['engine.pace'] = np.zeros(6)
    for y in range(3):
        for x in range(4):
            d['%d.%d.membrane.V' % (x, y)] = V[:, y, x]
    b = myokit.DataBlock2d.from_log(d)
    assert b.shape() == (6, 3, 4)
    np.testing.assert_array_equal(b.get2d('membrane.V'), V)
    np.testing.assert_array_equal(b.trace('membrane.V', 2, 1), V[:, 1, 2])
    p = str(tmp_path / 'b2.zip')
    b.save(p)
    c = myokit.DataBlock2d.load(p)
    np.testing.assert_array_equal(c.get2d('membrane.V'), V)
    e = c.to_log()
    np.testing.assert_array_equal(e['3.2.membrane.V'], V[:, 2, 3])


def test_block1d_from_log():
    d = myokit.DataLog()
    d['engine.time'] = np.arange(4.0)
    d.set_time_key('engine.time')
    for x in range(3):
        d['%d.membrane.V' % x] = np.arange(4.0) + 10 * x
    b = myokit.DataBlock1d.from_log(d)
    assert b.shape() == (4, 3)
    np.testing.assert_array_equal(b.trace('membrane.V', 2), np.arange(4.0)
                                  + 20)
    d['0.0.other'] = np.zeros(4)
    with pytest.raises(ValueError):
        myokit.DataBlock1d.from_log(d)


@pytest.mark.parametrize('precision', [myokit.DOUBLE_PRECISION,
                                       myokit.SINGLE_PRECISION])
def test_datalog_save_load_round_trip(tmp_path, precision):
    d = myokit.DataLog()
    d['engine.time'] = np.linspace(0, 1, 11)
    d.set_time_key('engine.time')
    d['0.membrane.V'] = np.linspace(-80, 20, 11)
    d['1.membrane.V'] = np.linspace(-70, 10, 11)
    p = str(tmp_path / 'log.zip')
    d.save(p, precision)
    e = myokit.DataLog.load(p)
    assert list(e.keys()) == list(d.keys())
    assert e.time_key() == 'engine.time'
    for k in d:
        if precision == myokit.DOUBLE_PRECISION:
            np.testing.assert_array_equal(e[k], d[k])
        else:
            np.testing.assert_array_equal(e[k], np.float32(d[k]))
    # Format (DataLog.save, myokit/_datalog.py:1051-1126)
    with zipfile.ZipFile(p) as z:
        head = z.read('structure.txt').decode().splitlines()
        body = z.read('data.bin')
    code = 'd' if precision == myokit.DOUBLE_PRECISION else 'f'
    assert head == ['3', '11', code, 'engine.time', 'engine.time',
                    '0.membrane.V', '1.membrane.V']
    dt = '<f8' if code == 'd' else '<f4'
    np.testing.assert_array_equal(
        np.frombuffer(body, dtype=dt)[11:22], np.asarray(
            d['0.membrane.V'], dtype=dt))


def test_datalog_load_errors(tmp_path):
    bad = str(tmp_path / 'bad.zip')
    with open(bad, 'wb') as f:
        f.write(b'not a zip')
    with pytest.raises(myokit.DataLogReadError):
        myokit.DataLog.load(bad)
    p = str(tmp_path / 'wrong.zip')
    with zipfile.ZipFile(p, 'w') as z:
        z.writestr('structure.txt', '2\n3\nd\n\na')
        z.writestr('data.bin', b'\0' * 48)
    with pytest.raises(myokit.DataLogReadError, match='number of fields'):
        myokit.DataLog.load(p)
    with zipfile.ZipFile(p, 'w') as z:
        z.writestr('structure.txt', '1\n30\nd\n\na')
        z.writestr('data.bin', b'\0' * 48)
    with pytest.raises(myokit.DataLogReadError, match='larger data size'):
        myokit.DataLog.load(p)


def test_has_nan_checks_final_value():
    # DataLog.has_nan (myokit/_datalog.py:366-374): the final value only
    d = myokit.DataLog()
    d['a'] = np.array([1.0, np.nan, 2.0])
    assert not d.has_nan()
    d['b'] = np.array([1.0, 2.0, np.nan])
    assert d.has_nan()
