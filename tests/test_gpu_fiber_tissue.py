"""
FiberTissueSimulation on the GPU: double precision against the fiber-tissue
oracle bit for bit, single precision within the stated tolerances, and the
behaviour the reference's tests check
(``myokit/tests/test_simulation_fiber_tissue.py``: logs, negative time,
pre/reset, protocol, run errors and progress). Models from the reference's
test data: ``dn-1985-normalised.mmt`` (fiber) and ``lr-1991.mmt``
(tissue).
"""
import os

import numpy as np
import pytest

import myokit_amd as myokit
from oracle.cgen import OracleFiberTissue

pytestmark = pytest.mark.gpu

DATA = os.path.join(os.path.dirname(__file__), 'data')


@pytest.fixture(scope='module')
def models():
    mf = myokit.load_model(os.path.join(DATA, 'dn-1985-normalised.mmt'))
    mt = myokit.load_model(os.path.join(DATA, 'lr-1991.mmt'))
    return mf, mt, myokit.pacing.blocktrain(1000, 2.0, offset=.01)


def small(models, precision=myokit.SINGLE_PRECISION):
    # the reference's shared small simulation (s1)
    mf, mt, p = models
    return myokit.FiberTissueSimulation(
        mf, mt, p, ncells_fiber=(8, 4), ncells_tissue=(8, 6), nx_paced=4,
        g_fiber=(235, 100), g_tissue=(9, 5), g_fiber_tissue=9, dt=0.0012,
        precision=precision)


def single(models):
    # the reference's shared 1 x 1 simulation (s0)
    _, mt, p = models
    return myokit.FiberTissueSimulation(
        mt, mt, p, ncells_fiber=(1, 1), ncells_tissue=(1, 1), nx_paced=1,
        g_fiber=(0, 0), g_tissue=(0, 0), g_fiber_tissue=0,
        precision=myokit.DOUBLE_PRECISION, dt=0.01)


def _v(log, nx, ny):
    return np.stack([np.asarray(log['%d.%d.membrane.V' % (x, y)])
                     for y in range(ny) for x in range(nx)], axis=1)


CASES = [
    # fiber model, tissue model, fiber, tissue, g_fiber, g_tissue, gft, dt,
    # duration
    ('dn', 'lr', (8, 4), (8, 6), (235, 100), (9, 5), 9, 0.0012, 5.0),
    ('lr', 'lr', (24, 3), (14, 9), (9, 6), (9, 5), 7, 0.01, 20.0),
]


@pytest.mark.parametrize('case', range(len(CASES)))
def test_fp64_matches_oracle_bit_exact(models, case):
    mf, mt, p = models
    names, nf, nt, gf, gt, gft, dt, dur = (CASES[case][:2],) + CASES[case][2:]
    fm = mf if names[0] == 'dn' else mt
    kw = dict(ncells_fiber=nf, ncells_tissue=nt, nx_paced=3, g_fiber=gf,
              g_tissue=gt, g_fiber_tissue=gft, dt=dt)
    s = myokit.FiberTissueSimulation(fm, mt, p,
                                     precision=myokit.DOUBLE_PRECISION, **kw)
    lf, lt = s.run(dur, logf=['engine.time', 'membrane.V'],
                   logt=['membrane.V'], log_interval=0.25)
    o = OracleFiberTissue(fm, mt, p, **kw)
    t, pace, vf, vt = o.run(dur, log_interval=0.25)
    np.testing.assert_array_equal(np.asarray(lf['engine.time']), t)
    df = np.max(np.abs(_v(lf, *nf) - vf))
    dtt = np.max(np.abs(_v(lt, *nt) - vt))
    print('fiber-tissue fp64 case %d: max|dV| fiber %g tissue %g mV (rows'
          ' %d, tissue V max %.1f)' % (case, df, dtt, len(t), np.max(vt)),
          flush=True)
    assert df == 0 and dtt == 0
    np.testing.assert_array_equal(np.array(s.fiber_state()), o.state_f)
    np.testing.assert_array_equal(np.array(s.tissue_state()), o.state_t)


def _activation(t, v, thr=-30.0):
    """First upward crossing of thr per cell, linearly interpolated."""
    out = np.full(v.shape[1], np.nan)
    for c in range(v.shape[1]):
        i = np.nonzero((v[:-1, c] < thr) & (v[1:, c] >= thr))[0]
        if len(i):
            k = i[0]
            out[c] = t[k] + (thr - v[k, c]) * (t[k + 1] - t[k]) / (
                v[k + 1, c] - v[k, c])
    return out


def test_fp32_accuracy(models):
    _, mt, p = models
    kw = dict(ncells_fiber=(24, 3), ncells_tissue=(14, 9), nx_paced=3,
              g_fiber=(9, 6), g_tissue=(9, 5), g_fiber_tissue=7, dt=0.01)
    s = myokit.FiberTissueSimulation(mt, mt, p, **kw)
    lf, lt = s.run(30.0, logf=['membrane.V'], logt=['membrane.V'],
                   log_interval=0.05)
    o = OracleFiberTissue(mt, mt, p, **kw)
    t, pace, vf, vt = o.run(30.0, log_interval=0.05)
    v32 = np.concatenate([_v(lf, 24, 3), _v(lt, 14, 9)], axis=1)
    v64 = np.concatenate([vf, vt], axis=1)
    rms = np.sqrt(np.mean((v32 - v64) ** 2))
    a32, a64 = _activation(t, v32), _activation(t, v64)
    assert np.array_equal(np.isnan(a32), np.isnan(a64))
    ok = ~np.isnan(a64)
    da = np.max(np.abs(a32[ok] - a64[ok]))
    print('fiber-tissue fp32: V RMS %.3g mV, activation max|dt| %.3g ms over'
          ' %d cells' % (rms, da, ok.sum()), flush=True)
    assert ok.sum() == v64.shape[1]     # every cell activated
    assert rms <= 0.5 and da <= 0.1


def test_basic(models):
    s = small(models)
    logf = ['engine.time', 'membrane.V', 'isi.isiCa']
    logt = ['membrane.V', 'ica.Ca_i', 'ica.ICa']
    logf, logt = s.run(.1, logf=logf, logt=logt, log_interval=0.01)
    nfx, nfy, ntx, nty = 8, 4, 8, 6
    assert len(logf) == 1 + 2 * nfx * nfy
    assert 'engine.time' in logf and '0.0.membrane.V' in logf
    assert '%d.%d.isi.isiCa' % (nfx - 1, nfy - 1) in logf
    assert len(logt) == 3 * ntx * nty
    assert '%d.%d.ica.ICa' % (ntx - 1, nty - 1) in logt
    assert len(logf['engine.time']) == 10
    assert all(len(v) == 10 for v in logt.values())


def test_negative_time(models):
    s = single(models)
    p = myokit.pacing.blocktrain(duration=1, level=1, period=2)
    s.set_protocol(p)
    s.set_time(-10)
    assert s.time() == -10
    df, dt = s.run(5, logf=['engine.pace'], logt=myokit.LOG_NONE)
    assert s.time() == -5
    assert np.all(np.asarray(df['engine.pace']) == 0)
    df, dt = s.run(6, logf=['engine.time', 'engine.pace'],
                   logt=myokit.LOG_NONE)
    pace = np.asarray(df['engine.pace'])
    assert s.time() == 1
    assert np.all(pace[:-1] == 0) and pace[-1] == 1


def test_pre_reset(models):
    s = small(models)
    sf, st = s.fiber_state(), s.tissue_state()
    assert sf == s.default_fiber_state() and st == s.default_tissue_state()
    s.run(5, logf=myokit.LOG_NONE, logt=myokit.LOG_NONE)
    assert s.time() == 5
    assert s.fiber_state() != s.default_fiber_state()
    assert s.tissue_state() != s.default_tissue_state()
    assert sf == s.default_fiber_state() and st == s.default_tissue_state()
    s.set_time(0)
    assert s.fiber_state() != s.default_fiber_state()
    s.reset()
    assert sf == s.fiber_state() and st == s.tissue_state()
    assert s.time() == 0
    s.pre(5)
    assert s.time() == 0
    assert sf != s.fiber_state() and st != s.tissue_state()
    sf, st = s.fiber_state(), s.tissue_state()
    assert sf == s.default_fiber_state() and st == s.default_tissue_state()
    s.run(5, logf=myokit.LOG_NONE, logt=myokit.LOG_NONE)
    assert s.fiber_state() != s.default_fiber_state()
    s.reset()
    assert s.time() == 0
    assert sf == s.fiber_state() and st == s.tissue_state()


def test_protocol(models):
    s = single(models)
    df, dt = s.run(5, logf=['membrane.V'], logt=myokit.LOG_NONE,
                   log_interval=1)
    assert df['0.0.membrane.V'][-1] > 0
    s.reset()
    x = 0.0123
    s.set_protocol(myokit.pacing.blocktrain(level=x, duration=1000,
                                            period=1000))
    df, dt = s.run(3, logf=['engine.pace'])
    assert np.all(np.asarray(df['engine.pace']) == x)
    s.set_protocol(None)
    s.reset()
    df, dt = s.run(3, logf=['engine.pace'])
    assert np.all(np.asarray(df['engine.pace']) == 0)


class _Cancel(myokit.ProgressReporter):
    def enter(self, msg=None):
        pass

    def exit(self):
        pass

    def update(self, progress):
        return False


def test_run_errors_and_progress(models):
    s = small(models)
    with pytest.raises(ValueError, match='can\'t be negative'):
        s.run(-1)
    s.run(1, log_interval=0)
    s.run(1, log_interval=-1)
    with pytest.raises(ValueError, match='subclass of myokit.ProgressR'):
        s.run(1, progress=12)
    t0, sf = s.time(), s.fiber_state()
    with pytest.raises(myokit.SimulationCancelledError):
        s.run(20, progress=_Cancel())
    assert s.time() == t0 and s.fiber_state() == sf
