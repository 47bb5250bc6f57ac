"""
Autotuning on the GPU: every candidate either runs or is skipped, the
simulation's state and time are left unchanged, and the tuned engine gives
the results of the default one.
"""
import numpy as np
import pytest

import myokit_amd as myokit
from myokit_amd import tune

pytestmark = pytest.mark.gpu


def test_autotune_sheet(tmp_path, monkeypatch):
    monkeypatch.setenv(tune.CACHE_ENV, str(tmp_path / 'tune.json'))
    m = myokit.load_model('example')

    def make():
        s = myokit.SimulationOpenCL(m, myokit.pacing.blocktrain(1000, 2),
                                    ncells=(128, 96), rl=True, precision=64)
        s.set_step_size(0.01)
        s.set_conductance(9.5, 9.5)
        s.set_paced_cells(4, 96, 0, 0)
        return s
    s = make()
    s.run(1.0)
    st, t = s.state(), s.time()
    r = tune.autotune(s, duration=0.2, reps=1, verbose=True)
    ran = [c for c, x in r['trials'] if x is not None]
    print('chosen', r['config'], 'of', len(ran), 'runnable', flush=True)
    assert r['config'] in ran and len(ran) >= 3
    np.testing.assert_array_equal(s.state(), st)
    assert s.time() == t
    # double precision: every driver gives the same numbers
    d1 = s.run_block(3.0, ['membrane.V'], log_interval=0.5)
    ref = make()
    ref.run(1.0)
    d2 = ref.run_block(3.0, ['membrane.V'], log_interval=0.5)
    np.testing.assert_array_equal(d1['membrane.V'], d2['membrane.V'])
    # the cache gives the same choice to a new simulation of this kind
    u = make()
    assert tune.apply_cached(u) == r['config']
