"""
The HIP exporter's runnable cable on the GPU: the program that
``__graft_entry__.build()`` exported and built (``tests/helpers.py``,
``EXPORT_CFG``: LR1991, double precision, 40 cells, dt = 2^-7 ms, g = 9.5,
5 paced cells) must reproduce the oracle's cable (the reference's
SimulationOpenCL semantics) bit for bit, in every state of every cell at
every logged time.
"""
import io
import os
import subprocess

import numpy as np
import pytest

import myokit_amd as myokit
from helpers import EXPORT_CFG, EXPORT_DIR, load
from oracle.cgen import OracleSimulation

pytestmark = pytest.mark.gpu


def test_exported_cable_matches_oracle():
    exe = os.path.join(EXPORT_DIR, 'cable')
    if not os.path.isfile(exe):
        raise RuntimeError('Exported cable not built: run build() first.')
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    head = lines[0].split(',')
    data = np.loadtxt(io.StringIO('\n'.join(lines[1:])), delimiter=',',
                      ndmin=2)
    n = EXPORT_CFG['n_cells']
    m = load('lr1991')
    states = [s.qname() for s in m.states()]
    assert head[:2] == ['engine.time', 'engine.pace']
    assert head[2:2 + len(states)] == ['0.' + s for s in states]
    o = OracleSimulation(m, myokit.pacing.blocktrain(1000, 2), ncells=n)
    o.gx = EXPORT_CFG['g']
    o.dt = EXPORT_CFG['dt']
    o.set_paced_cells(EXPORT_CFG['n_paced'])
    ref = o.run(EXPORT_CFG['duration'], log_vars=states,
                log_interval=EXPORT_CFG['log_interval'])
    np.testing.assert_array_equal(data[:, 0], ref['time'])
    np.testing.assert_array_equal(data[:, 1], ref['pace'])
    worst = 0.0
    for k, s in enumerate(states):
        got = data[:, 2 + k::len(states)]
        worst = max(worst, float(np.max(np.abs(got - ref[s]))))
    print('exported cable: %d rows x %d cells, max |d| over all states %g'
          % (len(data), n, worst), flush=True)
    assert worst == 0
    assert np.max(data[:, 2::len(states)][-1]) > 0     # the wave arrived
