#!/usr/bin/env python3
"""
Writes the golden fixtures in tests/golden/ by running the reference myokit
(from /root/reference) in the build container. The reference never travels
with the repository; only the small data files written here do.

Each fixture pins something no closed form or reference test vector pins:

* derivatives.npz   -- Model.evaluate_derivatives of LR1991 (example.mmt) and
                       Beeler-Reuter 1977 at random states and inputs: pins the
                       mmt parser and the expression evaluator.
* rl_forms.json     -- inf/tau variables that lib.hh finds after
                       convert_hh_states_to_inf_tau_form: pins RL detection.
* sim1d_*.npz       -- Simulation1d (CPU, double precision) cable runs. The
                       reference's own test pins SimulationOpenCL to
                       Simulation1d within 1e-13 in double precision
                       (test_simulation_opencl_vs_sim1d.py:118-136), so these
                       trajectories pin the 1D FE/RL stepping, coupling and
                       pacing semantics. Besides V they hold the logged
                       diffusion current (membrane.i_diff), an intermediary
                       (ica.ICa or isi.Isi) and engine.pace, which pin what
                       a log row holds: values at the row's time, before the
                       step from it. sim1d_br1977_reftest.npz is the
                       configuration of that reference test itself.
* pacing_py.json    -- myokit.PacingSystem sequences (time, next time, level)
                       for random event protocols: pins the event system.

Usage: python tests/golden/make_golden.py
"""
import json
import os
import sys
import warnings

import numpy as np

REF = '/root/reference'
HERE = os.path.dirname(os.path.abspath(__file__))
DATA = os.path.join(os.path.dirname(HERE), 'data')

sys.path.insert(0, REF)
warnings.simplefilter('ignore')
import myokit  # noqa: E402
import myokit.lib.hh as hh  # noqa: E402

MODELS = {
    'lr1991': os.path.join(REF, 'myokit', '_bin', 'example.mmt'),
    'br1977': os.path.join(DATA, 'beeler-1977-model.mmt'),
}


def derivatives():
    rng = np.random.default_rng(1)
    out = {}
    for name, path in MODELS.items():
        m = myokit.load_model(path)
        x0 = np.array(m.initial_values(True))
        states, inputs, derivs = [], [], []
        for k in range(40):
            x = x0 * (1 + 0.2 * rng.uniform(-1, 1, size=len(x0)))
            x[0] = rng.uniform(-90, 40)          # membrane potential
            pace = float(rng.integers(0, 2))
            idiff = float(rng.uniform(-5, 5))
            inp = {'pace': pace, 'diffusion_current': idiff, 'time': 0.0}
            d = m.evaluate_derivatives(list(x), inp)
            states.append(x)
            inputs.append([pace, idiff])
            derivs.append(d)
        out[name + '_states'] = np.array(states)
        out[name + '_inputs'] = np.array(inputs)
        out[name + '_derivs'] = np.array(derivs)
    np.savez(os.path.join(HERE, 'derivatives.npz'), **out)


def rl_forms():
    out = {}
    for name, path in MODELS.items():
        m = myokit.load_model(path)
        vm = m.label('membrane_potential')
        c = hh.convert_hh_states_to_inf_tau_form(m, vm)
        v = c.get(vm.qname())
        r = {}
        for s in c.states():
            res = hh.get_inf_and_tau(s, v)
            r[s.qname()] = None if res is None else [
                res[0].qname(), res[1].qname()]
        ab = {}
        for s in m.states():
            res = hh.get_alpha_and_beta(s, vm)
            ab[s.qname()] = None if res is None else [
                res[0].qname(), res[1].qname()]
        out[name] = {'inf_tau': r, 'alpha_beta': ab}
    with open(os.path.join(HERE, 'rl_forms.json'), 'w') as f:
        json.dump(out, f, indent=1, sort_keys=True)


def sim1d():
    cases = [
        # name, model, ncells, npaced, g, rl, dt, duration, log_interval,
        # protocol (level, start, duration, period, multiplier),
        # intermediary logged besides V and i_diff, stimulus amplitude
        # (None = the model's)
        ('sim1d_lr1991_fe', 'lr1991', 20, 5, 9.5, False, 0.005, 30.0, 0.5,
         (1, 0, 2, 1000, 0), 'ica.ICa', None),
        ('sim1d_lr1991_rl', 'lr1991', 20, 5, 9.5, True, 0.01, 30.0, 0.5,
         (1, 0, 2, 1000, 0), 'ica.ICa', None),
        ('sim1d_br1977_fe', 'br1977', 10, 2, 10.0, False, 0.005, 15.0, 0.5,
         (1, 1, 2, 1000, 0), 'isi.Isi', None),
        ('sim1d_lr1991_cfg1', 'lr1991', 200, 5, 9.5, False, 0.005, 60.0, 1.0,
         (1, 0, 2, 1000, 0), 'ica.ICa', None),
        # test_simulation_opencl_vs_sim1d.py:28-57: BR1977 with the
        # stimulus amplitude set to -80, blocktrain(1000, 2, offset=3),
        # 10 cells, 2 paced, dt 0.005, 15 ms logged every 0.5 ms; Simulation1d
        # uses its default conductance
        ('sim1d_br1977_reftest', 'br1977', 10, 2, None, False, 0.005, 15.0,
         0.5, (1, 3, 2, 1000, 0), 'isi.Isi', -80.0),
    ]
    for (name, model, n, npaced, g, rl, dt, dur, li, ev, inter,
         amp) in cases:
        m = myokit.load_model(MODELS[model])
        if amp is not None:
            m.get('membrane.i_stim.amplitude').set_rhs(amp)
        p = myokit.Protocol()
        p.schedule(*ev)
        s = myokit.Simulation1d(m, p, ncells=n, rl=rl)
        if g is None:
            g = s.conductance()
        s.set_conductance(g)
        s.set_paced_cells(npaced)
        s.set_step_size(dt)
        vname = m.label('membrane_potential').qname()
        dname = m.binding('diffusion_current').qname()
        qs = [vname, dname, inter]
        keys = ['engine.time', 'engine.pace'] + [
            str(i) + '.' + q for q in qs for i in range(n)]
        d = s.run(dur, log=keys, log_interval=li).npview()
        cols = {}
        for key, q in zip(['V', 'i_diff', 'inter'], qs):
            cols[key] = np.stack([d[str(i) + '.' + q] for i in range(n)],
                                 axis=1)
        np.savez(os.path.join(HERE, name + '.npz'),
                 time=d['engine.time'], pace=d['engine.pace'],
                 final_state=np.array(s.state()),
                 config=np.array([n, npaced, g, float(rl), dt, dur, li]),
                 event=np.array(ev, dtype=float),
                 inter_name=np.array(inter),
                 amplitude=np.array(np.nan if amp is None else amp),
                 **cols)
        print(name, cols['V'].shape)


def pacing():
    rng = np.random.default_rng(3)
    cases = []
    for k in range(30):
        p = myokit.Protocol()
        n = int(rng.integers(1, 4))
        t0 = 0.0
        rows = []
        for i in range(n):
            start = t0 + float(rng.integers(0, 20)) * 0.5
            dur = float(rng.integers(1, 8)) * 0.25
            if rng.uniform() < 0.5:
                period = dur + float(rng.integers(1, 10)) * 0.5
                mult = int(rng.integers(0, 4))
            else:
                period, mult = 0.0, 0
            level = float(rng.integers(1, 5)) * 0.5
            try:
                p.schedule(level, start, dur, period, mult)
            except myokit.SimultaneousProtocolEventError:
                continue
            rows.append([level, start, dur, period, mult])
            t0 = start + dur + 1
        times = np.sort(rng.uniform(0, 60, size=40))
        times = np.concatenate([[0.0], times])
        ps = myokit.PacingSystem(p)
        seq = []
        try:
            for t in times:
                lev = ps.advance(float(t))
                seq.append([float(t), ps.next_time(), lev])
            err = None
        except myokit.SimultaneousProtocolEventError:
            err = 'simultaneous'
        cases.append({'events': [[e.level(), e.start(), e.duration(),
                                  e.period(), e.multiplier()]
                                 for e in p.events()],
                      'times': [float(t) for t in times],
                      'sequence': seq, 'error': err})
    with open(os.path.join(HERE, 'pacing_py.json'), 'w') as f:
        json.dump(cases, f)


if __name__ == '__main__':
    derivatives()
    rl_forms()
    pacing()
    sim1d()
