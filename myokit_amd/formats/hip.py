"""
HIP exporters: a model's single-cell step as standalone HIP source for
gfx950, and a runnable cable simulation around it.

The counterparts in the reference are ``CudaKernelExporter`` /
``CudaKernelRLExporter`` (``myokit/formats/cuda/_exporter.py:12-65``, template
``cuda/template/kernel.cu``: one forward-Euler or Rush-Larsen step of one
cell, its default state) and ``OpenCLExporter`` / ``OpenCLRLExporter``
(``myokit/formats/opencl/_exporter.py:12-121``: a kernel plus a runnable
cable program). The cell step here is the one the simulation engine runs
(``generate_cell_function``: constants folded on the host, exact literals,
the portable double-precision functions of ``mk_math.h``), so an exported
double-precision kernel computes what ``SimulationOpenCL`` computes, bit
for bit.
"""
import os

from .._datalog import DOUBLE_PRECISION, SINGLE_PRECISION, float_str
from .._kernelgen import (_FLOAT_HELPERS, _read, generate_cell_function)


class HipKernelExporter:
    """Writes ``kernel.hip``: ``mk_export_step(state, time, dt, pace,
    idiff)``, one forward-Euler step of one cell in place, and
    ``mk_export_default_state(state)``.

    A variable must be bound to ``diffusion_current`` (the current from the
    cell to its neighbours: positive when the cell is a source) and the
    membrane potential labelled ``membrane_potential``.
    ``precision``: single (default, as the reference's GPU exporters) or
    double; ``native_maths``: the device's fast single-precision functions.
    """
    _rl = False
    _name = 'hip-kernel'

    def __init__(self, precision=SINGLE_PRECISION, native_maths=False):
        if precision not in (SINGLE_PRECISION, DOUBLE_PRECISION):
            raise ValueError('Only single and double precision are supported.')
        self._precision = precision
        self._native = bool(native_maths)

    def supports_model(self):
        return True

    def post_export_info(self):
        return ('Compile for an MI355X with: hipcc --offload-arch=gfx950 -O3'
                ' -ffp-contract=' + ('off' if self._precision ==
                                     DOUBLE_PRECISION else 'on')
                + ' -Xclang -target-feature -Xclang -packed-fp32-ops')

    def runnable(self, path, model, protocol=None):
        """Writes the export into the directory ``path`` (created if
        needed) and returns the paths of the files written."""
        from ..simulation import _prepare_cell_model
        path = os.path.abspath(os.path.expanduser(path))
        if os.path.exists(path) and not os.path.isdir(path):
            raise ValueError('Not a directory: ' + path)
        os.makedirs(path, exist_ok=True)
        model = model.clone()
        model.validate()
        if model.binding('diffusion_current') is None:
            raise ValueError('No variable bound to `diffusion_current`.')
        if model.label('membrane_potential') is None:
            raise ValueError('No variable labelled `membrane_potential`.')
        cm = _prepare_cell_model(model, self._rl, True, [], [])
        out = []
        for name, text in self._files(cm, protocol):
            fn = os.path.join(path, name)
            with open(fn, 'w') as f:
                f.write(text)
            out.append(fn)
        return out

    def _files(self, cm, protocol):
        return [('kernel.hip', self._kernel(cm))]

    def _kernel(self, cm):
        dbl = self._precision == DOUBLE_PRECISION
        states = list(cm.states)
        vm = cm.model.label('membrane_potential')
        init = [s.initial_value(as_float=True) for s in states]
        kind = self._name
        lines = [
            '/*',
            ' * kernel.hip -- model "' + str(cm.model.name()) + '", exported'
            ' by myokit_amd (' + kind + ').',
            ' *',
            ' * mk_export_step(state, time, dt, pace, idiff) advances one cell'
            ' by one',
            ' * step in place: forward Euler'
            + (', Rush-Larsen for the gating variables' if self._rl else '')
            + '. idiff is the',
            ' * diffusion current (outward: positive when the cell is a'
            ' source).',
            ' * mk_export_default_state(state) writes the initial state.',
            ' * ' + ('Double' if dbl else 'Single') + ' precision'
            + (', native maths' if self._native and not dbl else '') + '.',
            ' *',
            ' * State order (MK_EXPORT_NSTATE values):',
        ]
        for i, s in enumerate(states):
            lines.append(' *   %d  %s' % (i, s.qname()))
        lines += [
            ' */',
            '#pragma once',
            '#ifndef __HIPCC_RTC__',
            '#include <hip/hip_runtime.h>',
            '#endif',
            '#define MK_EXPORT_NSTATE ' + str(len(states)),
            '#define MK_EXPORT_IVM ' + str(states.index(vm)),
            '#define MK_MATH_DEVICE',
            'typedef ' + ('double' if dbl else 'float') + ' Real;',
            '',
        ]
        parts = ['\n'.join(lines), _read('mk_math.h')]
        if not dbl:
            parts.append(_FLOAT_HELPERS)
        parts.append(generate_cell_function(cm, self._precision,
                                            self._native))
        tail = [
            '/* One step of one cell, in place */',
            '__device__ __forceinline__ void mk_export_step(',
            '    Real* state, const Real time, const Real dt, const Real pace,',
            '    const Real idiff)',
            '{',
            '    mk_cell<false>(state, time, dt, pace, idiff, nullptr,'
            ' nullptr);',
            '}',
            '',
            '/* The model\'s initial state */',
            '__host__ __device__ inline void mk_export_default_state('
            'Real* state)',
            '{',
        ]
        for i, (s, v) in enumerate(zip(states, init)):
            tail.append('    state[%d] = %s;  /* %s */'
                        % (i, float_str(v), s.qname()))
        tail += ['}', '']
        parts.append('\n'.join(tail))
        return '\n'.join(parts)


class HipKernelRLExporter(HipKernelExporter):
    """Like :class:`HipKernelExporter`, with Rush-Larsen updates for the
    states in Hodgkin-Huxley form."""
    _rl = True
    _name = 'hip-kernel-rl'


_CABLE = r'''/*
 * cable.hip -- a cable of @N@ cells of the exported model on the GPU.
 *
 * Every step, from the potentials at t: the zero-flux stencil with
 * conductance g, then one step of every cell (kernel.hip), the first
 * @NPACED@ cells receiving the protocol's pacing level. Step k ends at
 * tmin + k dt and is taken with the difference to the current time, as in
 * myokit's simulations (sim_step, openclsim.c:1051-1162); the protocol's
 * events do not shorten steps here. Every @LOGEVERY@ steps the time, the
 * pacing level and the
 * state of every cell (at t, before the step) are printed as one CSV row
 * with 17 significant digits.
 *
 * Build: ./build.sh    Run: ./cable [duration]
 */
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "kernel.hip"

#define CHECK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { \
    fprintf(stderr, "HIP error %s at line %d\n", hipGetErrorString(e_), \
            __LINE__); exit(1); } } while (0)

/* Protocol events: level, start, duration, period, multiplier */
static const double events[][5] = {
@EVENTS@
};
static const int n_events = @NEVENTS@;

/* State names, in the order of MK_EXPORT_NSTATE */
static const char* const state_names[] = {@NAMES@};

/* Pacing level at time t: the level of the event (occurrence) that covers
 * t, the one with the latest start if several do. */
static double pace_at(double t)
{
    double level = 0, best = -HUGE_VAL;
    for (int k = 0; k < n_events; ++k) {
        const double* e = events[k];
        double s = e[1];
        if (e[3] > 0 && t >= s) {
            double r = std::floor((t - s) / e[3]);
            if (e[4] > 0 && r > e[4] - 1) r = e[4] - 1;
            s += r * e[3];
        }
        if (t >= s && t < s + e[2] && s > best) {
            best = s;
            level = e[0];
        }
    }
    return level;
}

/* One step of the cable: state src (SoA, state i of cell c at i * n + c)
 * to dst */
__global__ void __launch_bounds__(256)
cable_step(const int n, const int n_paced, const Real t, const Real dt,
           const Real pace, const Real g, const Real* __restrict__ src,
           Real* __restrict__ dst)
{
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= n) return;
    Real y[MK_EXPORT_NSTATE];
    for (int i = 0; i < MK_EXPORT_NSTATE; ++i) y[i] = src[i * n + c];
    const Real* V = src + MK_EXPORT_IVM * n;
    Real idiff = 0;
    if (n > 1) {
        if (c == 0) idiff = g * (V[0] - V[1]);
        else if (c == n - 1) idiff = g * (V[c] - V[c - 1]);
        else idiff = g * (((Real)2 * V[c] - V[c - 1]) - V[c + 1]);
    }
    mk_export_step(y, t, dt, c < n_paced ? pace : (Real)0, idiff);
    for (int i = 0; i < MK_EXPORT_NSTATE; ++i) dst[i * n + c] = y[i];
}

int main(int argc, char** argv)
{
    const int n = @N@, n_paced = @NPACED@, log_every = @LOGEVERY@;
    const double tmin = 0, dt = @DT@, g = @G@;
    const double tmax = argc > 1 ? atof(argv[1]) : @DURATION@;
    const int ns = MK_EXPORT_NSTATE;
    std::vector<Real> h((size_t)ns * n);
    Real s0[MK_EXPORT_NSTATE];
    mk_export_default_state(s0);
    for (int c = 0; c < n; ++c)
        for (int i = 0; i < ns; ++i) h[(size_t)i * n + c] = s0[i];
    Real *a = nullptr, *b = nullptr;
    CHECK(hipMalloc((void**)&a, h.size() * sizeof(Real)));
    CHECK(hipMalloc((void**)&b, h.size() * sizeof(Real)));
    CHECK(hipMemcpy(a, h.data(), h.size() * sizeof(Real),
                    hipMemcpyHostToDevice));
    printf("engine.time,engine.pace");
    for (int c = 0; c < n; ++c)
        for (int i = 0; i < ns; ++i) printf(",%d.%s", c, state_names[i]);
    printf("\n");
    const int grid = (n + 255) / 256;
    double t = tmin;
    for (long k = 0; t < tmax; ++k) {
        const double hk = tmin + (double)(k + 1) * dt - t;
        const double pace = pace_at(t);
        if (k % log_every == 0) {
            CHECK(hipMemcpy(h.data(), a, h.size() * sizeof(Real),
                            hipMemcpyDeviceToHost));
            printf("%.17g,%.17g", t, pace);
            for (int c = 0; c < n; ++c)
                for (int i = 0; i < ns; ++i)
                    printf(",%.17g", (double)h[(size_t)i * n + c]);
            printf("\n");
        }
        cable_step<<<grid, 256>>>(n, n_paced, (Real)t, (Real)hk, (Real)pace,
                                  (Real)g, a, b);
        CHECK(hipGetLastError());
        Real* tmp = a;
        a = b;
        b = tmp;
        t += hk;
    }
    CHECK(hipDeviceSynchronize());
    CHECK(hipFree(a));
    CHECK(hipFree(b));
    return 0;
}
'''

_BUILD = '''#!/bin/bash
# Builds the cable simulation for an MI355X (gfx950).
set -e
cd "$(dirname "$0")"
hipcc --offload-arch=gfx950 -O3 -ffp-contract=@CONTRACT@ \\
    -Xclang -target-feature -Xclang -packed-fp32-ops -o cable cable.hip
'''


class HipCableExporter(HipKernelExporter):
    """Writes ``kernel.hip``, ``cable.hip`` (a runnable cable simulation:
    ``n_cells`` cells, conductance ``g``, step ``dt``, the first ``n_paced``
    cells paced by the protocol, every state logged every ``log_interval``
    time units as CSV on stdout) and ``build.sh``. ``duration`` is the
    default run length (``./cable [duration]``)."""
    _name = 'hip'

    def __init__(self, precision=SINGLE_PRECISION, native_maths=False,
                 n_cells=256, dt=0.005, g=1.5, n_paced=5, duration=600.0,
                 log_interval=1.0):
        super().__init__(precision, native_maths)
        self._n = int(n_cells)
        self._dt = float(dt)
        self._g = float(g)
        self._n_paced = int(n_paced)
        self._duration = float(duration)
        self._log_every = max(1, int(round(float(log_interval) / self._dt)))
        if self._n < 1:
            raise ValueError('The cable needs at least one cell.')
        if self._dt <= 0:
            raise ValueError('The step size must be greater than zero.')

    def post_export_info(self):
        return 'Build with ./build.sh, run with ./cable [duration] > out.csv'

    def _files(self, cm, protocol):
        if protocol is None:
            from ..pacing import blocktrain
            protocol = blocktrain(1000, 0.5, offset=50)
        ev = protocol.to_arrays()
        rows = ['    {%s},' % ', '.join(float_str(x) for x in e)
                for e in ev] or ['    {0, 0, 0, 0, 0},']
        names = ', '.join('"' + s.qname() + '"' for s in cm.states)
        cable = _CABLE
        for key, val in (
                ('@N@', str(self._n)), ('@NPACED@', str(self._n_paced)),
                ('@LOGEVERY@', str(self._log_every)),
                ('@DT@', float_str(self._dt)), ('@G@', float_str(self._g)),
                ('@DURATION@', float_str(self._duration)),
                ('@EVENTS@', '\n'.join(rows)),
                ('@NEVENTS@', str(len(ev))), ('@NAMES@', names)):
            cable = cable.replace(key, val)
        build = _BUILD.replace('@CONTRACT@', 'off' if self._precision ==
                               DOUBLE_PRECISION else 'on')
        return [('kernel.hip', self._kernel(cm)), ('cable.hip', cable),
                ('build.sh', build)]


class HipCableRLExporter(HipCableExporter):
    """Like :class:`HipCableExporter`, with Rush-Larsen updates."""
    _rl = True
    _name = 'hip-rl'
