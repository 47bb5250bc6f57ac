"""
The portable double-precision functions (myokit_amd/csrc/mk_math.h) on the
device against the same functions in the CPU oracle, value for value, over
inputs no cell model reaches: exp at and beyond both range limits (results
that overflow, that round into the subnormals or to zero), at +-0, +-inf
and NaN; log, log10, sqrt and pow over the whole positive range
(subnormals, zero, inf) and at negative arguments; the trigonometric
functions up to |x| = 1e4 and near multiples of pi/2; asin and acos over
[-1, 1]. Each input is a per-cell field of a probe model whose logged
intermediaries apply the functions, so the device evaluates them in the
generated kernel exactly as a model's equations do. Each argument carries
a term 0 * v (v = -80, so the term is -0: x + (-0) is x for every x, +-0
included), which makes the results intermediaries rather than constants,
and constants cannot be logged.

NaN results are compared as NaN (a CPU and a GPU may set different NaN
bits); every other result must be equal, sign of zero included.
"""
import numpy as np
import pytest

import myokit_amd as myokit

pytestmark = pytest.mark.gpu

PROBE = '''
[[model]]
c.v = -80

[engine]
time = 0 bind time
pace = 0 bind pace

[c]
dot(v) = -0.001 * (v + 80) + engine.pace * 0
    label membrane_potential
xe = 0
xt = 0
xb = 1
xs = 0
z = 0 * v
e = exp(xe + z)
l = log(xb + z)
l10 = log10(xb + z)
sq = sqrt(xb + z)
pw = (xb + z) ^ (xe / 256)
s = sin(xt + z)
co = cos(xt + z)
t = tan(xt + z)
at = atan(xt + z)
as = asin(xs + z)
ac = acos(xs + z)
'''

OUT = ['c.e', 'c.l', 'c.l10', 'c.sq', 'c.pw', 'c.s', 'c.co', 'c.t', 'c.at',
       'c.as', 'c.ac']
N = 4096


def _inputs():
    rng = np.random.default_rng(2024)
    # exp: specials, the subnormal and overflow ranges, dense steps around
    # both limits, the clamp bounds, random arguments
    specials = [0.0, -0.0, np.inf, -np.inf, np.nan, 1.0, -1.0, 709.78,
                709.782712893384, 709.7827128933841, 710.0, 711.0, -746.0,
                -746.5, -745.1332191019411, -745.1332191019412, -708.4,
                -708.3964185322641, 5e-324, -5e-324, 1e300, -1e300]
    xe = list(specials)
    for lim in (709.782712893384, -745.1332191019411):
        x = lim
        for _ in range(200):
            x = np.nextafter(x, -np.inf)
        for _ in range(400):
            xe.append(x)
            x = np.nextafter(x, np.inf)
    xe += list(np.linspace(-746.5, -707, 800))
    xe += list(np.linspace(705, 711, 300))
    xe += list(rng.uniform(-800, 800, N - len(xe) - 800))
    xe += list(rng.standard_normal(800) * 30)
    xe = np.array(xe[:N])
    # log, log10, sqrt, pow: 10^[-320, 308], zero, subnormals, inf, negatives
    xb = 10.0 ** rng.uniform(-320, 308, N)
    xb[:12] = [0.0, -0.0, 5e-324, 2.2250738585072014e-308, 1.0, np.inf,
               -1.0, -1e-300, np.nan, 0.5, 2.0, 1e308]
    # trigonometric: |x| up to 1e4, near multiples of pi/2, +-0
    xt = rng.uniform(-1e4, 1e4, N)
    k = np.arange(-200, 200)
    xt[:400] = k * (np.pi / 2)
    xt[400:800] = k * (np.pi / 2) + rng.uniform(-1e-9, 1e-9, 400)
    xt[800:1600] = rng.uniform(-4, 4, 800)
    xt[1600:1604] = [0.0, -0.0, 1e-300, -1e-300]
    # asin, acos: [-1, 1] and just outside
    xs = rng.uniform(-1, 1, N)
    xs[:8] = [-1.0, 1.0, 0.0, -0.0, np.nextafter(1, 2), -np.nextafter(1, 2),
              0.5, -0.5]
    return {'c.xe': xe, 'c.xt': xt, 'c.xb': xb, 'c.xs': xs}


def test_portable_math_device_equals_oracle():
    from oracle.cgen import OracleSimulation
    m = myokit.parse_model(PROBE)
    fields = _inputs()
    p = myokit.Protocol()
    s = myokit.SimulationOpenCL(m, p, ncells=N, diffusion=False,
                                precision=myokit.DOUBLE_PRECISION)
    s.set_step_size(0.01)
    for name, v in fields.items():
        s.set_field(name, v)
    # NaN results are the point here, not a numerical error to report
    d = s.run_block(0.02, OUT, log_interval=0.01, report_nan=False)
    assert s.engine_driver() == 'cells'
    o = OracleSimulation(myokit.parse_model(PROBE), myokit.Protocol(),
                         ncells=N, diffusion=False)
    o.dt = 0.01
    for name, v in fields.items():
        o.fields[name] = np.asarray(v, dtype=float)
    r = o.run(0.02, log_vars=list(OUT), log_interval=0.01)
    rows = len(r['time'])
    assert rows >= 1
    np.testing.assert_array_equal(d['time'], r['time'])
    for var in OUT:
        got = np.asarray(d[var], dtype=float)
        assert got.shape == (rows, N)
        ref = np.asarray(r[var], dtype=float).reshape(rows, N)
        nan = np.isnan(ref)
        bad = ~((got == ref) & (np.signbit(got) == np.signbit(ref))) & ~nan
        bad |= nan != np.isnan(got)
        print(var, 'rows', rows, 'NaN', int(nan[0].sum()), 'inf',
              int(np.isinf(ref[0]).sum()), 'zero', int((ref[0] == 0).sum()),
              'differ', int(bad.sum()), flush=True)
        if bad.any():
            i = np.argwhere(bad)[0]
            raise AssertionError('%s differs at cell %d: device %r, oracle '
                                 '%r' % (var, i[1], got[tuple(i)],
                                         ref[tuple(i)]))


def _ulps(got, ref):
    """|got - ref| in units of the single-precision spacing at ref."""
    ref32 = ref.astype(np.float32)
    sp = np.spacing(np.abs(ref32)).astype(np.float64)
    return np.abs(got - ref) / sp


def test_single_precision_math_against_numpy():
    # The single-precision kernel's functions (hardware v_exp_f32 with the
    # rounding-error correction of mk_expf, and the device library's log,
    # sqrt, pow, sin, cos, tan, atan, asin, acos) against numpy in double
    # precision on the same float inputs. Bars: OpenCL's single-precision
    # bounds, under which the reference's kernels run (exp, log, log10 and
    # sqrt 3 ulp; sin, cos, asin and acos 4; tan and atan 5; pow 16), where
    # the result is a normal float, sin, cos and tan also passing within
    # 2^-24 absolute near their zeros. Overflow to inf, underflow to 0 and
    # NaN must appear at the same inputs (away from the last ulps of the
    # range).
    m = myokit.parse_model(PROBE)
    f = _inputs()
    n = N
    rng = np.random.default_rng(99)
    xe = f['c.xe'].copy()
    fin = np.isfinite(xe)
    xe[fin] = np.clip(xe[fin], -120, 120)
    xe[:64] = np.linspace(-104, -85, 64)        # into the float subnormals
    xe[64:128] = np.linspace(85, 89, 64)        # past the float range
    xb = 10.0 ** rng.uniform(-37, 37, n)
    xb[:8] = [0.0, -1.0, np.inf, np.nan, 1.0, 2.0, 0.5, 1e-30]
    xt = f['c.xt']
    xs = f['c.xs']
    fields = {'c.xe': xe, 'c.xt': xt, 'c.xb': xb, 'c.xs': xs}
    s = myokit.SimulationOpenCL(m, myokit.Protocol(), ncells=n,
                                diffusion=False,
                                precision=myokit.SINGLE_PRECISION)
    s.set_step_size(0.01)
    for name, v in fields.items():
        s.set_field(name, v)
    d = s.run_block(0.01, OUT, log_interval=0.01, report_nan=False)
    x = {k: np.asarray(v, dtype=np.float32).astype(np.float64)
         for k, v in fields.items()}
    with np.errstate(all='ignore'):
        ref = {
            'c.e': np.exp(x['c.xe']), 'c.l': np.log(x['c.xb']),
            'c.l10': np.log10(x['c.xb']), 'c.sq': np.sqrt(x['c.xb']),
            'c.pw': np.power(x['c.xb'], (x['c.xe'] / 256).astype(
                np.float32).astype(np.float64)),
            'c.s': np.sin(x['c.xt']), 'c.co': np.cos(x['c.xt']),
            'c.t': np.tan(x['c.xt']), 'c.at': np.arctan(x['c.xt']),
            'c.as': np.arcsin(x['c.xs']), 'c.ac': np.arccos(x['c.xs'])}
    bars = {'c.e': 3, 'c.l': 3, 'c.l10': 3, 'c.sq': 3, 'c.pw': 16,
            'c.s': 4, 'c.co': 4, 'c.t': 5, 'c.at': 5, 'c.as': 4, 'c.ac': 4}
    tiny = 2.0 ** -126
    for var in OUT:
        got = np.asarray(d[var], dtype=np.float64)[0]
        r = ref[var]
        with np.errstate(over='ignore'):    # out-of-range results: inf
            r32 = r.astype(np.float32).astype(np.float64)
        # special values where the double result is clearly outside the
        # float range (not within a few ulp of its edges)
        big = np.abs(r) > 3.0e38
        small = (np.abs(r) < 1e-39) & np.isfinite(r)
        nan = np.isnan(r)
        assert np.array_equal(np.isnan(got), nan), var
        assert np.all(np.isinf(got[big & (np.abs(r) > 3.5e38)])), var
        if var in ('c.e', 'c.pw'):
            assert np.all(got[small & (np.abs(r) < 1e-46)] == 0), var
        ok = np.isfinite(r32) & ~nan & (np.abs(r32) >= tiny) & ~big
        u = _ulps(got[ok], r[ok])
        if var in ('c.s', 'c.co', 'c.t'):
            u = np.where(np.abs(got[ok] - r[ok]) <= 2.0 ** -24, 0, u)
        print(var, 'checked', int(ok.sum()), 'max ulp %.2f' % float(
            np.max(u) if len(u) else 0), flush=True)
        assert np.max(u) <= bars[var], (var, float(np.max(u)))
