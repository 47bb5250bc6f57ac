"""
Builds ``libmyokit_amd.so`` (the host library, C ABI in
``include/myokit_amd.h``) in-tree with hipcc.

The per-model kernels are not built here: they are generated at run time and
compiled with hiprtc for the device's architecture (gfx950 on MI355X).
``check_kernel_compiles`` compiles one generated kernel for gfx950 without a
device, as a build-time check of the generator and the driver templates.

Usage: ``python -m myokit_amd._build``
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, 'csrc', 'mkh_host.cpp')
OUT = os.path.join(HERE, 'libmyokit_amd.so')

HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')

FLAGS = [
    '-O2', '-fPIC', '-shared', '-std=c++17', '-Wall',
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '--offload-arch=gfx950',
]


def build_library(verbose=False):
    """Compiles the host library; returns the path of the .so."""
    cmd = [HIPCC] + FLAGS + ['-I', os.path.join(ROOT, 'include'),
                             '-o', OUT + '.tmp', SRC, '-lhiprtc']
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('Building libmyokit_amd.so failed:\n' + r.stderr)
    if verbose and r.stderr.strip():
        print(r.stderr)
    os.replace(OUT + '.tmp', OUT)
    return OUT


def check_kernel_compiles(precision=32, coupling=1, adjoint=False, rl=True,
                          fields=(), adjoint_precision=None):
    """Generates the LR1991 kernel (with the adjoint kernels if ``adjoint``)
    and compiles it for gfx950 with hiprtc (no device needed)."""
    import ctypes
    from . import _native, _kernelgen
    from .simulation import _prepare_cell_model
    from ._parser import load_model
    model = load_model('example')
    cm = _prepare_cell_model(model, rl=rl, diffusion=True,
                             fields=list(fields), quantities=['membrane.V'])
    src = _kernelgen.generate_kernel(cm, precision, coupling,
                                     adjoint=adjoint,
                                     adjoint_precision=adjoint_precision)
    prog = ctypes.c_void_p()
    log = ctypes.create_string_buffer(1 << 16)
    _native.check(_native.lib().mkh_program_build(
        src.encode(), _kernelgen.compile_options(precision).encode(), None,
        -1, ctypes.byref(prog), log, 1 << 16), 'mkh_program_build')
    _native.lib().mkh_program_free(prog)
    return True


if __name__ == '__main__':
    print(build_library(verbose=True))
    if '--check' in sys.argv:
        for p in (32, 64):
            check_kernel_compiles(p)
        print('kernels compile for gfx950')
