"""
Exporters of a model to standalone HIP source (the role of
``myokit.formats.exporter``, ``myokit/formats/__init__.py:96-115``, for its
GPU kernel formats ``cuda-kernel``, ``cuda-kernel-rl``, ``opencl`` and
``opencl-rl``).

``exporter(name)`` returns an exporter; ``exporters()`` lists the names:

``hip-kernel``, ``hip-kernel-rl``
    ``kernel.hip``: the model's single-cell step as a device function, with
    its default state (forward Euler, or Rush-Larsen where the model's
    gates allow it).
``hip``, ``hip-rl``
    ``kernel.hip`` plus ``cable.hip``, a runnable cable simulation on the
    GPU, and ``build.sh``.
"""
from .hip import (HipCableExporter, HipCableRLExporter,  # noqa
                  HipKernelExporter, HipKernelRLExporter)

_EXPORTERS = {
    'hip': HipCableExporter,
    'hip-rl': HipCableRLExporter,
    'hip-kernel': HipKernelExporter,
    'hip-kernel-rl': HipKernelRLExporter,
}


def exporter(name, **kwargs):
    """Returns an instance of the exporter ``name`` (keyword arguments go
    to its constructor)."""
    try:
        cls = _EXPORTERS[name]
    except KeyError:
        raise KeyError('Exporter not found: ' + str(name))
    return cls(**kwargs)


def exporters():
    """Names of the available exporters."""
    return sorted(_EXPORTERS.keys())
