"""
Factory functions for pacing protocols (``myokit/pacing.py``).
"""
from ._protocol import Protocol


def blocktrain(period, duration, offset=0, level=1.0, limit=0):
    """A train of block pulses: ``duration`` long, every ``period``, first at
    ``offset``; ``limit`` pulses in total (0 = unlimited)."""
    p = Protocol()
    p.schedule(level, offset, duration, period, limit)
    return p


def bpm2bcl(bpm, m=1e-3):
    """Converts beats per minute to a basic cycle length (in ms by
    default)."""
    return 60.0 / (m * bpm)
