"""
Progress reporting interface (``myokit/_progress.py:12-146``): a reporter is
used as ``with reporter.job(msg): ... reporter.update(fraction)``, and
``update`` returning False cancels the simulation.
"""
import contextlib
import sys
import time


class ProgressReporter:
    """Base class for progress reporters."""

    def enter(self, msg=None):
        pass

    def exit(self):
        pass

    @contextlib.contextmanager
    def job(self, msg=None):
        self.enter(msg)
        try:
            yield
        finally:
            self.exit()

    def update(self, progress):
        raise NotImplementedError


class ProgressPrinter(ProgressReporter):
    """Prints a line every ``digits``-rounded percentage change."""

    def __init__(self, digits=1):
        self._digits = int(digits)
        self._last = None
        self._t0 = None

    def enter(self, msg=None):
        self._t0 = time.time()
        self._last = None
        if msg:
            print(msg)

    def update(self, f):
        p = round(100 * f, self._digits)
        if p != self._last:
            self._last = p
            print('[{:.1f}s] {} % done'.format(time.time() - self._t0, p))
            sys.stdout.flush()
        return True


class Timeout(ProgressReporter):
    """Cancels a simulation after ``seconds`` of wall time."""

    def __init__(self, seconds):
        self._max = float(seconds)
        self._t0 = None

    def enter(self, msg=None):
        self._t0 = time.time()

    def update(self, f):
        return time.time() - self._t0 < self._max
