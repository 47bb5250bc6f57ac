"""
Pacing protocols and the event-based pacing system.

``Protocol`` / ``ProtocolEvent`` follow the validation and ordering rules of
``myokit/_protocol.py`` (``Protocol.add:52-75``, ``ProtocolEvent:536-589``).
``PacingSystem`` restates the event system of ``myokit/_sim/pacing.h``
(``ESys_ScheduleEvent:191-213``, ``ESys_AdvanceTime:482-548``, float-tolerant
comparisons ``ESys_eq/ESys_geq:69-80``). The device library carries its own
C++ implementation of the same rules; this one serves Python callers
(``Protocol.value_at_times``) and the tests that compare the two.
"""
import sys

import numpy as np

from ._err import ProtocolEventError, SimultaneousProtocolEventError

_EPS = sys.float_info.epsilon


def _eq(a, b):
    if a == b:
        return True
    s = max(abs(a), abs(b))
    return abs(a - b) / s < _EPS


def _geq(a, b):
    return a >= b or _eq(a, b)


class ProtocolEvent:
    """A block of stimulus: ``level`` from ``start`` for ``duration``,
    optionally repeating every ``period`` (``multiplier`` times, 0 = always).
    """

    def __init__(self, level, start, duration, period=0, multiplier=0):
        self._next = None
        self._level = float(level)
        self._start = float(start)
        self._duration = float(duration)
        self._period = float(period)
        self._multiplier = int(multiplier)
        if self._start < 0:
            raise ProtocolEventError(
                'An event can not start at a negative time.')
        if self._duration < 0:
            raise ProtocolEventError(
                'An event can not have a negative duration.')
        if self._period < 0:
            raise ProtocolEventError(
                'An event\'s period must be zero or greater.')
        if self._multiplier < 0:
            raise ProtocolEventError(
                'An event\'s multiplier must be zero or greater')
        if self._period == 0 and self._multiplier > 0:
            raise ProtocolEventError(
                'Non-recurring events can not specify a multiplier.')
        if float(self._multiplier) != float(multiplier):
            raise ProtocolEventError(
                'The event multiplier must be an integer.')
        if self._period > 0 and self._duration > self._period:
            raise ProtocolEventError(
                'A recurring event can not have a duration that\'s longer'
                ' than its period.')

    def level(self):
        return self._level

    def start(self):
        return self._start

    def duration(self):
        return self._duration

    def period(self):
        return self._period

    def multiplier(self):
        return self._multiplier

    def stop(self):
        return self._start + self._duration

    def clone(self):
        return ProtocolEvent(self._level, self._start, self._duration,
                             self._period, self._multiplier)

    def characteristic_time(self):
        if self._period == 0:
            return self._start + self._duration
        if self._multiplier > 0:
            return self._start + self._multiplier * self._period
        return self._period

    def code(self):
        return (f'{self._level!r} {self._start!r} {self._duration!r} '
                f'{self._period!r} {self._multiplier}')

    def __str__(self):
        return ('Event(level=' + str(self._level) + ', start='
                + str(self._start) + ', duration=' + str(self._duration)
                + ', period=' + str(self._period) + ', multiplier='
                + str(self._multiplier) + ')')


class Protocol:
    """An ordered sequence of pacing events."""

    def __init__(self):
        self._head = None

    def add(self, e):
        """Schedules an event, keeping the list ordered by start time."""
        if self._head is None:
            self._head = e
            return
        if e._start < self._head._start:
            e._next = self._head
            self._head = e
            return
        f = self._head
        while f._next is not None and e._start >= f._next._start:
            f = f._next
        if e._start == f._start:
            raise SimultaneousProtocolEventError(
                'Two events cannot (re-)start at the same time: Error at'
                ' time t=' + str(e._start) + '.')
        e._next = f._next
        f._next = e

    def schedule(self, level, start, duration, period=0, multiplier=0):
        self.add(ProtocolEvent(level, start, duration, period, multiplier))

    def add_step(self, level, duration):
        self.schedule(level, self.characteristic_time(), duration)

    def events(self):
        out = []
        e = self._head
        while e is not None:
            out.append(e)
            e = e._next
        return out

    def head(self):
        return self._head

    def clone(self):
        p = Protocol()
        for e in self.events():
            p.add(e.clone())
        return p

    def characteristic_time(self):
        tmax = 0
        for e in self.events():
            tmax = max(tmax, e.characteristic_time())
        return tmax

    def in_words(self):
        return '\n'.join(str(e) for e in self.events())

    def code(self):
        lines = ['[[protocol]]', '# Level Start Length Period Multiplier']
        lines.extend(e.code() for e in self.events())
        return '\n'.join(lines) + '\n'

    def __eq__(self, other):
        return isinstance(other, Protocol) and self.code() == other.code()

    def __str__(self):
        return self.code()

    def value_at_times(self, times):
        if len(times) == 0:
            return []
        times = np.asarray(times)
        if np.any(times[1:] < times[:-1]):
            raise ValueError(
                'The argument `times` must contain a non-decreasing sequence'
                ' of time points.')
        if times[0] < 0:
            raise ValueError('Times cannot be negative.')
        p = PacingSystem(self)
        return [p.advance(t) for t in times]

    def to_arrays(self):
        """Returns the events as an ``(n, 5)`` float64 array of
        ``level, start, duration, period, multiplier`` in queue order."""
        ev = self.events()
        a = np.zeros((len(ev), 5))
        for i, e in enumerate(ev):
            a[i] = (e._level, e._start, e._duration, e._period,
                    e._multiplier)
        return a


class _Ev:
    __slots__ = ('level', 'duration', 'start', 'period', 'multiplier',
                 'next')

    def __init__(self, level, start, duration, period, multiplier):
        self.level = level
        self.start = start
        self.duration = duration
        self.period = period
        self.multiplier = multiplier
        self.next = None


def _schedule(head, add):
    add.next = None
    if head is None:
        return head if add is None else add
    if add.start < head.start:
        add.next = head
        return add
    e = head
    while e.next is not None and add.start >= e.next.start:
        e = e.next
    if add.start == e.start:
        raise SimultaneousProtocolEventError(
            'E-Pacing error: Event scheduled or re-occuring at the same time'
            ' as another event.')
    add.next = e.next
    e.next = add
    return head


class PacingSystem:
    """Event-based pacing: tracks the pacing level as time advances."""

    def __init__(self, protocol=None, initial_time=0):
        self._time = float(initial_time)
        self._head = None
        self._fire = None
        self._tnext = self._time
        self._tdown = self._time
        self._level = 0.0
        if protocol is not None:
            evs = [_Ev(e._level, e._start, e._duration, e._period,
                       float(e._multiplier)) for e in protocol.events()]
            head = None
            for e in evs:
                head = _schedule(head, e)
            self._head = head
        self.advance(self._time)

    def time(self):
        return self._time

    def next_time(self):
        return self._tnext

    def pace(self):
        return self._level

    def advance(self, new_time):
        new_time = float(new_time)
        if new_time < self._time:
            raise ValueError('New time is before current time.')
        self._time = new_time
        while _geq(self._time, self._tnext):
            if self._fire is not None and _geq(self._tnext, self._tdown):
                self._fire = None
                self._level = 0.0
            if self._head is not None and _geq(self._tnext,
                                               self._head.start):
                f = self._fire = self._head
                self._head = f.next
                self._tdown = f.start + f.duration
                self._level = f.level
                if f.period > 0:
                    if f.multiplier != 1:
                        if f.multiplier > 1:
                            f.multiplier -= 1
                        f.start += f.period
                        self._head = _schedule(self._head, f)
                    else:
                        f.period = 0
                if self._head is not None and _eq(self._head.start,
                                                  self._tdown):
                    self._tdown = self._head.start
            self._tnext = float('inf')
            if self._fire is not None and self._tnext > self._tdown:
                self._tnext = self._tdown
            if self._head is not None and self._tnext > self._head.start:
                self._tnext = self._head.start
        return self._level
