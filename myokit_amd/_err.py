"""
Exception types of the tissue engine.

The names mirror the exceptions that ``myokit.SimulationOpenCL`` users catch
(``myokit/_err.py``: ``MyokitError:19``, ``ParseError``, ``SimulationError:376``,
``SimulationCancelledError:386``, ``FindNanError:143``, ``ProtocolEventError``,
``SimultaneousProtocolEventError``), so that ``except myokit.X`` code keeps
working when ``myokit_amd`` is imported under the name ``myokit``.
"""


class MyokitError(Exception):
    """Base class for every error raised by this package."""


class ParseError(MyokitError):
    """Raised when an ``mmt`` source cannot be parsed."""

    def __init__(self, name, line=0, char=0, desc=None):
        self.name = name
        self.line = int(line)
        self.char = int(char)
        self.desc = desc
        msg = f'{name} on line {self.line}'
        if desc:
            msg += ': ' + str(desc)
        super().__init__(msg)


class IntegrityError(MyokitError):
    """Raised when a model is invalid (unresolved names, cycles, ...)."""


class MissingRhsError(IntegrityError):
    """Raised when a variable has no right-hand side."""


class IllegalReferenceError(IntegrityError):
    """Raised when a name cannot be resolved in its scope."""


class CyclicalDependencyError(IntegrityError):
    """Raised when the equations of a model contain a cycle."""


class NumericalError(MyokitError):
    """Raised when evaluating an expression fails numerically."""


class ProtocolEventError(MyokitError):
    """Raised for an invalid protocol event."""


class SimultaneousProtocolEventError(MyokitError):
    """Raised when two protocol events (re-)start at the same time."""


class SimulationError(MyokitError):
    """Raised when a simulation fails (e.g. NaN found in the results)."""


class SimulationCancelledError(MyokitError):
    """Raised when a progress reporter cancels a simulation."""

    def __init__(self, message='Operation cancelled by user.'):
        super().__init__(message)


class FindNanError(MyokitError):
    """Raised when the origin of a NaN cannot be determined."""


class DeviceError(MyokitError, RuntimeError):
    """Raised when the HIP device library reports an error."""


class DataLogReadError(MyokitError):
    """Raised when a DataLog file cannot be read."""


class DataBlockReadError(MyokitError):
    """Raised when a DataBlock file cannot be read."""
