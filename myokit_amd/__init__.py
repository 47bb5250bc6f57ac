"""
myokit_amd -- myokit's multi-cell tissue simulator (``SimulationOpenCL``)
rebuilt for AMD Instinct MI355X (gfx950) with HIP.

``import myokit_amd as myokit`` gives the parts of the myokit API that the
tissue path needs: loading models (``load_model('example')``,
``parse_model``), protocols and pacing, ``DataLog`` / ``prepare_log``, the
precision and log flags, the exception types, and ``SimulationOpenCL``
itself, whose kernels are generated per model and compiled with hiprtc.
"""
from ._err import (  # noqa
    CyclicalDependencyError, DataBlockReadError, DataLogReadError,
    DeviceError, FindNanError,
    IllegalReferenceError,
    IntegrityError, MissingRhsError, MyokitError, NumericalError,
    ParseError,
    ProtocolEventError, SimulationCancelledError, SimulationError,
    SimultaneousProtocolEventError)
from ._datablock import DataBlock1d, DataBlock2d  # noqa
from ._datalog import (  # noqa
    DOUBLE_PRECISION, LOG_ALL, LOG_BOUND, LOG_DERIV, LOG_INTER, LOG_NONE,
    LOG_STATE, SINGLE_PRECISION, DataLog, prepare_log, split_key)
from ._expressions import (  # noqa
    ACos, And, ASin, ATan, Ceil, Cos, Derivative, Divide, Equal, Exp,
    Expression, Floor, If, Less, LessEqual, Log, Log10, Minus, More,
    MoreEqual, Multiply, Name, Not, NotEqual, Number, Or, Piecewise, Plus,
    Power, PrefixMinus, PrefixPlus, Quotient, Remainder, Sin, Sqrt, Tan, Abs)
from ._model import Component, Model, Variable  # noqa
from ._parser import (  # noqa
    load, load_model, load_protocol, parse, parse_expression, parse_model,
    parse_protocol)
from ._progress import ProgressPrinter, ProgressReporter, Timeout  # noqa
from ._protocol import PacingSystem, Protocol, ProtocolEvent  # noqa
from . import pacing  # noqa
from .simulation import SimulationOpenCL  # noqa
from .fiber_tissue import FiberTissueSimulation  # noqa

__version__ = '0.1.0'
