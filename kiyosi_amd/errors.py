"""Exception classes mirroring kiyosi's Python error mapping.

crates/kiyosi-python/src/error.rs:53-85 maps ``Error::InvalidInput`` to
``InvalidInputError``, ``UnsupportedOperation`` to ``UnsupportedOperationError``,
``ComputationFailure`` to ``ComputationFailureError``, ``MissingHistory`` to
``MissingHistoryError`` (all subclasses of ``KiyosiError``) and ``Io`` to
``OSError`` (``InterruptedError`` when cancelled).
"""

from . import _abi


class KiyosiError(Exception):
    """Base class of kiyosi pricing errors."""


class InvalidInputError(KiyosiError, ValueError):
    """Invalid contract, market or engine input."""


class UnsupportedOperationError(KiyosiError):
    """The engine does not support this model, curve or operation."""


class ComputationFailureError(KiyosiError, ArithmeticError):
    """A non-finite path, payment or result."""


class MissingHistoryError(KiyosiError):
    """Prior barrier history is required but unknown."""


class RootNotBracketedError(KiyosiError):
    """Error::RootNotBracketed: a solver's bounds do not bracket the target."""


class NonConvergenceError(KiyosiError, ArithmeticError):
    """Error::NonConvergence: a solver or series stopped without converging."""


class DeviceError(OSError):
    """A HIP runtime failure on the GPU."""


def raise_for_status(status: int, message: str) -> None:
    if status == _abi.OK:
        return
    if status == _abi.INVALID_INPUT:
        raise InvalidInputError(message)
    if status == _abi.UNSUPPORTED_OPERATION:
        raise UnsupportedOperationError(message)
    if status == _abi.COMPUTATION_FAILURE:
        raise ComputationFailureError(message)
    if status == _abi.INTERRUPTED:
        raise InterruptedError(message)
    if status == _abi.MISSING_HISTORY:
        raise MissingHistoryError(message)
    if status == _abi.NON_CONVERGENCE:
        raise NonConvergenceError(message)
    raise DeviceError(message)
