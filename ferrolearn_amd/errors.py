"""Python mirror of ferrolearn-core/src/error.rs (FerroError) for this path."""
from __future__ import annotations

from . import _lib as L


class FerroError(Exception):
    """Base class, one subclass per FerroError variant used on this path."""
    status = -1


class ShapeMismatch(FerroError):
    status = L.FL_SHAPE_MISMATCH


class InsufficientSamples(FerroError):
    status = L.FL_INSUFFICIENT_SAMPLES


class InvalidParameter(FerroError):
    status = L.FL_INVALID_PARAMETER


class NumericalInstability(FerroError):
    status = L.FL_NUMERICAL_INSTABILITY


class DeviceError(FerroError):
    """A HIP runtime failure (no counterpart in the reference)."""
    status = L.FL_DEVICE_ERROR


_BY_STATUS = {c.status: c for c in (ShapeMismatch, InsufficientSamples, InvalidParameter,
                                    NumericalInstability, DeviceError)}


def check(status: int) -> None:
    if status != L.FL_OK:
        raise _BY_STATUS.get(status, FerroError)(L.last_error())
