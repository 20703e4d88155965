"""Exceptions, as the reference's Python package defines them (python/scatrs/errors.py:1-80)."""

from __future__ import annotations

__all__ = [
    "InvalidInputError",
    "NotConvergedError",
    "OutsideValidatedDomainError",
    "ScatrsError",
    "NativeLibraryError",
]


class ScatrsError(Exception):
    """Base class for every error raised by scatrs_amd."""


class InvalidInputError(ScatrsError, ValueError):
    """An input is outside its valid range (nothing was computed)."""


class NotConvergedError(ScatrsError, ArithmeticError):
    """A computation did not reach the requested accuracy."""

    def __init__(self, message: str, achieved: float | None = None,
                 requested: float | None = None) -> None:
        super().__init__(message)
        self.achieved = achieved
        self.requested = requested


class OutsideValidatedDomainError(ScatrsError, ValueError):
    """Inputs are valid but outside the domain where accuracy has been validated."""


class NativeLibraryError(ScatrsError, RuntimeError):
    """The HIP library is missing, failed to load, or a device call failed."""
