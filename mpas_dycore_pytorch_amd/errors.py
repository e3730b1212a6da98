"""Error types of the public API.

Same names, bases and constructor keywords as the reference hierarchy
(``mpas_pytorch/exceptions.py:30-154``) so callers that catch the
reference's errors keep working.
"""

from __future__ import annotations

from typing import Iterable


class MpasError(Exception):
    """Umbrella type for every user-facing failure."""


class MpasConfigError(MpasError, ValueError):
    """Invalid configuration value (also a ValueError, as in the reference)."""


class MpasMissingFieldError(MpasConfigError):
    """A required state or mesh field is absent."""

    def __init__(self, message: str | None = None, *, fields: Iterable[str] | None = None,
                 source: str = "unknown") -> None:
        self.fields: tuple[str, ...] = tuple(fields or ())
        self.source = str(source)
        if message is None and self.fields:
            message = f"Missing required fields from {self.source}: {self.fields}"
        super().__init__(message if message is not None else "")


class MpasMeshMismatchError(MpasConfigError):
    """State and mesh dimensions disagree."""

    def __init__(self, message: str | None = None, *, dim_name: str | None = None,
                 init_value: int | None = None, grid_value: int | None = None) -> None:
        self.dim_name, self.init_value, self.grid_value = dim_name, init_value, grid_value
        if message is None and dim_name is not None:
            message = f"Mesh dimension mismatch for {dim_name}: state={init_value}, mesh={grid_value}"
        super().__init__(message if message is not None else "")


class MpasDeviceError(MpasConfigError):
    """The GPU runtime or the HIP library is unavailable."""

    def __init__(self, message: str | None = None, *, requested: str | None = None) -> None:
        self.requested = "" if requested is None else str(requested)
        if message is None and self.requested:
            message = f"{self.requested!r} requested but no ROCm GPU is visible"
        super().__init__(message if message is not None else "")


class MpasStrictModeError(MpasConfigError):
    """Strict validation rejected an input."""


class MpasUnsupportedCapabilityError(MpasConfigError):
    """A configuration outside what the GPU step implements."""

    def __init__(self, message: str | None = None, *, capability: str, requested: object) -> None:
        self.capability = capability
        self.requested = requested
        super().__init__(message or f"Unsupported {capability}: {requested!r}")


class MpasNumericalAgreementError(MpasError):
    """A numerical-agreement contract failed."""
