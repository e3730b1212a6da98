"""MI355X (gfx950) split-explicit dynamics step for the MPAS-A dycore.

Drop-in GPU replacement of the reference's ``atm_srk3`` large step
(``mpas_pytorch/dynamics/time_integration.py:4897``): hand-written HIP
kernels behind a C ABI (``include/mpas_hip.h``), with this package as the
Python host that mirrors the reference interface.

Importing the package does not touch the GPU; the HIP library is loaded on
the first step and its absence is an error (there is no CPU fallback).
"""

from .errors import (  # noqa: F401
    MpasConfigError,
    MpasDeviceError,
    MpasError,
    MpasMeshMismatchError,
    MpasMissingFieldError,
    MpasNumericalAgreementError,
    MpasStrictModeError,
    MpasUnsupportedCapabilityError,
)
from .time_integration import Srk3Stepper, atm_srk3  # noqa: F401
from .atmosphere_model import run_simulation, run_timestep  # noqa: F401

__all__ = ["atm_srk3", "Srk3Stepper", "run_timestep", "run_simulation", "MpasError", "MpasConfigError",
           "MpasDeviceError",
           "MpasMeshMismatchError", "MpasMissingFieldError", "MpasNumericalAgreementError",
           "MpasStrictModeError", "MpasUnsupportedCapabilityError"]
