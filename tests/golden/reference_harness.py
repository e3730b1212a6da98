"""Drive the reference implementation to write golden vectors.

Runs only in the build container, where the reference tree is mounted at
``/root/reference``; nothing on the GPU side imports this file.  It imports
the reference's hot-path modules (``mpas_pytorch.dynamics``,
``mpas_pytorch.autodiff`` and the modules they import) without executing the
two package ``__init__`` files whose I/O imports (netCDF4, Python 3.11
``datetime.UTC`` / ``typing.Self``) are unavailable here; no stand-in is
provided for any missing library.  Bytecode writing is disabled so the
reference tree is left untouched.
"""

from __future__ import annotations

import importlib
import os
import sys
import types

sys.dont_write_bytecode = True
os.environ.setdefault("ATEN_CPU_CAPABILITY", "default")

REFERENCE_ROOT = "/root/reference"


def import_reference():
    """Return the reference modules used to write fixtures."""
    if "mpas_pytorch" not in sys.modules:
        for name, sub in (("mpas_pytorch", "mpas_pytorch"), ("mpas_pytorch.utils", "mpas_pytorch/utils")):
            pkg = types.ModuleType(name)
            pkg.__path__ = [os.path.join(REFERENCE_ROOT, sub)]
            sys.modules[name] = pkg
    mods = {}
    for short, full in (
        ("ti", "mpas_pytorch.dynamics.time_integration"),
        ("icd", "mpas_pytorch.dynamics.init_coupled_diagnostics"),
        ("sd", "mpas_pytorch.dynamics.solve_diagnostics"),
        ("vr", "mpas_pytorch.dynamics.vector_reconstruction"),
        ("cfg", "mpas_pytorch.config_defaults"),
        ("const", "mpas_pytorch.constants"),
    ):
        mods[short] = importlib.import_module(full)
    return types.SimpleNamespace(**mods)
