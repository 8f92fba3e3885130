"""MI355X-native drop-in for fdfdpy's 2D FDFD solve path.

``from fdfdpy_amd import Simulation`` mirrors ``from fdfdpy import Simulation``
(fdfdpy/__init__.py:3). The re-export is resolved on first access, so that
host-only modules of the package (``fdfdpy_amd.configs``) import without
PyTorch.
"""

name = "fdfdpy_amd"

__all__ = ["Simulation", "InputError", "name"]


def __getattr__(attr):
    if attr in ("Simulation", "InputError"):
        from . import simulation
        return getattr(simulation, attr)
    raise AttributeError(f"module {__name__!r} has no attribute {attr!r}")
