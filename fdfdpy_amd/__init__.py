"""MI355X-native drop-in for fdfdpy's 2D FDFD solve path.

``from fdfdpy_amd import Simulation`` mirrors ``from fdfdpy import Simulation``
(fdfdpy/__init__.py:3).
"""

from .simulation import InputError, Simulation

name = "fdfdpy_amd"

__all__ = ["Simulation", "InputError", "name"]
