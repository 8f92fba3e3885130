"""Physical constants and defaults, same values as fdfdpy/constants.py:3-10."""

from numpy import sqrt

EPSILON_0 = 8.85418782e-12
MU_0 = 1.25663706e-6
C_0 = sqrt(1 / EPSILON_0 / MU_0)
ETA_0 = sqrt(MU_0 / EPSILON_0)

DEFAULT_MATRIX_FORMAT = 'csr'
# the device direct solver; the reference's names 'pardiso' and 'scipy' are
# accepted as aliases of it (fdfdpy/constants.py:9 defaults to 'pardiso')
DEFAULT_SOLVER = 'hip'
DEFAULT_LENGTH_SCALE = 1e-6  # microns
