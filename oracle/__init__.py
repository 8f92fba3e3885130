"""CPU oracle for the FDFD solve path (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker of the HIP path
or as the timed CPU baseline. The product package ``fdfdpy_amd`` never imports
it.
"""

from .fdfd_oracle import *  # noqa: F401,F403
