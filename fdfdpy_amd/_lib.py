"""ctypes binding of libfdfd_hip.so (the C ABI declared in include/fdfd_hip.h).

The library is built in-tree by ``fdfdpy_amd/csrc/build.py``. Loading fails
loudly when it is missing: there is no CPU fallback for any entry point.
"""

from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_double, c_int, c_int32, c_size_t, c_void_p

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libfdfd_hip.so")

FDFD_OK = 0
FDFD_ERR_ARG = 1
FDFD_ERR_HIP = 2
FDFD_ERR_NOT_CONVERGED = 3
FDFD_ERR_NUMERIC = 4

POL = {"Ez": 0, "Hz": 1}
PREC = {"fp64": 0, "c64": 1}
FACTOR_PIVOT = 0x100     # FDFD_FACTOR_PIVOT
FACTOR_STEPWISE = 0x200  # FDFD_FACTOR_STEPWISE


class FdfdGrid(ctypes.Structure):
    _fields_ = [("nx", c_int32), ("ny", c_int32), ("npml_x", c_int32), ("npml_y", c_int32),
                ("pol", c_int32), ("reserved", c_int32),
                ("omega", c_double), ("dl", c_double), ("L0", c_double)]


class FdfdSolveStats(ctypes.Structure):
    _fields_ = [("iterations", c_int32), ("converged", c_int32), ("stalled", c_int32),
                ("zero_rhs", c_int32), ("max_rel_residual", c_double), ("restarts", c_int32),
                ("sweeps", c_int32)]


class FdfdError(RuntimeError):
    """A non-zero status returned by the C ABI."""

    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


class NotConverged(FdfdError):
    pass


# name -> (restype, argtypes); every entry point of include/fdfd_hip.h
_G = POINTER(FdfdGrid)
SIGNATURES = {
    "fdfd_abi_version": (c_int, []),
    "fdfd_last_error": (ctypes.c_char_p, []),
    "fdfd_block_padding": (c_int, []),
    "fdfd_op_bytes": (c_size_t, [_G]),
    "fdfd_factor_bytes": (c_size_t, [_G, c_int]),
    "fdfd_factor_workspace_bytes": (c_size_t, [_G, c_int]),
    "fdfd_solve_workspace_bytes": (c_size_t, [_G, c_int, c_int]),
    "fdfd_assemble": (c_int, [_G, c_void_p, c_void_p, c_void_p]),
    "fdfd_sfactors": (c_int, [_G, c_void_p, c_void_p, c_void_p, c_void_p]),
    "fdfd_flux": (c_int, [c_int, c_int, c_double, c_int, c_void_p, c_void_p, c_int, c_int, c_int, c_void_p,
                          c_void_p]),
    "fdfd_apply": (c_int, [_G, c_void_p, c_void_p, c_void_p, c_int, c_void_p]),
    "fdfd_residual": (c_int, [_G, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p]),
    "fdfd_factor": (c_int, [_G, c_void_p, c_void_p, c_int, c_void_p, c_void_p]),
    "fdfd_factor_batch": (c_int, [_G, POINTER(c_void_p), POINTER(c_void_p), c_int, c_int, POINTER(c_void_p),
                                  c_void_p]),
    "fdfd_factor_wrap": (c_int, [_G, c_void_p, c_void_p, c_int, c_void_p, c_void_p]),
    "fdfd_factor_status": (c_int, [_G, c_int, c_void_p, c_void_p]),
    "fdfd_sweep": (c_int, [_G, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_int, c_void_p, c_void_p]),
    "fdfd_solve": (c_int, [_G, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_int, c_int, c_double,
                           c_int, c_void_p, POINTER(FdfdSolveStats), POINTER(c_double), c_void_p]),
    "fdfd_solve_static": (c_int, [_G, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_int, c_int, c_double,
                                  c_int, POINTER(FdfdSolveStats), POINTER(c_double), c_void_p]),
    "fdfd_norms": (c_int, [_G, c_void_p, c_int, c_void_p, POINTER(c_double), c_void_p]),
    "fdfd_stretch_scale": (c_int, [_G, c_void_p, c_int, c_int, c_void_p]),
    "fdfd_grad_eps": (c_int, [_G, c_void_p, c_void_p, c_void_p, c_void_p]),
    "fdfd_kerr_update": (c_int, [_G, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "fdfd_fields": (c_int, [_G, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p]),
    "fdfd_selftest_mfma": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "fdfd_engine_create": (c_void_p, [c_int]),
    "fdfd_engine_destroy": (c_int, [c_void_p]),
    "fdfd_engine_submit": (c_int, [c_void_p, _G, c_void_p, c_void_p, c_int, c_void_p, POINTER(ctypes.c_uint64)]),
    "fdfd_engine_wait": (c_int, [c_void_p, ctypes.c_uint64]),
    "fdfd_engine_launches": (ctypes.c_uint64, [c_void_p]),
    "fdfd_set_spin_cap": (None, [ctypes.c_uint]),
    "fdfd_set_step_symmetry": (c_int, [c_int]),
    "fdfd_set_twist": (c_int, [c_int]),
    "fdfd_twist_block": (c_int, [_G]),
}

_lib = None


def lib():
    """The loaded library; raises when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                "(the HIP library is required; there is no CPU fallback)")
        # torch first: the library then binds to the HIP runtime torch
        # already holds (loaded before torch, the library's engine could not
        # create its stream on the device)
        import torch  # noqa: F401
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != FDFD_OK:
        msg = lib().fdfd_last_error().decode(errors="replace")
        if rc == FDFD_ERR_NOT_CONVERGED:
            raise NotConverged(rc, msg)
        if rc == FDFD_ERR_ARG:
            raise ValueError(msg)
        raise FdfdError(rc, msg)
    return rc
