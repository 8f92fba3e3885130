"""The C ABI library loads and exports every declared symbol (no GPU calls)."""

import ctypes
import os
import re

import pytest

from conftest import ROOT
from fdfdpy_amd import _lib as L

HEADER = os.path.join(ROOT, "include", "fdfd_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fdfd_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    names = declared_functions()
    assert len(names) >= 18
    for n in names:
        assert hasattr(lib, n), n
        assert n in L.SIGNATURES, f"{n} has no ctypes signature"
    version = int(re.search(r"#define FDFD_ABI_VERSION (\d+)", open(HEADER).read()).group(1))
    assert lib.fdfd_abi_version() == version


def test_sizes():
    lib = L.lib()
    g = L.FdfdGrid(512, 512, 20, 20, 0, 0, 1.2e15, 0.02, 1e-6)
    assert lib.fdfd_block_padding() == 32
    assert lib.fdfd_op_bytes(ctypes.byref(g)) == 5 * 512 * 512 * 16
    # block inverses, then the complex FP64 inverse of the 2n x 2n wrap capacitance matrix
    kinv = 16 * 1024 * 1024
    assert lib.fdfd_factor_bytes(ctypes.byref(g), 0) == 512 * 2 * 512 * 512 * 8 + kinv
    assert lib.fdfd_factor_bytes(ctypes.byref(g), 1) == 512 * 2 * 512 * 512 * 4 + kinv
    # block size padded to the tile edge
    g2 = L.FdfdGrid(300, 300, 20, 20, 0, 0, 1.2e15, 0.02, 1e-6)
    assert lib.fdfd_factor_bytes(ctypes.byref(g2), 0) == 300 * 2 * 320 * 320 * 8 + 16 * 600 * 600
    assert lib.fdfd_solve_workspace_bytes(ctypes.byref(g2), 0, 4) > 16 * 300 * 300 * 4


@pytest.mark.parametrize("bad", [
    dict(nx=1), dict(ny=0), dict(ny=2049), dict(npml_x=-1), dict(pol=2), dict(omega=-1.0), dict(dl=0.0),
])
def test_invalid_grid_rejected_without_gpu(bad):
    lib = L.lib()
    f = dict(nx=64, ny=64, npml_x=10, npml_y=10, pol=0, reserved=0, omega=1.2e15, dl=0.02, L0=1e-6)
    f.update(bad)
    g = L.FdfdGrid(**f)
    assert lib.fdfd_op_bytes(ctypes.byref(g)) == 0
    assert lib.fdfd_last_error()
    rc = lib.fdfd_assemble(ctypes.byref(g), None, None, None)
    assert rc == L.FDFD_ERR_ARG
    with pytest.raises(ValueError):
        L.check(rc)


def test_null_pointers_rejected():
    lib = L.lib()
    g = L.FdfdGrid(64, 64, 10, 10, 0, 0, 1.2e15, 0.02, 1e-6)
    assert lib.fdfd_factor(ctypes.byref(g), None, None, 0, None, None) == L.FDFD_ERR_ARG
    assert lib.fdfd_factor(ctypes.byref(g), None, None, 7, None, None) == L.FDFD_ERR_ARG
    st = L.FdfdSolveStats()
    assert lib.fdfd_solve(ctypes.byref(g), None, None, 0, None, None, 1, 0, 1e-10, 10, None,
                          ctypes.byref(st), None, None) == L.FDFD_ERR_ARG
    assert b"fdfd_solve" in lib.fdfd_last_error()
