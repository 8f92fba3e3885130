"""Input validation of the drop-in Simulation (tests/test_simulation.py of the reference).

Every case raises before any device work, so these run without a GPU. The
reference asserts (AssertionError) where its own tests expect ValueError;
InputError is both.
"""

import numpy as np
import pytest

from fdfdpy_amd.simulation import InputError, Simulation


@pytest.fixture
def good():
    return dict(omega=100, eps_r=np.ones((100, 50)), dl=0.001, NPML=[10, 10], pol='Hz', L0=1e-4)


def make(g, **kw):
    a = dict(g)
    a.update(kw)
    return Simulation(a['omega'], a['eps_r'], a['dl'], a['NPML'], a['pol'], a['L0'])


@pytest.mark.parametrize("kw", [
    dict(omega=-100), dict(omega=[100, 200, 300]),                 # test_freq
    dict(eps_r=-np.ones((100, 50))), dict(eps_r=list(np.ones((100, 50)))),  # test_eps
    dict(dl=-0.001), dict(dl=[1e-4, 1e-5]),                        # test_dl
    dict(NPML=10), dict(NPML=[10, 10, 10]), dict(NPML=[200, 200]),  # test_NPML
    dict(pol=5), dict(pol='WrongPolarization'),                    # test_pol
    dict(L0=-1.0), dict(NPML=[-1, 3]), dict(eps_r=np.full((100, 50), np.nan)),
])
def test_bad_inputs_raise_value_and_assertion_error(good, kw):
    with pytest.raises(ValueError):
        make(good, **kw)
    with pytest.raises(AssertionError):
        make(good, **kw)
    with pytest.raises(InputError):
        make(good, **kw)


def test_package_import_surface():
    """from fdfdpy_amd import Simulation, as from fdfdpy import Simulation
    (fdfdpy/__init__.py:3)."""
    import fdfdpy_amd
    from fdfdpy_amd import InputError as IE
    from fdfdpy_amd import Simulation as S
    assert S is Simulation and IE is InputError
    assert fdfdpy_amd.name == "fdfdpy_amd"


def test_grid_beyond_the_largest_block_raises_input_error():
    """The direct solver's block is the shorter grid side, at most 2048 cells;
    a larger grid is refused before any device work."""
    with pytest.raises(InputError):
        Simulation(1e15, np.ones((2056, 2050)), 0.02, [20, 20], 'Ez')
