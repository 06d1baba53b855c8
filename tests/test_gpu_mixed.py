"""Mixed-precision form (FP32 reconstruction, limiters and Riemann solves;
FP64 state, accumulation, pressure recovery and CT) against the FP64 oracle.

Bar of the mixed form: relative L1(rho) <= 1e-5 against the FP64 oracle,
normalised max|div B| <= 1e-12.
"""
import numpy as np
import pytest

from dragon_amd import problems as P

from helpers import compare_states, gpu_grid, oracle_grid

pytestmark = pytest.mark.gpu


def run_mixed(problem, steps=None, t=None, nf=8):
    og = oracle_grid(problem)
    gg = gpu_grid(problem, precision=1)
    if t is not None:
        n_o, n_g = og.advance(t), gg.advance(t)
    else:
        n_o, n_g = og.advance(1e9, True, steps), gg.advance(1e9, True, steps)
    res = compare_states(gg, og, problem[0], nf)
    divb = gg.divb() if nf == 8 else 0.0
    print(problem[0], n_o, n_g, "rel L1", res["rel_l1"], "divB", divb)
    return res, divb


def test_mixed_orszag_tang_2d():
    res, divb = run_mixed(P.orszag_tang_2d(64), 40)
    assert res["rel_l1"][0] <= 1e-5
    assert divb <= 1e-12


def test_mixed_turbulence_3d():
    res, divb = run_mixed(P.turbulence_3d(16), 12)
    assert res["rel_l1"][0] <= 1e-5
    assert divb <= 1e-12


def test_mixed_sod_1d_full():
    """C1 in full: 1D Sod, 1024 cells, HLLC + minmod + verify fallback, to t = 0.2."""
    pb = P.sod_1d(1024)
    res, _ = run_mixed(pb, t=pb[6]["t_end"], nf=5)
    assert res["rel_l1"][0] <= 1e-5


def test_mixed_kelvin_helmholtz_2d():
    """C2 switch set (hydro CTU 4-solve, HLLC, van Leer, periodic) at 64^2, 60 steps."""
    res, _ = run_mixed(P.kelvin_helmholtz_2d(64), 60, nf=5)
    assert res["rel_l1"][0] <= 1e-5


def test_mixed_hydro_3d_ctu12():
    """3D hydro CTU 12-solve (diagonal advection) at 16^3, 10 steps."""
    res, _ = run_mixed(P.diagonal_3d(16), 10, nf=5)
    assert res["rel_l1"][0] <= 1e-5


@pytest.mark.parametrize("shape", [(2, 3), (2, 1, 3)])
def test_mixed_tiny_grids(shape):
    """The mixed form's walkers and wavefronts on grids of a few cells."""
    from test_gpu_parity import tiny_mhd
    res, divb = run_mixed(tiny_mhd(shape, 1), 2)
    assert res["rel_l1"][0] <= 1e-5
    assert divb <= 1e-12


@pytest.mark.parametrize("key,value", [("limiter", v) for v in range(5)] + [("ct_energy", v) for v in (0, 1, 2)])
def test_mixed_switches_mhd(key, value):
    """The mixed form under each limiter and CT energy mode (2D MHD)."""
    from test_gpu_parity import tiny_mhd
    pb = tiny_mhd((24, 20), 1)
    pb[2][key] = value
    res, divb = run_mixed(pb, 6)
    assert res["rel_l1"][0] <= 1e-5
    assert divb <= 1e-12


@pytest.mark.parametrize("solver", [0, 1, 2, 4])
def test_mixed_solvers_hydro(solver):
    """The mixed form's exact, HLL, HLLC and HLLE solvers (2D hydro)."""
    from test_gpu_parity import random_hydro
    res, _ = run_mixed(random_hydro((24, 20), solver), 6, nf=5)
    assert res["rel_l1"][0] <= 1e-5
