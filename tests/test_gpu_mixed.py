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


def run_mixed(problem, steps):
    og = oracle_grid(problem)
    gg = gpu_grid(problem, precision=1)
    n_o = og.advance(1e9, True, steps)
    n_g = gg.advance(1e9, True, steps)
    res = compare_states(gg, og, problem[0], 8)
    divb = gg.divb()
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
