"""HIP path vs the CPU oracle on the same seeded inputs (FP64 drop-in).

Bar of the FP64 drop-in: relative L1 per variable <= 1e-10 after the run.
The FP64 kernels are built without FMA contraction, so in practice the
results are expected to be bit-identical; the bit count is printed.
"""
import numpy as np
import pytest

import oracle as O
from dragon_amd import problems as P

from helpers import compare_states, gpu_grid, oracle_grid

pytestmark = pytest.mark.gpu
TOL = 1e-10


def run_pair(problem, steps=None, t=None, nf=8):
    og = oracle_grid(problem)
    gg = gpu_grid(problem)
    if t is not None:
        n_o = og.advance(t)
        n_g = gg.advance(t)
    else:
        n_o = og.advance(1e9, True, steps)
        n_g = gg.advance(1e9, True, steps)
    assert n_o == n_g, (n_o, n_g)
    res = compare_states(gg, og, problem[0], nf)
    print(problem[0], "steps", n_o, res)
    assert np.all(res["rel_l1"] <= TOL), res
    if "B_bit_equal" in res:
        assert res["B_max_abs"] <= 1e-12, res
    return res


def test_c1_sod_1d_full():
    """C1: 1D Sod, 1024 cells, HLLC + minmod + verify fallback, to t=0.2."""
    pb = P.sod_1d(1024)
    res = run_pair(pb, t=pb[6]["t_end"], nf=5)
    assert res["bit_equal"]


def test_c2_kh_2d_small():
    """C2 switch set (hydro CTU, HLLC, van Leer, periodic) at 64^2, 60 steps."""
    run_pair(P.kelvin_helmholtz_2d(64), steps=60, nf=5)


def test_c3_orszag_tang_2d_small():
    """C3 switch set (MHD CTU + CT, HLLD, MC, beta-gated) at 64^2, 40 steps."""
    run_pair(P.orszag_tang_2d(64), steps=40)


def test_c4_turbulence_3d_small():
    """C4 switch set (3D MHD CTU + CT, HLLD, van Leer) at 16^3, 12 steps."""
    run_pair(P.turbulence_3d(16), steps=12)


def test_c5_rotor_patch():
    """C5 switch set (MHD, gamma=1.4, outflow, minmod) on one 48^2 patch, 30 steps."""
    run_pair(P.rotor_2d(48, B0=4.0, r0=0.1, u0=1.5), steps=30)


def test_hydro_3d_ctu12():
    """3D hydro CTU 12-solve (diagonal advection) at 16^3, 10 steps."""
    run_pair(P.diagonal_3d(16), steps=10, nf=5)


def test_fp64_parity_100_steps_2d_mhd():
    """The drop-in bar: <= 1e-10 relative L1 per variable after 100 steps."""
    run_pair(P.orszag_tang_2d(48), steps=100)


def rect_mhd_3d(shape=(40, 9, 70)):
    """Non-cubic 3D MHD (CTU + HLLD + CT, van Leer, periodic) with a random
    state, 40 x 9 x 70 cells by default."""
    g = P.ghosts_for(True)
    w, B = P.empty_state(shape, g, True)
    rng = np.random.default_rng(7)
    wi = P.interior(w, g, shape)
    wi[..., 0] = 1.0 + 0.3 * rng.random(shape)
    wi[..., 1] = 1.0 + 0.3 * rng.random(shape)
    wi[..., 2:5] = 0.5 * (rng.random(shape + (3,)) - 0.5)
    P.interior(B, g, shape, extra=1)[..., 0] = 0.2
    cfg = dict(mhd=1, gamma=5.0 / 3.0, solver=5, limiter=2, ctu=1, ct_energy=2)
    return (shape, tuple(1.0 / n for n in shape), cfg, [(1, 63)], w, B, {"init_B": True})


def test_mhd_3d_rectangular():
    """rect_mhd_3d: several i-runs, j-tiles and k-tiles of the blocked EMF pass,
    partial tiles on every side."""
    run_pair(rect_mhd_3d(), steps=5)


def tiny_mhd(shape, bc):
    """Random MHD state (CTU + HLLD + CT, van Leer) on a grid of a few cells,
    outflow (0) or periodic (1) on every face."""
    g = P.ghosts_for(True)
    w, B = P.empty_state(shape, g, True)
    rng = np.random.default_rng(3)
    wi = P.interior(w, g, shape)
    wi[..., 0] = 1.0 + 0.3 * rng.random(shape)
    wi[..., 1] = 1.0 + 0.3 * rng.random(shape)
    wi[..., 2:5] = 0.3 * (rng.random(shape + (3,)) - 0.5)
    P.interior(B, g, shape, extra=1)[..., 0] = 0.2
    cfg = dict(mhd=1, gamma=5.0 / 3.0, solver=5, limiter=2, ctu=1, ct_energy=2)
    allf = {2: 15, 3: 63}[len(shape)]
    return (shape, tuple(1.0 / n for n in shape), cfg, [(bc, allf)], w, B, {"init_B": True})


def random_hydro(shape, solver, ctu=1, muscl=1):
    """Random hydro state, periodic on every face, the given Riemann solver id."""
    g = P.ghosts_for(False, ctu=bool(ctu), muscl=bool(muscl))
    w, B = P.empty_state(shape, g, False)
    rng = np.random.default_rng(5)
    wi = P.interior(w, g, shape)
    wi[..., 0] = 1.0 + 0.3 * rng.random(shape)
    wi[..., 1] = 1.0 + 0.3 * rng.random(shape)
    wi[..., 2:5] = 0.3 * (rng.random(shape + (3,)) - 0.5)
    cfg = dict(mhd=0, gamma=5.0 / 3.0, solver=solver, limiter=2, ctu=ctu, muscl=muscl)
    allf = {1: 3, 2: 15, 3: 63}[len(shape)]
    return (shape, tuple(1.0 / n for n in shape), cfg, [(1, allf)], w, B, {})


@pytest.mark.parametrize("solver", [0, 1, 2, 3, 4, 5])
def test_solver_choice_hydro(solver):
    """The runtime solver choice (Flux.cpp:34-45): HLLX and HLLC take HLLC, HLL
    and HLLE themselves, and every other id -- the exact solver's and HLLD's,
    which is MHD only -- the exact solver."""
    run_pair(random_hydro((12, 10), solver), steps=4, nf=5)


@pytest.mark.parametrize("solver", [0, 1, 2, 3, 4, 5])
def test_solver_choice_mhd(solver):
    """The runtime solver choice for MHD (Flux.cpp:46-53): HLL, HLLE, and HLLD
    for every other id."""
    pb = tiny_mhd((6, 5), 1)
    pb[2]["solver"] = solver
    run_pair(pb, steps=4)


SWITCHES = ([("limiter", v) for v in range(5)] + [("cfl_mode", v) for v in (0, 1, 2)] +
            [("ct_energy", v) for v in (0, 1, 2)])


@pytest.mark.parametrize("key,value", SWITCHES)
def test_switches_mhd(key, value):
    """Each Config.h switch value on 2D MHD (CTU + HLLD + CT): the five limiters,
    the CFL modes (max, add, power mean) and the three CT energy modes."""
    pb = tiny_mhd((9, 7), 1)
    pb[2][key] = value
    run_pair(pb, steps=4)


@pytest.mark.parametrize("key,value", [("limiter", v) for v in range(5)] + [("cfl_mode", v) for v in (0, 1, 2)] +
                         [("muscl", 0), ("ctu", 0)])
def test_switches_hydro(key, value):
    """The same on 2D hydro (CTU + HLLC), plus first order (no MUSCL, no CTU)
    and MUSCL without CTU."""
    if key == "muscl":
        pb = random_hydro((9, 7), 5, ctu=0, muscl=0)
    else:
        pb = random_hydro((9, 7), 5)
        pb[2][key] = value
    run_pair(pb, steps=4, nf=5)


def with_bc(pb, kind):
    faces = {1: 3, 2: 15, 3: 63}[len(pb[0])]
    return (pb[0], pb[1], pb[2], [(kind, faces)], pb[4], pb[5], pb[6])


@pytest.mark.parametrize("kind", [0, 2])
@pytest.mark.parametrize("case", ["mhd3", "mhd2", "hydro3"])
def test_steps_with_boundaries(case, kind):
    """Whole steps under outflow (0) and reflective (2) boundaries on every face
    (the configurations run periodic or outflow): ghost values feed the MUSCL
    stencils, the CTU corrections and the corner EMFs at the box edges."""
    pb = {"mhd3": lambda: tiny_mhd((6, 5, 7), 1), "mhd2": lambda: tiny_mhd((9, 7), 1),
          "hydro3": lambda: random_hydro((6, 5, 7), 5)}[case]()
    run_pair(with_bc(pb, kind), steps=3, nf=5 if case.startswith("hydro") else 8)


@pytest.mark.parametrize("cfl", [0.3, 1.2])
def test_mhd_double_rarefaction(cfl):
    """2D MHD with two strong rarefactions (v = -+3 about x = 1/2, p = 0.02):
    HLLD's physical-safety pivot to HLLE, and at CFL 1.2 rejected steps and
    dt halving, against the oracle."""
    g = P.ghosts_for(True)
    shape = (32, 4)
    w, B = P.empty_state(shape, g, True)
    wi = P.interior(w, g, shape)
    wi[..., 0] = 1.0
    wi[..., 1] = 0.02
    wi[:16, :, 2] = -3.0
    wi[16:, :, 2] = 3.0
    P.interior(B, g, shape, extra=1)[..., 0] = 0.05
    cfg = dict(mhd=1, gamma=5 / 3, solver=5, limiter=2, ctu=1, ct_energy=2, cfl_coeff=cfl)
    run_pair((shape, (1 / 32, 1 / 4), cfg, [(0, 15)], w, B, {"init_B": True}), steps=15)


def test_mhd_1d():
    """1D MHD (cell-centred B, no CT), MUSCL + HLLD, outflow."""
    g = P.ghosts_for(True, ctu=False)
    shape = (16,)
    w, B = P.empty_state(shape, g, True)
    rng = np.random.default_rng(9)
    wi = P.interior(w, g, shape)
    wi[..., 0] = 1 + 0.3 * rng.random(shape)
    wi[..., 1] = 1 + 0.3 * rng.random(shape)
    wi[..., 2:5] = 0.3 * (rng.random(shape + (3,)) - 0.5)
    wi[..., 5] = 0.5
    wi[..., 6] = 0.3 * rng.random(shape)
    wi[..., 7] = 0.2
    run_pair((shape, (1 / 16,), dict(mhd=1, gamma=5 / 3, solver=5, limiter=2, ctu=0), [(0, 3)], w, B, {}), steps=3)


@pytest.mark.parametrize("shape", [(1, 1), (2, 3), (1, 1, 1), (2, 1, 3)])
@pytest.mark.parametrize("bc", [0, 1])
def test_tiny_grids(shape, bc):
    """Grids of one to three cells per axis (fewer cells than ghost layers):
    every stencil, walker, wavefront and ghost pass at its smallest."""
    run_pair(tiny_mhd(shape, bc), steps=2)


@pytest.mark.parametrize("nz", [2, 3])
def test_thin_contiguous_axis(nz):
    """Ghost passes on the contiguous axis run one thread per ghost cell when
    the axis holds at least ng = 3 cells (every cell they read is then
    interior); below that the reference's in-order g loop, whose periodic
    images read ghosts written earlier in the same pass.  Both sides of the
    switch against the oracle."""
    run_pair(rect_mhd_3d((10, 12, nz)), steps=3)


@pytest.mark.parametrize("case,chunk", [("c2", 7), ("c3", 7), ("c4", 5), ("rect", 7), ("rect", 64)])
def test_windowed_walkers_exact(case, chunk, monkeypatch):
    """The FP32 build's strided-axis walkers keep the three MUSCL cells in
    registers and read the cell two ahead before each Riemann solve.  The same
    code instantiated for FP64 (DG_MARCH_WINDOW_CHECK) must return the oracle's
    bits: an exact check of the window's cell bookkeeping (first cell of a
    walker, outer ghost layers, partial last walkers) that the tolerance of the
    mixed form could not give.  Short walkers (DG_MARCH_CHUNK) put many walker
    starts and a partial walker on every line."""
    monkeypatch.setenv("DG_MARCH_WINDOW_CHECK", "1")
    monkeypatch.setenv("DG_MARCH_CHUNK", str(chunk))
    if case == "c2":
        run_pair(P.kelvin_helmholtz_2d(64), steps=20, nf=5)
    elif case == "c3":
        run_pair(P.orszag_tang_2d(48), steps=20)
    elif case == "c4":
        run_pair(P.turbulence_3d(16), steps=6)
    else:
        run_pair(rect_mhd_3d(), steps=4)


def fallback_case(name, n, dim, try_hlle):
    """Riemann problems run at CFL 1.5, where HLLC's fluxes fail the verify
    check of Flux.cpp:17-97 (L - F dt/dx or R + F dt/dx unphysical) on some
    faces: the HLLE / exact-solver fallback runs, and some attempts fail and
    halve dt.  Toro's test 3 (p 1000 | 0.01) and Einfeldt's 1-2-3 problem
    (v -/+2, p 0.4).  dim 2: the same data along x, periodic in y."""
    L, R = {"toro3": ((1.0, 1000.0, 0.0), (1.0, 0.01, 0.0)), "123": ((1.0, 0.4, -2.0), (1.0, 0.4, 2.0))}[name]
    shape = (n,) if dim == 1 else (n, 8)
    g = P.ghosts_for(False)
    w, _ = P.empty_state(shape, g, False)
    x = (np.arange(n) + 0.5) / n
    wi = P.interior(w, g, shape)
    for q in range(3):
        col = np.where(x < 0.5, L[q], R[q])
        wi[..., (0, 1, 2)[q]] = col if dim == 1 else col[:, None]
    cfg = dict(mhd=0, gamma=5.0 / 3.0, solver=5, limiter=0, ctu=1, verify_fallback=1, fallback_try_hlle=int(try_hlle),
               cfl_coeff=1.5)
    bcs = [] if dim == 1 else [(0, 3), (1, 12)]
    d = (1.0 / n,) if dim == 1 else (1.0 / n, 1.0 / n)
    return shape, d, cfg, bcs, w, None, {}


@pytest.mark.parametrize("name,dim,try_hlle", [("toro3", 1, False), ("toro3", 1, True), ("123", 1, False), ("toro3", 2, False),
                                               ("123", 2, True)])
def test_verify_fallback_paths(name, dim, try_hlle):
    """The verify-fallback chain (HLLC -> [HLLE] -> exact, Flux.cpp:17-97) and
    the dt halving of rejected attempts on the grid, bit for bit with the
    oracle, 1D and 2D CTU, with and without the HLLE stage."""
    res = run_pair(fallback_case(name, 200, dim, try_hlle), steps=40, nf=5)
    assert res["bit_equal"]
