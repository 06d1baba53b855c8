"""Known-answer tests of the oracle, ported from the reference's own test
suite (values and tolerances as there: approx = rel 1e-15, abs 1e-18,
Testing/Testing.hpp:22-26)."""
import math

import numpy as np
import pytest

import oracle as O

G = 5.0 / 3.0
H = O.Cfg(mhd=0)
M = O.Cfg(mhd=1)


def approx(a, b, rel=1e-15, ab=1e-18):
    return abs(a - b) <= max(ab, rel * max(abs(a), abs(b)))


def close(a, b, rel=1e-15, ab=1e-18):
    return all(approx(x, y, rel, ab) for x, y in zip(a, b))


def st(rho, vx, vy, vz, p, Bx=0.0, By=0.0, Bz=0.0):
    return O.s8(rho, p, vx, vy, vz, Bx, By, Bz)


# ---------------------------------------------- FluidElement_Tests.cpp:137-219
def test_conversion():
    W = st(2.0, 3.0, 4.0, 0.0, 10.0)
    U = O.to_cons(H, W)
    assert approx(U[0], 2.0) and approx(U[2], 6.0) and approx(U[3], 8.0) and approx(U[4], 0.0)
    assert approx(U[1], 40.0)
    assert close(O.to_prim(H, U), W)
    W = st(2.0, 3.0, 4.0, 0.0, 10.0, 1.0, 2.0, 3.0)
    U = O.to_cons(M, W)
    assert approx(U[1], 40.0 + 14.0 / (8 * math.pi))
    assert close(O.to_prim(M, U), W)


def test_flux():
    W = st(2.0, 3.0, 4.0, 5.0, 10.0)
    F = O.flux_prim(H, W)
    U = O.to_cons(H, W)
    assert approx(F[0], 6.0) and approx(F[2], 28.0) and approx(F[3], 24.0) and approx(F[4], 30.0)
    assert approx(F[1], 3.0 * (U[1] + 10.0))
    W = st(2.0, 3.0, 4.0, 5.0, 10.0, 1.0, 2.0, 3.0)
    F = O.flux_prim(M, W)
    assert approx(F[0], 6.0)
    assert approx(F[2], 28.0 + 12 / (8 * math.pi))
    assert approx(F[3], 24.0 - 2 / (4 * math.pi))
    assert approx(F[4], 30.0 - 3 / (4 * math.pi))
    assert F[5] == 0.0
    assert approx(F[6], 3.0 * 2.0 - 4.0 * 1.0) and approx(F[7], 3.0 * 3.0 - 5.0 * 1.0)


def test_enthalpy_and_speeds():
    W = st(2.0, 3.0, 4.0, 0.0, 10.0)
    assert approx(O.speeds(H, W)[4], 25.0)
    assert approx(O.speeds(H, W)[0], math.sqrt(G * 5.0))
    W = st(2.0, 3.0, 4.0, 0.0, 10.0, 1, 2, 3)
    s = O.speeds(M, W)
    assert approx(s[4], 25.0 + 14.0 / (8 * math.pi))
    assert approx(s[1], math.sqrt(14 / (8 * math.pi)))
    c2, a2, ax2 = G * 5.0, 14 / (8 * math.pi), 1 / (8 * math.pi)
    assert approx(s[3], math.sqrt((c2 + a2) / 2 + math.sqrt((c2 + a2) ** 2 / 4 - ax2 * c2)))


def test_physicality():  # FluidElement_Tests.cpp:221-271
    assert O.is_physical(H, st(1, 0, 0, 0, 1))
    assert not O.is_physical(H, st(0, 0, 0, 0, 1))
    assert not O.is_physical(H, st(1, 0, 0, 0, -1))
    assert not O.is_physical(H, st(1, math.nan, 0, 0, 1))
    assert not O.is_physical(M, st(1, 0, 0, 0, 1, math.inf, 0, 0))


# ------------------------------------------------- Riemann_Exact_Tests.cpp
def test_exact_sod_star_state():
    L, R = st(1.0, 0, 0, 0, 1.0), st(0.125, 0, 0, 0, 0.1)
    p, u, rl, rr, trrs = O.exact_star(H, L, R)
    assert approx(p, 0.29394518766601785) and approx(u, 0.8411948521688083)
    assert approx(rl, 0.4796890587209175) and approx(rr, 0.22980574931194703)
    W0 = O.exact_sample(H, L, R, 0.0)
    assert approx(W0[0], rl) and approx(W0[2], u) and approx(W0[1], p)
    assert close(O.exact_sample(H, L, R, -100.0), L) and close(O.exact_sample(H, L, R, 100.0), R)
    fan = O.exact_sample(H, L, R, -0.5)
    assert rl < fan[0] < 1.0 and p < fan[1] < 1.0 and 0.0 < fan[2] < u


def test_exact_supersonic_upwind():
    for v, side in ((10.0, 0), (-10.0, 1)):
        L, R = st(1.0, v, 2.0, -3.0, 1.0), st(0.125, v, -8.0, 9.0, 0.1)
        F = O.riemann_solver(H, L, R, "EXACT")
        assert close(F, O.flux_prim(H, L if side == 0 else R))


# -------------------------------------------------------- HLL_Tests.cpp
def test_hll_manual_wave_speeds():
    L, R = st(1.0, 0, 0, 0, 1.0), st(0.125, 0, 0, 0, 0.1)
    UL, UR = O.to_cons(H, L), O.to_cons(H, R)
    FL, FR = O.flux_prim(H, L), O.flux_prim(H, R)
    SL, SR = -1.0, 2.0
    exp = (SR * FL - SL * FR + SL * SR * (UR - UL)) / (SR - SL)
    assert close(O.riemann_solver(H, L, R, "HLL_S", SL, SR), exp)
    assert close(O.riemann_solver(H, L, R, "HLL_S", 0.1, 2.0), FL)
    assert close(O.riemann_solver(H, L, R, "HLL_S", -2.0, -0.1), FR)
    assert close(O.riemann_solver(H, L, R, "HLLC_S", 0.1, 2.0), FL)
    assert close(O.riemann_solver(H, L, R, "HLLC_S", -2.0, -0.1), FR)
    assert np.all(np.isfinite(O.riemann_solver(H, L, R, "HLLC_S", -1.0, 2.0)))


def test_equal_state_and_stationary_contact():
    W = st(1.0, 0.75, 0.2, -0.1, 1.0)
    for s in ("HLL", "HLLC", "HLLE"):
        assert close(O.riemann_solver(H, W, W, s), O.flux_prim(H, W))
    Wm = st(1.0, 0.75, 0.2, -0.1, 1.0, 0.4, -0.3, 0.2)
    assert close(O.riemann_solver(M, Wm, Wm, "HLLD"), O.flux_prim(M, Wm))
    L, R = st(1.0, 0, 0, 0, 1.0), st(2.0, 0, 0, 0, 1.0)
    exp = np.array([0, 0, 1.0, 0, 0, 0, 0, 0])
    assert close(O.riemann_solver(H, L, R, "HLLC"), exp)
    assert close(O.riemann_solver(M, L, R, "HLLD"), exp)


def test_supersonic_upwind_all_solvers():
    for v, side in ((10.0, 0), (-10.0, 1)):
        L, R = st(1.0, v, 0, 0, 1.0), st(0.125, v, 0, 0, 0.1)
        for s in ("HLL", "HLLC", "HLLE"):
            assert close(O.riemann_solver(H, L, R, s), O.flux_prim(H, L if side == 0 else R))
        L = st(1.0, v, 0.5, -0.3, 1.0, 0.4, -0.3, 0.2)
        R = st(0.125, v, -0.2, 0.4, 0.1, 0.4, 0.1, -0.5)
        assert close(O.riemann_solver(M, L, R, "HLLD"), O.flux_prim(M, L if side == 0 else R))


@pytest.mark.parametrize("pair", [
    (st(1.0, -0.7, 0.6, -0.2, 1.0, 0.4, 0.3, -0.2), st(0.8, 0.5, -0.4, 0.3, 0.7, 0.4, -0.2, 0.5)),
    (st(1.0, -0.7, 0.6, -0.2, 1.0, -0.4, 0.3, -0.2), st(0.8, 0.5, -0.4, 0.3, 0.7, -0.4, -0.2, 0.5)),
    (st(1.0, -1.2, 0.1, 0.5, 0.6, 0.0, 0.7, -0.4), st(1.2, 1.0, -0.3, 0.2, 0.9, 0.0, -0.5, 0.3)),
    (st(5.99924, 19.5975, 0.5, -0.2, 460.894, 1.0, 0.2, -0.6), st(5.99242, -6.19633, -0.4, 0.3, 46.0950, 1.0, -0.3, 0.4)),
])
def test_hlld_star_regions_finite(pair):  # HLL_Tests.cpp:241-260
    L, R = pair
    F = O.riemann_solver(M, L, R, "HLLD")
    assert np.all(np.isfinite(F))
    assert F[5] == 0.0
    UL, UR = O.to_cons(M, L), O.to_cons(M, R)
    for U in (UL - F * 0.001, UR + F * 0.001):
        W = O.to_prim(M, U)
        assert W[0] > 0 and W[1] > 0


def test_verify_fallback_safety_check():
    """Riemann_Tests.cpp:150-180: a physical flux passes the check unchanged;
    for L/R = (rho 1, vx -/+3, p 0.4) at dt/dx = 1 an unphysical HLLC update
    pivots to HLLE or the exact flux."""
    W = st(1.0, 0.2, -0.3, 0.4, 1.0)
    c = O.Cfg(mhd=0, solver=O.HLL)
    assert close(O.riemann_flux(c, W, W, 0.01), O.riemann_flux(c, W, W, 0.0))
    L, R = st(1.0, -3.0, 0.2, -0.1, 0.4), st(1.0, 3.0, -0.3, 0.4, 0.4)
    sel = O.riemann_flux(H, L, R, 0.0)
    fb = O.riemann_flux(H, L, R, 1.0)
    ok = lambda U: O.to_prim(H, U)[0] > 0 and O.to_prim(H, U)[1] > 0  # noqa: E731
    if not (ok(O.to_cons(H, L) - sel) and ok(O.to_cons(H, R) + sel)):
        assert np.array_equal(fb, O.riemann_solver(H, L, R, "HLLE")) or np.array_equal(fb, O.riemann_solver(H, L, R, "EXACT"))
    assert np.all(np.isfinite(fb))


# -------------------------------------------------------- TVD_Tests.cpp
@pytest.mark.parametrize("lim,cases", [
    (O.MINMOD, [((2.0, 3.0), 2.0), ((3.0, 2.0), 2.0), ((-2.0, -3.0), -2.0), ((-2.0, 3.0), 0.0)]),
    (O.MC, [((1.0, 3.0), 2.0), ((3.0, 1.0), 2.0), ((-1.0, -3.0), -2.0), ((1.0, -3.0), 0.0)]),
    (O.VANLEER, [((2.0, 4.0), 8.0 / 3.0), ((-2.0, -4.0), -8.0 / 3.0), ((2.0, -4.0), 0.0)]),
    (O.SUPERBEE, [((1.0, 4.0), 2.0), ((4.0, 1.0), 2.0), ((-1.0, -4.0), -2.0), ((1.0, -4.0), 0.0)]),
    (O.VANALBADA, [((2.0, 4.0), 12.0 / 5.0), ((-2.0, -4.0), -12.0 / 5.0), ((2.0, -4.0), 0.0)]),
])
def test_scalar_limiters(lim, cases):
    c = O.Cfg(limiter=lim)
    for (a, b), e in cases:
        assert approx(O.limit(c, a, b), e)


def test_minmod_identity():  # TVD_Tests.cpp:71-76
    rng = np.random.default_rng(0)
    c = O.Cfg(limiter=O.MINMOD)
    for _ in range(100):
        a, b = rng.integers(-1000000, 1000001, 2) * 1e-3
        cc = -(a + b)
        assert approx(O.limit(c, a, -b) + O.limit(c, a, -cc), a)


def test_muscl_closed_forms():
    c = O.Cfg(mhd=0, limiter=O.MINMOD)
    W = st(1.0, 0.2, -0.3, 0.4, 2.0)
    Lo, Ro = O.muscl(c, W, W, W, 0.25)
    assert close(Lo, W) and close(Ro, W)
    Lo, Ro = O.muscl(c, st(1, 1, 2, 3, 4), st(2, 2, 3, 4, 5), st(3, 3, 4, 5, 6), 0.0)
    assert close(Lo, st(1.5, 1.5, 2.5, 3.5, 4.5)) and close(Ro, st(2.5, 2.5, 3.5, 4.5, 5.5))
    dt = 0.1
    Lo, Ro = O.muscl(c, st(1, 0, 0, 0, 1), st(2, 0, 0, 0, 2), st(3, 0, 0, 0, 3), dt)
    UL, UR = O.to_cons(c, st(1.5, 0, 0, 0, 1.5)), O.to_cons(c, st(2.5, 0, 0, 0, 2.5))
    corr = (O.flux_prim(c, st(2.5, 0, 0, 0, 2.5)) - O.flux_prim(c, st(1.5, 0, 0, 0, 1.5))) * (0.5 * dt)
    assert close(Lo, O.to_prim(c, UL - corr)) and close(Ro, O.to_prim(c, UR - corr))


def test_muscl_first_order_fallback():
    """An unphysical half-step state falls back to first order (MUSCL.cpp:56-60)."""
    c = O.Cfg(mhd=0, limiter=O.SUPERBEE)
    wC = st(1.0, 0, 0, 0, 1e-3)
    Lo, Ro = O.muscl(c, st(1.0, -5, 0, 0, 1e-3), wC, st(1.0, 5, 0, 0, 1e-3), 0.9)
    assert close(Lo, wC) and close(Ro, wC)


# -------------------------------------------------------- CFL_Tests.cpp
def test_cfl_speeds():
    W = st(2.0, 3.0, -4.0, 5.0, 5.0)
    a = math.sqrt(G * 5.0 / 2.0)
    sx, sy, sz = (3 + a) / 0.25, (4 + a) / 0.10, (5 + a) / 0.50
    assert approx(O.cfl_speed(O.Cfg(mhd=0, cfl_mode=O.CFL_ADD), W, 0.25, 0.10, 0.50), sx + sy + sz)
    assert approx(O.cfl_speed(O.Cfg(mhd=0, cfl_mode=O.CFL_MAX), W, 0.25, 0.10, 0.50), max(sx, sy, sz))
    Wm = st(2.0, 3.0, -4.0, 5.0, 5.0, 2.0, -1.0, 3.0)
    am = O.speeds(M, Wm)[3]
    sx, sy, sz = (3 + am) / 0.25, (4 + am) / 0.10, (5 + am) / 0.50
    assert approx(O.cfl_speed(O.Cfg(mhd=1, cfl_mode=2), Wm, 0.25, 0.10, 0.50), math.sqrt(sx * sx + sy * sy + sz * sz), 1e-14)


def test_cfl_time_1d_ignores_outer_ghosts():  # CFL_Tests.cpp:244-257
    g = O.Grid((5,), (0.25,), O.Cfg(mhd=0), ghosts=2)
    w = g.get_w()
    w[:] = st(1.0, 0.1, 0, 0, 1.0)
    w[0] = st(1.0, 1000.0, 0, 0, 1.0)
    w[-1] = st(1.0, 1000.0, 0, 0, 1.0)
    g.set_w(w)
    assert approx(g.cfl(), 0.3 * 0.25 / (0.1 + math.sqrt(G)))
    w[1] = st(1.0, 9.0, 0, 0, 1.0)  # ghost layer 1 counts
    g.set_w(w)
    assert approx(g.cfl(), 0.3 * 0.25 / (9.0 + math.sqrt(G)))


def test_cfl_time_3d_mhd():  # CFL_Tests.cpp:405-431
    g = O.Grid((3, 4, 5), (0.5, 0.25, 5.0), O.Cfg(mhd=1, ctu=0, muscl=0), ghosts=1)
    w = g.get_w()
    Wm = st(1.0, 2.0, 3.0, 4.0, 1.0, 1.0, -2.0, 0.5)
    w[:] = Wm
    g.set_w(w)
    a = O.speeds(M, Wm)[3]
    speed = (2 + a) / 0.5 + (3 + a) / 0.25 + (4 + a) / 5.0
    assert approx(g.cfl(), 0.3 / speed)


def test_cfl_zero_speed_raises():
    g = O.Grid((4, 4), (0.1, 0.1), O.Cfg(mhd=0), ghosts=2)
    with pytest.raises(O.OracleError, match="max signal speed is zero"):
        g.cfl()


# -------------------------------------------------------- face masks
def test_face_mask():  # GhostFill.cpp:14-49
    assert O.face_mask("X") == O.FX and O.face_mask("XY+") == O.FX | O.YP
    assert O.face_mask("LR") == O.FX and O.face_mask("UD") == O.FY and O.face_mask("z-") == O.ZN
