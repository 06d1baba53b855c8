"""The oracle's GhostFill boundaries against the reference's own boundary
tests (Testing/Core/Boundary/*.cpp), restated case by case: tagged interior
states, a tagged "garbage" state G in the ghosts, one boundary applied, then
the same assertions on cells and face B.

A single element is applied alone (``Outflow(X).apply(grid)`` in the
reference) by covering every other face with Ignore, which switches off the
BoundaryList's implicit outflow (BoundaryList.hpp:104-110, Ignore.cpp:12-18).
"""
import numpy as np
import pytest

import oracle as O

XN, XP, YN, YP, ZN, ZP = 1, 2, 4, 8, 16, 32
X, Y, Z = 3, 12, 48


def tagged(tag):
    """make_tagged_state (Grid_Tests.cpp:31-37): make_state(rho, vx, vy, vz, p)."""
    return O.s8(tag, tag + 40, tag + 10, tag + 20, tag + 30, tag + 50, tag + 60, tag + 70)


def tagged_vec(tag):
    return np.array([tag + 100, tag + 200, tag + 300], dtype=float)


G = tagged(-666)
W_FIX = tagged(2112)


def reflect(w, ax, mhd=True):
    """reflectX/Y/Z (Reflecting_Tests.cpp): flip v_ax and the transverse cell B."""
    w = w.copy()
    w[2 + ax] *= -1
    if mhd:
        for q in range(3):
            if q != ax:
                w[5 + q] *= -1
    return w


def close(a, b):
    return np.allclose(a, b, rtol=1e-13, atol=1e-13)


class TGrid:
    """fill_1D/2D/3D (Boundary_Tests.cpp:53-110) on an oracle grid."""

    def __init__(self, shape, ghosts, mhd=1):
        self.shape, self.mhd = shape, mhd
        self.dim = len(shape)
        self.G = O.Grid(shape, (1.0,) * self.dim, O.Cfg(mhd=mhd), ghosts=ghosts)
        self.g = self.G.g
        self.fill()

    def fill(self):
        g, n = self.g, self.shape
        w = self.G.get_w()
        w[:] = G
        s = 1.0 / max(n)
        for idx in np.ndindex(*n):
            if self.dim == 1:
                tag = idx[0] * s
            elif self.dim == 2:
                tag = idx[0] * s + idx[1] * s * s
            else:
                tag = idx[0] * s + idx[1] * s * s + idx[2] * s ** 3 + 1
            w[tuple(g + i for i in idx)] = tagged(tag)
        self.G.set_w(w)
        if self.mhd and self.dim >= 2:
            B = self.G.get_B()
            for idx in np.ndindex(*B.shape[:-1]):
                ii = [q - g for q in idx]
                tag = 10 * ii[0] + ii[1] if self.dim == 2 else 100 * ii[0] + 10 * ii[1] + ii[2]
                B[idx] = tagged_vec(tag)
            self.G.set_B(B)
        self.load()

    def load(self):
        self.W = self.G.get_w()
        self.B = self.G.get_B() if (self.mhd and self.dim >= 2) else None

    def apply(self, *elems, alone=True):
        """elems: (kind, faces, corners, gated, state). alone: no implicit outflow."""
        self.G.clear_bc()
        covered = 0
        for kind, faces, *_ in elems:
            covered |= faces
            if kind == O.PERIODIC:
                covered |= (X if faces & X else 0) | (Y if faces & Y else 0) | (Z if faces & Z else 0)
        if alone and (63 & ~covered):
            self.G.add_bc(O.IGNORE, 63 & ~covered)
        for kind, faces, corners, gated, state in elems:
            self.G.add_bc(kind, faces, corners, gated, state)
        self.G.apply_bc()
        self.load()

    def w(self, *idx):
        return self.W[tuple(self.g + i for i in idx)]

    def b(self, *idx):
        return self.B[tuple(self.g + i for i in idx)]


def E(kind, faces, corners=True, gated=False, state=None):
    return (kind, faces, corners, gated, state)


# ------------------------------------------------------------------ outflow
@pytest.mark.parametrize("mhd", [0, 1])
def test_outflow_1d(mhd):  # Outflow_Tests.cpp:51-77
    t = TGrid((4,), 2, mhd)
    t.apply(E(O.OUTFLOW, X))
    assert close(t.w(-1), t.w(0)) and close(t.w(-2), t.w(0))
    assert close(t.w(4), t.w(3)) and close(t.w(5), t.w(3))
    t.fill()
    t.apply(E(O.OUTFLOW, XN))
    assert close(t.w(-1), t.w(0)) and close(t.w(-2), t.w(0))
    assert close(t.w(4), G) and close(t.w(5), G)
    t.fill()
    t.apply(E(O.OUTFLOW, XP))
    assert close(t.w(4), t.w(3)) and close(t.w(5), t.w(3))
    assert close(t.w(-1), G) and close(t.w(-2), G)


@pytest.mark.parametrize("mhd", [0, 1])
def test_outflow_2d(mhd):  # Outflow_Tests.cpp:79-110
    t = TGrid((3, 4), 1, mhd)
    t.apply(E(O.OUTFLOW, X))
    for j in range(4):
        assert close(t.w(-1, j), t.w(0, j)) and close(t.w(3, j), t.w(2, j))
        if mhd:
            assert t.b(-1, j)[1] == t.b(0, j)[1] and t.b(4, j)[1] == t.b(3, j)[1]
    t.fill()
    t.apply(E(O.OUTFLOW, Y, corners=False))
    for i in range(3):
        assert close(t.w(i, -1), t.w(i, 0)) and close(t.w(i, 4), t.w(i, 3))
        if mhd:
            assert t.b(i, -1)[0] == t.b(i, 0)[0] and t.b(i, 5)[0] == t.b(i, 4)[0]
    assert close(t.w(-1, -1), G)
    t.fill()
    t.apply(E(O.OUTFLOW, X | Y))
    assert close(t.w(-1, -1), t.w(0, 0))


@pytest.mark.parametrize("mhd", [0, 1])
def test_outflow_3d(mhd):  # Outflow_Tests.cpp:112-180
    t = TGrid((3, 4, 5), 1, mhd)
    t.apply(E(O.OUTFLOW, X))
    for j in range(4):
        for k in range(5):
            assert close(t.w(-1, j, k), t.w(0, j, k)) and close(t.w(3, j, k), t.w(2, j, k))
            if mhd:
                assert t.b(-1, j, k)[1] == t.b(0, j, k)[1] and t.b(-1, j, k)[2] == t.b(0, j, k)[2]
                assert t.b(3, j, k)[1] == t.b(2, j, k)[1] and t.b(4, j, k)[2] == t.b(3, j, k)[2]
    t.fill()
    t.apply(E(O.OUTFLOW, YN))
    for i in range(3):
        for k in range(5):
            assert close(t.w(i, -1, k), t.w(i, 0, k)) and close(t.w(i, 4, k), G)
            if mhd:
                assert t.b(i, -1, k)[0] == t.b(i, 0, k)[0] and t.b(i, -1, k)[2] == t.b(i, 0, k)[2]
    t.fill()
    t.apply(E(O.OUTFLOW, YP))
    for i in range(3):
        for k in range(5):
            assert close(t.w(i, -1, k), G) and close(t.w(i, 4, k), t.w(i, 3, k))
            if mhd:
                assert t.b(i, 4, k)[0] == t.b(i, 3, k)[0] and t.b(i, 4, k)[2] == t.b(i, 3, k)[2]
    t.fill()
    t.apply(E(O.OUTFLOW, Z))
    for i in range(3):
        for j in range(4):
            assert close(t.w(i, j, -1), t.w(i, j, 0)) and close(t.w(i, j, 5), t.w(i, j, 4))
            if mhd:
                assert t.b(i, j, -1)[0] == t.b(i, j, 0)[0] and t.b(i, j, -1)[1] == t.b(i, j, 0)[1]
                assert t.b(i, j, 5)[0] == t.b(i, j, 4)[0] and t.b(i, j, 5)[1] == t.b(i, j, 4)[1]
    t.apply(E(O.OUTFLOW, Z, corners=False))
    assert close(t.w(-1, 1, -1), G)
    t.fill()
    t.apply(E(O.OUTFLOW, X | Y | Z))
    assert close(t.w(-1, -1, -1), t.w(0, 0, 0))


# ------------------------------------------------------------ gated outflow
@pytest.mark.parametrize("mhd", [0, 1])
def test_outflow_gated_1d(mhd):  # OutflowGate_Tests.cpp: 1D
    t = TGrid((4,), 2, mhd)

    def run(cell, v, faces):
        t.fill()
        t.W[t.g + cell, 2] = v
        t.G.set_w(t.W)
        t.apply(E(O.OUTFLOW, faces, gated=True))

    run(0, 5.0, XN)
    assert t.w(-1)[2] == 0.0 and t.w(-2)[2] == 0.0
    blocked = t.w(-1).copy()
    blocked[2] = t.w(0)[2]
    assert close(blocked, t.w(0))
    run(0, -5.0, XN)
    assert t.w(-1)[2] == -5.0 and t.w(-2)[2] == -5.0
    run(3, -5.0, XP)
    assert t.w(4)[2] == 0.0 and t.w(5)[2] == 0.0
    run(3, 5.0, XP)
    assert t.w(4)[2] == 5.0 and t.w(5)[2] == 5.0


@pytest.mark.parametrize("mhd", [0, 1])
@pytest.mark.parametrize("ax", [0, 1])
def test_outflow_gated_2d(mhd, ax):  # OutflowGate_Tests.cpp: 2D X and Y
    t = TGrid((3, 4), 2, mhd)
    n = (3, 4)
    neg, pos = (XN, XP) if ax == 0 else (YN, YP)
    other = n[1 - ax]

    def cell(a, b):  # a along ax, b transverse
        return (a, b) if ax == 0 else (b, a)

    def run(edge, v, faces):
        t.fill()
        for b in range(other):
            t.W[tuple(t.g + q for q in cell(edge, b)) + (2 + ax,)] = v
        t.G.set_w(t.W)
        t.apply(E(O.OUTFLOW, faces, gated=True))

    run(0, 5.0, neg)
    for b in range(other):
        assert t.w(*cell(-1, b))[2 + ax] == 0 and t.w(*cell(-2, b))[2 + ax] == 0
        blocked = t.w(*cell(-1, b)).copy()
        blocked[2 + ax] = t.w(*cell(0, b))[2 + ax]
        assert close(blocked, t.w(*cell(0, b)))
    run(0, -5.0, neg)
    for b in range(other):
        assert t.w(*cell(-1, b))[2 + ax] == -5 and t.w(*cell(-2, b))[2 + ax] == -5
    last = n[ax] - 1
    run(last, -5.0, pos)
    for b in range(other):
        assert t.w(*cell(last + 1, b))[2 + ax] == 0 and t.w(*cell(last + 2, b))[2 + ax] == 0
    run(last, 5.0, pos)
    for b in range(other):
        assert t.w(*cell(last + 1, b))[2 + ax] == 5 and t.w(*cell(last + 2, b))[2 + ax] == 5


# ----------------------------------------------------------------- periodic
@pytest.mark.parametrize("mhd", [0, 1])
def test_periodic_1d(mhd):  # Periodic_Tests.cpp:26-34
    t = TGrid((4,), 2, mhd)
    t.apply(E(O.PERIODIC, X))
    assert close(t.w(-1), t.w(3)) and close(t.w(-2), t.w(2))
    assert close(t.w(4), t.w(0)) and close(t.w(5), t.w(1))


@pytest.mark.parametrize("mhd", [0, 1])
def test_periodic_2d(mhd):  # Periodic_Tests.cpp:35-84
    t = TGrid((3, 4), 2, mhd)
    t.apply(E(O.PERIODIC, X))
    for j in range(4):
        assert close(t.w(-1, j), t.w(2, j)) and close(t.w(3, j), t.w(0, j))
        if mhd:
            assert close(t.b(-1, j), t.b(2, j)) and close(t.b(3, j), t.b(0, j)) and close(t.b(4, j), t.b(1, j))
    t.fill()
    t.apply(E(O.PERIODIC, Y))
    for i in range(3):
        assert close(t.w(i, -1), t.w(i, 3)) and close(t.w(i, 4), t.w(i, 0))
        if mhd:
            assert close(t.b(i, -1), t.b(i, 3)) and close(t.b(i, 4), t.b(i, 0)) and close(t.b(i, 5), t.b(i, 1))
    t.fill()
    t.apply(E(O.PERIODIC, Y, corners=False))
    assert close(t.w(-1, -1), G)
    t.fill()
    t.apply(E(O.PERIODIC, X | Y))
    assert close(t.w(-1, -1), t.w(2, 3))


@pytest.mark.parametrize("mhd", [0, 1])
def test_periodic_3d(mhd):  # Periodic_Tests.cpp:86-163
    t = TGrid((3, 4, 5), 1, mhd)
    t.apply(E(O.PERIODIC, X))
    for j in range(4):
        for k in range(5):
            assert close(t.w(-1, j, k), t.w(2, j, k)) and close(t.w(3, j, k), t.w(0, j, k))
            if mhd:
                assert close(t.b(-1, j, k), t.b(2, j, k)) and close(t.b(3, j, k), t.b(0, j, k))
                assert close(t.b(4, j, k), t.b(1, j, k))
    t.fill()
    t.apply(E(O.PERIODIC, X))
    assert close(t.w(-1, 1, -1), G)
    t.fill()
    t.apply(E(O.PERIODIC, Y))
    for i in range(3):
        for k in range(5):
            assert close(t.w(i, -1, k), t.w(i, 3, k)) and close(t.w(i, 4, k), t.w(i, 0, k))
            if mhd:
                assert close(t.b(i, -1, k), t.b(i, 3, k)) and close(t.b(i, 4, k), t.b(i, 0, k))
                assert close(t.b(i, 5, k), t.b(i, 1, k))
    t.fill()
    t.apply(E(O.PERIODIC, Z))
    for i in range(3):
        for j in range(4):
            assert close(t.w(i, j, -1), t.w(i, j, 4)) and close(t.w(i, j, 5), t.w(i, j, 0))
            if mhd:
                assert close(t.b(i, j, -1), t.b(i, j, 4)) and close(t.b(i, j, 5), t.b(i, j, 0))
                assert close(t.b(i, j, 6), t.b(i, j, 1))
    t.fill()
    t.apply(E(O.PERIODIC, X | Y | Z))
    assert close(t.w(-1, -1, -1), t.w(2, 3, 4))


# --------------------------------------------------------------- reflective
@pytest.mark.parametrize("mhd", [0, 1])
def test_reflective_1d(mhd):  # Reflecting_Tests.cpp:38-73
    t = TGrid((4,), 2, mhd)
    for faces in (X, XN, XP):
        t.fill()
        t.apply(E(O.REFLECTIVE, faces))
        if faces & XN:
            assert close(t.w(-1), reflect(t.w(0), 0, mhd)) and close(t.w(-2), reflect(t.w(1), 0, mhd))
        if faces & XP:
            assert close(t.w(4), reflect(t.w(3), 0, mhd)) and close(t.w(5), reflect(t.w(2), 0, mhd))


@pytest.mark.parametrize("mhd", [0, 1])
def test_reflective_2d(mhd):  # Reflecting_Tests.cpp:88-134
    t = TGrid((3, 4), 1, mhd)
    t.apply(E(O.REFLECTIVE, X, corners=False))
    for j in range(4):
        assert close(t.w(-1, j), reflect(t.w(0, j), 0, mhd)) and close(t.w(3, j), reflect(t.w(2, j), 0, mhd))
        if mhd:
            assert t.b(-1, j)[0] == t.b(1, j)[0] and t.b(-1, j)[1] == -t.b(0, j)[1]
            assert t.b(4, j)[0] == t.b(2, j)[0] and t.b(3, j)[1] == -t.b(2, j)[1]
    assert close(t.w(-1, -1), G)
    t.fill()
    t.apply(E(O.REFLECTIVE, Y))
    for i in range(3):
        assert close(t.w(i, -1), reflect(t.w(i, 0), 1, mhd)) and close(t.w(i, 4), reflect(t.w(i, 3), 1, mhd))
        if mhd:
            assert t.b(i, -1)[0] == -t.b(i, 0)[0] and t.b(i, -1)[1] == t.b(i, 1)[1]
            assert t.b(i, 4)[0] == -t.b(i, 3)[0] and t.b(i, 5)[1] == t.b(i, 3)[1]
    t.fill()
    t.apply(E(O.REFLECTIVE, X | Y))
    assert close(t.w(-1, -1), reflect(reflect(t.w(0, 0), 0, mhd), 1, mhd))


@pytest.mark.parametrize("mhd", [0, 1])
def test_reflective_3d(mhd):  # Reflecting_Tests.cpp:155-247
    t = TGrid((3, 4, 5), 1, mhd)
    n = (3, 4, 5)
    for ax, corners in ((0, True), (1, True), (2, False)):
        t.fill()
        t.apply(E(O.REFLECTIVE, 3 << (2 * ax), corners=corners))
        for idx in np.ndindex(*[n[d] for d in range(3) if d != ax]):
            def c(a):
                li = list(idx)
                li.insert(ax, a)
                return tuple(li)
            assert close(t.w(*c(-1)), reflect(t.w(*c(0)), ax, mhd))
            assert close(t.w(*c(n[ax])), reflect(t.w(*c(n[ax] - 1)), ax, mhd))
            if mhd:
                bm, b0, b1 = t.b(*c(-1)), t.b(*c(0)), t.b(*c(1))
                assert bm[ax] == b1[ax]
                assert all(bm[q] == -b0[q] for q in range(3) if q != ax)
                bg, bl, bg2 = t.b(*c(n[ax])), t.b(*c(n[ax] - 1)), t.b(*c(n[ax] + 1))
                assert bg2[ax] == bl[ax]
                assert all(bg[q] == -bl[q] for q in range(3) if q != ax)


# -------------------------------------------------------------------- fixed
@pytest.mark.parametrize("mhd", [0, 1])
def test_fixed_1d(mhd):  # FixedBound_Tests.cpp:27-56
    t = TGrid((4,), 2, mhd)
    W = W_FIX if mhd else np.concatenate([W_FIX[:5], np.zeros(3)])
    t.apply(E(O.FIXED, X, state=W_FIX))
    for c in (-1, -2, 4, 5):
        assert close(t.w(c)[: 8 if mhd else 5], W[: 8 if mhd else 5])
    t.fill()
    t.apply(E(O.FIXED, XN, state=W_FIX))
    assert close(t.w(-1)[:5], W[:5]) and close(t.w(4), G)
    t.fill()
    t.apply(E(O.FIXED, XP, state=W_FIX))
    assert close(t.w(-1), G) and close(t.w(5)[:5], W[:5])


@pytest.mark.parametrize("mhd", [0, 1])
def test_fixed_2d_3d(mhd):  # FixedBound_Tests.cpp:58-162
    nf = 8 if mhd else 5
    t = TGrid((3, 4), 1, mhd)
    t.apply(E(O.FIXED, X, corners=False, state=W_FIX))
    for j in range(4):
        assert close(t.w(-1, j)[:nf], W_FIX[:nf]) and close(t.w(3, j)[:nf], W_FIX[:nf])
        if mhd:
            assert t.b(-1, j)[1] == W_FIX[6] and t.b(3, j)[1] == W_FIX[6]
    assert close(t.w(-1, -1), G)
    t.fill()
    t.apply(E(O.FIXED, Y, state=W_FIX))
    for i in range(3):
        assert close(t.w(i, -1)[:nf], W_FIX[:nf]) and close(t.w(i, 4)[:nf], W_FIX[:nf])
        if mhd:
            assert t.b(i, -1)[0] == W_FIX[5] and t.b(i, 4)[0] == W_FIX[5]
    t.fill()
    t.apply(E(O.FIXED, X | Y, state=W_FIX))
    assert close(t.w(-1, -1)[:nf], W_FIX[:nf])
    t = TGrid((3, 4, 5), 1, mhd)
    t.apply(E(O.FIXED, Y, corners=False, state=W_FIX))
    for i in range(3):
        for k in range(5):
            assert close(t.w(i, -1, k)[:nf], W_FIX[:nf]) and close(t.w(i, 4, k)[:nf], W_FIX[:nf])
            if mhd and k + 1 < 5:
                assert t.b(i, -1, k)[0] == W_FIX[5] and t.b(i, 4, k)[0] == W_FIX[5]
            if mhd and i + 1 < 3:
                assert t.b(i, -1, k)[2] == W_FIX[7] and t.b(i, 4, k)[2] == W_FIX[7]
    assert close(t.w(-1, -1, 1), G)
    t.fill()
    t.apply(E(O.FIXED, Z, state=W_FIX))
    for i in range(3):
        for j in range(4):
            assert close(t.w(i, j, -1)[:nf], W_FIX[:nf]) and close(t.w(i, j, 5)[:nf], W_FIX[:nf])
            if mhd:
                assert t.b(i, j, -1)[0] == W_FIX[5] and t.b(i, j, -1)[1] == W_FIX[6]
                assert t.b(i, j, 5)[0] == W_FIX[5] and t.b(i, j, 5)[1] == W_FIX[6]
    t.fill()
    t.apply(E(O.FIXED, X | Y | Z, state=W_FIX))
    assert close(t.w(-1, -1, -1)[:nf], W_FIX[:nf])


# ------------------------------------------------------- ignore, composition
@pytest.mark.parametrize("mhd", [0, 1])
def test_ignore(mhd):  # Boundary_Tests.cpp:290-345
    t = TGrid((4,), 2, mhd)
    t.apply(E(O.IGNORE, X))
    assert all(close(t.w(c), G) for c in (-1, -2, 4, 5))
    t = TGrid((3, 4), 1, mhd)
    B0 = None if t.B is None else t.B.copy()
    t.apply(E(O.IGNORE, X | Y))
    g = t.g
    for i in range(-g, 3 + g):
        for j in range(-g, 4 + g):
            if i < 0 or i >= 3 or j < 0 or j >= 4:
                assert close(t.w(i, j), G)
    if mhd:
        assert np.array_equal(t.B, B0)


@pytest.mark.parametrize("mhd", [0, 1])
def test_ignore_blocks_implicit_outflow(mhd):  # Boundary_Tests.cpp:347-375
    t = TGrid((3, 4), 1, mhd)
    t.apply(E(O.IGNORE, X), alone=False)
    for j in range(4):
        assert close(t.w(-1, j), G) and close(t.w(3, j), G)
    for i in range(3):
        assert close(t.w(i, -1), t.w(i, 0)) and close(t.w(i, 4), t.w(i, 3))


@pytest.mark.parametrize("mhd", [0, 1])
def test_missing_faces_default_to_outflow(mhd):  # Boundary_Tests.cpp:205-223
    t = TGrid((3, 4), 1, mhd)
    t.apply(E(O.REFLECTIVE, X), alone=False)
    assert close(t.w(-1, 0), reflect(t.w(0, 0), 0, mhd)) and close(t.w(3, 0), reflect(t.w(2, 0), 0, mhd))
    assert close(t.w(0, -1), t.w(0, 0)) and close(t.w(0, 4), t.w(0, 3))


@pytest.mark.parametrize("mhd", [0, 1])
def test_composition_3d(mhd):  # Boundary_Tests.cpp:226-248
    t = TGrid((3, 4, 5), 1, mhd)
    t.apply(E(O.REFLECTIVE, ZN), E(O.PERIODIC, Y), E(O.REFLECTIVE, ZP), alone=False)
    assert close(t.w(-1, 0, 0), t.w(0, 0, 0)) and close(t.w(3, 0, 0), t.w(2, 0, 0))
    assert close(t.w(0, -1, 0), t.w(0, 3, 0)) and close(t.w(0, 4, 0), t.w(0, 0, 0))
    assert close(t.w(0, 0, -1), reflect(t.w(0, 0, 0), 2, mhd)) and close(t.w(0, 0, 5), reflect(t.w(0, 0, 4), 2, mhd))
    assert close(t.w(-1, -1, -1), reflect(t.w(0, 3, 0), 2, mhd))


@pytest.mark.parametrize("mhd", [0, 1])
def test_composition_order(mhd):  # Boundary_Tests.cpp:250-268
    t = TGrid((3, 4, 5), 1, mhd)
    t.apply(E(O.REFLECTIVE, X | Y | Z), E(O.PERIODIC, Y), alone=False)
    assert close(t.w(0, -1, 0), t.w(0, 3, 0)) and close(t.w(0, 4, 0), t.w(0, 0, 0))
    assert close(t.w(0, 0, -1), reflect(t.w(0, 0, 0), 2, mhd)) and close(t.w(0, 0, 5), reflect(t.w(0, 0, 4), 2, mhd))
    assert close(t.w(-1, -1, -1), reflect(reflect(t.w(0, 3, 0), 0, mhd), 2, mhd))


@pytest.mark.parametrize("mhd", [0, 1])
def test_composition_overlap_order(mhd):  # Boundary_Tests.cpp:269-279
    t = TGrid((3, 4), 1, mhd)
    nf = 8 if mhd else 5
    t.apply(E(O.FIXED, X, state=W_FIX), E(O.REFLECTIVE, XN), alone=False)
    assert close(t.w(-1, 0), reflect(t.w(0, 0), 0, mhd))
    assert close(t.w(3, 0)[:nf], W_FIX[:nf])
