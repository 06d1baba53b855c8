"""GhostFill boundaries on the GPU vs the oracle: every kind (Outflow, gated
Outflow, Periodic, Reflective, Fixed, Ignore), single faces and whole
dimensions, with and without corner ghosts, and ordered compositions with the
implicit outflow (BoundaryList.hpp:79-110), in 1D/2D/3D, hydro and MHD.
Cells and face B, ghosts included, must be bit-identical."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

XN, XP, YN, YP, ZN, ZP = 1, 2, 4, 8, 16, 32
X, Y, Z = 3, 12, 48
OUT, PER, REF, FIX, IGN, GATED = "out", "per", "ref", "fix", "ign", "gated"

SHAPES = {1: (7,), 2: (5, 6), 3: (4, 5, 6)}


def cases(dim):
    allf = {1: X, 2: X | Y, 3: X | Y | Z}[dim]
    singles = [XN, XP] + ([YN, YP] if dim >= 2 else []) + ([ZN, ZP] if dim >= 3 else [])
    out = []
    for kind in (OUT, GATED, PER, REF, FIX, IGN):
        for corners in (True, False):
            out.append([(kind, allf, corners)])
        for f in singles:
            out.append([(kind, f, True)])
    if dim >= 2:
        out.append([(REF, X, True)])
        out.append([(FIX, X, True), (REF, XN, True)])
        out.append([(REF, allf, True), (PER, Y, True)])
        out.append([(IGN, X, True)])
        out.append([(GATED, YP, False), (PER, X, True)])
    if dim == 3:
        out.append([(REF, ZN, True), (PER, Y, True), (REF, ZP, True)])
        out.append([(PER, X | Z, False), (FIX, YN, True), (GATED, YP, True)])
    return out


def make_elements(D, spec, state):
    els = []
    for kind, f, corners in spec:
        if kind == OUT:
            els.append(D.Outflow(f, corners))
        elif kind == GATED:
            els.append(D.Outflow.Gated(f, corners))
        elif kind == PER:
            els.append(D.Periodic(f, corners))
        elif kind == REF:
            els.append(D.Reflective(f, corners))
        elif kind == FIX:
            els.append(D.Fixed(state, f, corners))
        else:
            els.append(D.Ignore(f))
    return els


def add_oracle(og, spec, state):
    kmap = {OUT: O.OUTFLOW, GATED: O.OUTFLOW, PER: O.PERIODIC, REF: O.REFLECTIVE, FIX: O.FIXED, IGN: O.IGNORE}
    for kind, f, corners in spec:
        og.add_bc(kmap[kind], f, corners, kind == GATED, state if kind == FIX else None)


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("mhd", [0, 1])
def test_boundaries_match_oracle(dim, mhd):
    import dragon_amd as D
    shape = SHAPES[dim]
    rng = np.random.default_rng(100 * dim + mhd)
    state = O.s8(1.7, 0.9, 0.3, -0.2, 0.1, 0.5, -0.4, 0.25)
    n_checked = 0
    for spec in cases(dim):
        og = O.Grid(shape, (0.1,) * dim, O.Cfg(mhd=mhd), ghosts=1)
        w = rng.normal(0, 1, og.wshape)
        w[..., 0] = np.abs(w[..., 0]) + 0.5
        w[..., 1] = np.abs(w[..., 1]) + 0.5
        if not mhd:
            w[..., 5:] = 0.0
        og.set_w(w)
        B = None
        if og.has_B:
            B = rng.normal(0, 1, og.bshape)
            og.set_B(B)
        add_oracle(og, spec, state)
        og.apply_bc()
        gg = D.Grid(shape, (0.1,) * dim, 1, D.Config(mhd=mhd))
        assert gg.g == og.g
        gg.set_state(w, B)
        gg.boundary = D.BoundaryList(*make_elements(D, spec, state))
        gg.apply_boundary()
        wg, wo = gg.get_w(), og.get_w()
        nf = 8 if mhd else 5
        assert np.array_equal(wg[..., :nf], wo[..., :nf]), (dim, mhd, spec, np.argwhere(wg[..., :nf] != wo[..., :nf])[:5])
        if og.has_B:
            assert np.array_equal(gg.get_B(), og.get_B()), (dim, mhd, spec)
        n_checked += 1
    print(f"dim {dim} mhd {mhd}: {n_checked} boundary lists bit-identical", flush=True)
