"""Boundary lists with elements evaluated on the host (user GhostFill
subclasses, Boundary/GhostFill.hpp:16-30, e.g. the reference's Jet in
Examples/CollidingFlows/Jets/Jets.hpp:12-39).  A host element written to do
exactly what the built-in Fixed does must give the same bits as Fixed itself
and as the oracle with Fixed at the same place in the list, wherever it sits,
over an advance of several unsplit steps, and over split steps, where the
reference applies the list before every directional sweep
(Hydro/Godunov/Splitting.cpp:219-299)."""
import numpy as np
import pytest

import oracle as O
from dragon_amd import problems as P

from helpers import gpu_grid, oracle_cfg

pytestmark = pytest.mark.gpu


class HostFixed:
    """Fixed.cpp's cell fill with corner ghosts, written on the host."""

    @staticmethod
    def make(state, faces):
        import dragon_amd as D

        class _HostFixed(D.GhostFill):
            def apply(self, grid):
                g, w = grid.g, grid.w
                for ax in range(grid.dim):
                    for side, bit in ((0, 1 << (2 * ax)), (1, 2 << (2 * ax))):
                        if self.faces & bit:
                            idx = [slice(None)] * grid.dim
                            idx[ax] = slice(0, g) if side == 0 else slice(w.shape[ax] - g, w.shape[ax])
                            w[tuple(idx)] = state

        return _HostFixed(faces)


def run(pb, els, steps, split=False):
    import dragon_amd as D
    G = gpu_grid(pb)
    G.boundary = D.BoundaryList(*els)
    n = G.advance_split(1e9, True, steps) if split else G.advance(1e9, True, steps)
    return G, n


def run_oracle(pb, els, steps, split=False):
    """The oracle with the same list (device elements only: Fixed in place of
    the host element)."""
    import dragon_amd as D
    shape, d, cfg, bcs, w, B, meta = pb
    G = O.Grid(shape, d, oracle_cfg(cfg, 8), ghosts=1)
    G.set_w(w)
    kinds = {D.Periodic: O.PERIODIC, D.Fixed: O.FIXED, D.Reflective: O.REFLECTIVE, D.Outflow: O.OUTFLOW}
    for e in els:
        G.add_bc(kinds[type(e)], e.faces, True, False, e.state if isinstance(e, D.Fixed) else None)
    n = G.advance_split(1e9, True, steps) if split else G.advance(1e9, True, steps)
    return G, n


def lists(layout, inflow):
    import dragon_amd as D
    dev = [D.Periodic(D.Y), D.Fixed(inflow, D.X_negative), D.Reflective(D.Y_negative), D.Outflow(D.X_positive)]
    host = list(dev)
    if layout == "first":
        dev = [D.Fixed(inflow, D.X_negative)] + [e for e in dev if not isinstance(e, D.Fixed)]
        host = [HostFixed.make(inflow, D.X_negative)] + [e for e in host if not isinstance(e, D.Fixed)]
    elif layout == "last":
        dev = [e for e in dev if not isinstance(e, D.Fixed)] + [D.Fixed(inflow, D.X_negative)]
        host = [e for e in host if not isinstance(e, D.Fixed)] + [HostFixed.make(inflow, D.X_negative)]
    else:
        host[1] = HostFixed.make(inflow, D.X_negative)
    return dev, host


@pytest.mark.parametrize("layout", ["first", "middle", "last"])
@pytest.mark.parametrize("split", [False, True])
def test_host_element_matches_device_fixed_and_oracle(layout, split):
    pb = P.kelvin_helmholtz_2d(64)
    inflow = np.array([1.5, 2.5, 0.3, 0.0, 0.0, 0.0, 0.0, 0.0])
    dev, host = lists(layout, inflow)
    a, na = run(pb, dev, 8, split)
    b, nb = run(pb, host, 8, split)
    o, no = run_oracle(pb, dev, 8, split)
    assert na == nb == no
    wa, wb, wo = a.get_w(), b.get_w(), o.get_w()
    g, sh = a.g, pb[0]
    sl = tuple(slice(g, g + s) for s in sh)
    print(layout, "split" if split else "unsplit", "steps", na, "host == device", np.array_equal(wa, wb),
          "host == oracle (interior)", np.array_equal(wb[sl][..., :5], wo[sl][..., :5]), flush=True)
    assert np.array_equal(wa, wb)
    assert np.array_equal(wb[sl][..., :5], wo[sl][..., :5])


def test_host_element_error_propagates():
    """A host element that raises stops the advance with its own exception."""
    import dragon_amd as D

    class Broken(D.GhostFill):
        def apply(self, grid):
            raise ValueError("user boundary failed")

    G = gpu_grid(P.kelvin_helmholtz_2d(32))
    G.boundary = D.BoundaryList(D.Periodic(), Broken(D.X_negative))
    with pytest.raises(ValueError, match="user boundary failed"):
        G.advance_split(1e-3)
    with pytest.raises(ValueError, match="user boundary failed"):
        G.advance(1e-3)
