"""Boundary lists with elements evaluated on the host (user GhostFill
subclasses, Boundary/GhostFill.hpp:16-30, e.g. the reference's Jet in
Examples/CollidingFlows/Jets/Jets.hpp:12-39).  A host element written to do
exactly what the built-in Fixed does must give the same bits as Fixed itself,
wherever it sits in the list, over an advance of several steps."""
import numpy as np
import pytest

from dragon_amd import problems as P

from helpers import gpu_grid

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


def run(pb, els, steps):
    import dragon_amd as D
    G = gpu_grid(pb)
    G.boundary = D.BoundaryList(*els)
    n = G.advance(1e9, True, steps)
    return G, n


@pytest.mark.parametrize("layout", ["first", "middle", "last"])
def test_host_element_matches_device_fixed(layout):
    import dragon_amd as D
    pb = P.kelvin_helmholtz_2d(64)
    inflow = np.array([1.5, 2.5, 0.3, 0.0, 0.0, 0.0, 0.0, 0.0])
    dev = [D.Periodic(D.Y), D.Fixed(inflow, D.X_negative), D.Reflective(D.Y_negative), D.Outflow(D.X_positive)]
    host = list(dev)
    pos = {"first": 0, "middle": 1, "last": 3}[layout]
    if layout == "first":
        dev = [D.Fixed(inflow, D.X_negative)] + [e for e in dev if not isinstance(e, D.Fixed)]
        host = [HostFixed.make(inflow, D.X_negative)] + [e for e in host if not isinstance(e, D.Fixed)]
    elif layout == "last":
        dev = [e for e in dev if not isinstance(e, D.Fixed)] + [D.Fixed(inflow, D.X_negative)]
        host = [e for e in host if not isinstance(e, D.Fixed)] + [HostFixed.make(inflow, D.X_negative)]
    else:
        host[pos] = HostFixed.make(inflow, D.X_negative)
    a, na = run(pb, dev, 8)
    b, nb = run(pb, host, 8)
    assert na == nb
    wa, wb = a.get_w(), b.get_w()
    print(layout, "steps", na, "bit-identical", np.array_equal(wa, wb), flush=True)
    assert np.array_equal(wa, wb)


def test_host_element_refused_by_split_step():
    import dragon_amd as D
    pb = P.kelvin_helmholtz_2d(32)
    G = gpu_grid(pb)
    G.boundary = D.BoundaryList(D.Periodic(), HostFixed.make(np.ones(8), D.X_negative))
    with pytest.raises(D.DragonError) as e:
        G.advance_split(1e-3)
    assert e.value.code == D.DG_ERR_UNSUPPORTED
