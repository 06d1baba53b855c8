"""Shared test helpers: build an oracle grid and a HIP grid from the same
initial arrays and compare them."""
import numpy as np

import oracle as O

KIND = {0: O.OUTFLOW, 1: O.PERIODIC, 2: O.REFLECTIVE, 3: O.FIXED, 4: O.IGNORE}


def oracle_cfg(cfg: dict, threads=1) -> O.Cfg:
    c = O.Cfg(threads=threads)
    for k, v in cfg.items():
        if hasattr(c, k):
            setattr(c, k, v)
    return c


def oracle_grid(problem, threads=1):
    shape, d, cfg, bcs, w, B, meta = problem
    G = O.Grid(shape, d, oracle_cfg(cfg, threads), ghosts=1)
    G.set_w(w)
    if B is not None:
        G.set_B(B)
    for bc in bcs:
        G.add_bc(KIND[bc[0]], bc[1], True, False, bc[2] if len(bc) > 2 else None)
    if meta.get("init_B"):
        G.init_B()
    return G


def gpu_grid(problem, precision=0):
    import dragon_amd as D
    shape, d, cfg, bcs, w, B, meta = problem
    conf = D.Config(precision=precision)
    for k, v in cfg.items():
        setattr(conf, k, v)
    G = D.Grid(shape, d, 1, conf)
    G.set_state(w, B)
    els = []
    for bc in bcs:
        cls = {0: D.Outflow, 1: D.Periodic, 2: D.Reflective, 4: D.Ignore}.get(bc[0])
        els.append(D.Fixed(bc[2], bc[1]) if bc[0] == 3 else cls(bc[1]))
    G.boundary = D.BoundaryList(*els)
    if meta.get("init_B"):
        G.initialize_B_fields()
    return G


def interior(a, g, shape, extra=0):
    return a[tuple(slice(g, g + s + extra) for s in shape)]


def rel_l1(a, b):
    """Per-variable relative L1 (sum|a-b| / sum|b|); identically-zero
    reference fields report the absolute max difference instead."""
    out = []
    for q in range(a.shape[-1]):
        den = np.abs(b[..., q]).sum()
        num = np.abs(a[..., q] - b[..., q]).sum()
        out.append(num / den if den > 0 else np.abs(a[..., q] - b[..., q]).max())
    return np.array(out)


def compare_states(gg, og, shape, nf):
    wg, wo = gg.get_w(), og.get_w()
    g = og.g
    a = interior(wg, g, shape)[..., :nf]
    b = interior(wo, g, shape)[..., :nf]
    res = {"rel_l1": rel_l1(a, b), "bit_equal": bool(np.array_equal(a, b)),
           "max_abs": float(np.abs(a - b).max())}
    if og.has_B:
        Bg, Bo = gg.get_B(), og.get_B()
        fa, fb = interior(Bg, g, shape, 1), interior(Bo, g, shape, 1)
        res["B_bit_equal"] = bool(np.array_equal(fa, fb))
        res["B_max_abs"] = float(np.abs(fa - fb).max())
    return res
