"""Batched patches (a fixed sequence of launches per step, each over every
patch, per-patch dt) against the oracle: every patch advanced by the batch
must be bit-identical to the same patch advanced alone by the oracle's
Grid::advance, one CFL step at a time; graph replays must equal eager
launches; tiles must not change results; a patch that cannot step reports
its failure without disturbing the others."""
import numpy as np
import pytest

import oracle as O
from dragon_amd import problems as P

from helpers import oracle_cfg

pytestmark = pytest.mark.gpu

# (nx, ny, B0, r0, u0): every size class (largest side < 64, < 128, < 192,
# >= 192), non-square patches, x and y faces beyond one wavefront
PATCHES = [(16, 16, 4.0, 0.1, 1.5), (24, 20, 3.0, 0.09, 1.2), (33, 17, 5.0, 0.12, 2.0), (40, 40, 2.5, 0.1, 1.0),
           (12, 30, 4.5, 0.11, 1.8), (70, 66, 3.5, 0.1, 1.4), (20, 131, 4.0, 0.08, 1.6), (130, 24, 2.8, 0.12, 1.1),
           (200, 24, 3.2, 0.1, 1.3)]


def build(precision=0, patches=PATCHES, bc="outflow"):
    import dragon_amd as D
    pbs = [P.rotor_2d(nx, B0, r0, u0, ny=ny) for nx, ny, B0, r0, u0 in patches]
    cfg = D.Config(precision=precision)
    for k, v in pbs[0][2].items():
        setattr(cfg, k, v)
    b = D.Batch([pb[0] for pb in pbs], [pb[1] for pb in pbs], cfg, bc=bc)
    for p, pb in enumerate(pbs):
        b.set_state(p, pb[4], pb[5])
    b.initialize_B_fields()
    return b, pbs


def oracle_patch(pb, bc="outflow", threads=1):
    shape, d, cfg, bcs, w, B, meta = pb
    G = O.Grid(shape, d, oracle_cfg(cfg, threads), ghosts=1)
    if bc == "periodic":
        G.clear_bc()
        G.add_bc(1, O.FALL)
    G.set_w(w)
    G.set_B(B)
    G.init_B()
    return G


def test_batch_fp64_matches_oracle_per_patch():
    b, pbs = build(0)
    assert b.g == 3
    b.run(3)
    b.run(2, graph=True)
    b.sync()
    for p, pb in enumerate(pbs):
        G = oracle_patch(pb)
        for _ in range(5):
            G.advance(1e9, True, 1)
        w, B = b.get_state(p)
        st = b.stats(p)
        same_w = np.array_equal(w, G.get_w())
        same_B = np.array_equal(B, G.get_B())
        print(f"patch {p} {pb[0]}: steps {st['steps']} fails {st['fails']} dt {st['dt']:.4g} cells equal {same_w} faces equal {same_B}",
              flush=True)
        assert st["steps"] == 5 and st["status"] == 0
        assert same_w and same_B


def test_batch_fp64_100_steps_every_size_class():
    """The drop-in bar on batched patches: 100 steps (60 eager, 40 replayed
    from a HIP graph) of one patch per size class, the largest near the top
    of the configuration's range (256 x 250), each bit-identical to the
    oracle advancing that patch alone."""
    patches = [(40, 36, 4.0, 0.1, 1.5), (100, 90, 3.0, 0.09, 1.2), (150, 64, 5.0, 0.12, 2.0), (256, 250, 2.6, 0.11, 1.7),
               (197, 201, 4.4, 0.08, 1.1)]
    b, pbs = build(0, patches)
    b.run(60)
    b.run(40, graph=True)
    b.sync()
    for p, pb in enumerate(pbs):
        G = oracle_patch(pb, threads=8)
        for _ in range(100):
            G.advance(1e9, True, 1)
        w, B = b.get_state(p)
        st = b.stats(p)
        same = np.array_equal(w, G.get_w()) and np.array_equal(B, G.get_B())
        print(f"patch {pb[0]}: steps {st['steps']} fails {st['fails']} bit-equal {same}", flush=True)
        assert st["steps"] == 100 and st["status"] == 0 and same


def test_batch_periodic_matches_oracle_per_patch():
    b, pbs = build(0, PATCHES[:3] + PATCHES[6:], bc="periodic")
    b.run(3, graph=True)
    b.sync()
    for p, pb in enumerate(pbs):
        G = oracle_patch(pb, "periodic")
        for _ in range(3):
            G.advance(1e9, True, 1)
        w, B = b.get_state(p)
        print(f"periodic patch {p} {pb[0]}: {b.stats(p)}", flush=True)
        assert b.stats(p)["steps"] == 3
        assert np.array_equal(w, G.get_w()) and np.array_equal(B, G.get_B()), p


def test_batch_graph_replays_match_eager():
    a, _ = build(0)
    g, _ = build(0)
    for rep in range(3):
        a.run(2)
        a.sync()
        g.run(2, graph=True)
        g.sync()
        for p in range(len(PATCHES)):
            wa, Ba = a.get_state(p)
            wg, Bg = g.get_state(p)
            assert np.array_equal(wa, wg) and np.array_equal(Ba, Bg), (rep, p)
    print("graph replays == eager launches over 3 replays", flush=True)


def test_batch_tiles_do_not_change_results():
    """x-sweep tiles and y-sweep segment widths (set, or chosen by the
    autotuner) leave every patch's results bit for bit unchanged."""
    a, _ = build(0)
    t = a.tiles()
    a.run(2)
    a.sync()
    b, _ = build(0)
    tiles, ms = b.autotune(reps=1)
    print("autotuned tiles", tiles, "ms per class launch", ms, flush=True)
    assert all(1 <= x <= 4096 for x in tiles)
    print("autotuned y-sweep segment widths", b.ysweep(), flush=True)
    assert all(w in (0, 16, 32, 64) for w in b.ysweep())
    for xt, yw in (([3, 5, 7, 9], [16, 32, 64, 0]), ([64, 2, 11, 4], [0, 16, 0, 64])):
        c, _ = build(0)
        c.set_tiles(xt)
        c.set_ysweep(yw)
        c.run(2)
        c.sync()
        for p in range(len(PATCHES)):
            wa, Ba = a.get_state(p)
            wc, Bc = c.get_state(p)
            assert np.array_equal(wa, wc) and np.array_equal(Ba, Bc), (xt, yw, p)


def test_batch_mixed_close_to_fp64():
    a, pbs = build(0)
    m, _ = build(1)
    a.run(10)
    m.run(10)
    a.sync()
    m.sync()
    for p, pb in enumerate(pbs):
        g = 3
        wa, Ba = a.get_state(p)
        wm, Bm = m.get_state(p)
        sl = tuple(slice(g, g + s) for s in pb[0])
        ra, rm = wa[sl][..., 0], wm[sl][..., 0]
        l1 = np.abs(rm - ra).sum() / np.abs(ra).sum()
        Bi = Bm[tuple(slice(g, g + s + 1) for s in pb[0])]
        d = pb[1]
        div = (Bi[1:, :-1, 0] - Bi[:-1, :-1, 0]) / d[0] + (Bi[:-1, 1:, 1] - Bi[:-1, :-1, 1]) / d[1]
        ndiv = np.abs(div).max() * min(d) / np.abs(Bi[..., :2]).max()
        print(f"patch {p}: mixed rel L1(rho) {l1:.3g}, normalised max|div B| {ndiv:.3g}", flush=True)
        assert l1 <= 1e-5 and ndiv <= 1e-12


def test_batch_failed_patch_reports_and_keeps_state():
    """A patch with a negative pressure cannot step at any dt: each step
    halves its dt on the device (one attempt per step) down to the minimum,
    then it reports DG_ERR_DT_UNDERFLOW, keeps its state and stops; the other
    patches step normally all along."""
    import dragon_amd as D
    b, pbs = build(0, PATCHES[:4])
    w, B = b.get_state(2)
    w[3 + 5, 3 + 6, 1] = -1.0
    b.set_state(2, w, B)
    before = b.get_state(2)
    nsteps = 50  # dt ~ 1e-3 reaches the 1e-14 tolerance after ~37 halvings
    b.run(nsteps)
    with pytest.raises(D.DragonError) as e:
        b.sync()
    assert e.value.code == D.DG_ERR_DT_UNDERFLOW and e.value.where == (2,)
    after = b.get_state(2)
    assert np.array_equal(before[0], after[0]) and np.array_equal(before[1], after[1])
    st = b.stats(2)
    print("failed patch", st, flush=True)
    assert st["status"] == D.DG_ERR_DT_UNDERFLOW and st["steps"] == 0 and 30 <= st["fails"] < nsteps
    for p in (0, 1, 3):
        G = oracle_patch(pbs[p])
        for _ in range(nsteps):
            G.advance(1e9, True, 1)
        assert b.stats(p)["steps"] == nsteps
        assert np.array_equal(b.get_state(p)[0], G.get_w()), p


@pytest.mark.parametrize("limiter", [0, 4])
def test_batch_mixed_periodic_div_b(limiter):
    """Mixed batch, periodic patches, minmod and van Albada, 20 steps: the
    periodic image faces take identical updates, so normalised max|div B|
    stays at round-off."""
    import dragon_amd as D
    pbs = [P.rotor_2d(nx, B0, r0, u0, ny=ny) for nx, ny, B0, r0, u0 in PATCHES]
    cfg = D.Config(precision=1)
    for k, v in pbs[0][2].items():
        setattr(cfg, k, v)
    cfg.limiter = limiter
    b = D.Batch([pb[0] for pb in pbs], [pb[1] for pb in pbs], cfg, bc="periodic")
    for p, pb in enumerate(pbs):
        b.set_state(p, pb[4], pb[5])
    b.initialize_B_fields()
    b.run(20)
    b.sync()
    worst = 0.0
    for p, pb in enumerate(pbs):
        w, B = b.get_state(p)
        g = b.g
        nx, ny = pb[0]
        Bx = B[g:g + nx + 1, g:g + ny, 0]
        By = B[g:g + nx, g:g + ny + 1, 1]
        div = (Bx[1:, :] - Bx[:-1, :]) / pb[1][0] + (By[:, 1:] - By[:, :-1]) / pb[1][1]
        worst = max(worst, np.abs(div).max() * min(pb[1]) / np.abs(B[..., :2]).max())
    print(f"mixed batch periodic limiter {limiter}: normalised max|div B| {worst:.3g}", flush=True)
    assert worst <= 1e-12
