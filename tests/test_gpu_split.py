"""Strang split step (Grid2D/3D::split_step, Splitting.cpp:118-299; hydro) on
the GPU against the oracle's split step, bit for bit: 2D Kelvin-Helmholtz
with the rotating X/Y order, 3D random hydro over all six sweep orders, a
non-cubic 3D box; the mixed form close to FP64; MHD refused."""
import numpy as np
import pytest

from dragon_amd import problems as P

from helpers import compare_states, gpu_grid, oracle_grid

pytestmark = pytest.mark.gpu


def random_hydro_3d(shape, seed):
    g = P.ghosts_for(False)
    w, _ = P.empty_state(shape, g, False)
    rng = np.random.default_rng(seed)
    wi = P.interior(w, g, shape)
    wi[..., 0] = 1.0 + 0.5 * rng.random(shape)
    wi[..., 1] = 1.0 + 0.5 * rng.random(shape)
    wi[..., 2:5] = 0.6 * (rng.random(shape + (3,)) - 0.5)
    cfg = dict(mhd=0, gamma=5.0 / 3.0, solver=5, limiter=2, ctu=1)
    return shape, tuple(1.0 / s for s in shape), cfg, [(1, 63)], w, None, {}


def run_split_pair(pb, steps):
    og = oracle_grid(pb)
    gg = gpu_grid(pb)
    n_o = og.advance_split(1e9, True, steps)
    n_g = gg.advance_split(1e9, True, steps)
    assert n_o == n_g, (n_o, n_g)
    res = compare_states(gg, og, pb[0], 5)
    print(pb[0], "split steps", n_o, res, flush=True)
    assert res["bit_equal"], res
    return gg


def test_split_kh_2d_matches_oracle():
    run_split_pair(P.kelvin_helmholtz_2d(64), 20)


def test_split_hydro_3d_matches_oracle():
    run_split_pair(random_hydro_3d((16, 16, 16), 3), 12)  # every one of the six orders twice


def test_split_hydro_3d_rectangular_matches_oracle():
    run_split_pair(random_hydro_3d((20, 7, 70), 5), 7)


def test_split_mixed_close_to_fp64():
    pb = P.kelvin_helmholtz_2d(128)
    a, m = gpu_grid(pb, 0), gpu_grid(pb, 1)
    a.advance_split(1e9, True, 30)
    m.advance_split(1e9, True, 30)
    g = a.g
    ra = a.get_w()[g:-g, g:-g, 0]
    rm = m.get_w()[g:-g, g:-g, 0]
    l1 = np.abs(rm - ra).sum() / np.abs(ra).sum()
    print("split KH 128^2 mixed vs FP64 rel L1(rho) after 30 steps:", l1, flush=True)
    assert l1 <= 1e-5


def test_split_refuses_mhd():
    import dragon_amd as D
    G = gpu_grid(P.orszag_tang_2d(32))
    with pytest.raises(D.DragonError) as e:
        G.split_step(1e-3)
    assert e.value.code == D.DG_ERR_UNSUPPORTED and "Unsplit" in str(e.value)
