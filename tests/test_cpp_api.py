"""The C++ mirror of the Grid / Boundary API (include/dragon_gpu.hpp) and of
DRAGONHOARD (include/dragon_hoard.hpp), driven by tests/cpp/grid_driver (built
by __graft_entry__.build()): argument checks on the CPU, and on the GPU a Sod
tube and an Orszag-Tang run through the C++ classes, bit-identical to the
oracle, host boundary elements and frames / restarts."""
import os
import subprocess

import numpy as np
import pytest

from dragon_amd import problems as P

from helpers import oracle_grid

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "grid_driver")


def need_driver():
    if not os.path.exists(DRIVER):
        pytest.fail("tests/cpp/grid_driver is not built (run __graft_entry__.build())")


def test_cpp_api_argument_checks():
    need_driver()
    r = subprocess.run([DRIVER, "abi"], capture_output=True, text=True, timeout=60)
    print(r.stdout)
    assert r.returncode == 0 and "abi checks passed" in r.stdout


def run_driver(pb, t_end, bc, tmp_path):
    shape, d, cfg, bcs, w, B, meta = pb
    dim = len(shape)
    n = list(shape) + [1] * (3 - dim)
    dd = list(d) + [1.0] * (3 - dim)
    data = [np.ascontiguousarray(w).ravel()]
    if cfg.get("mhd") and dim >= 2:
        data.append(np.ascontiguousarray(B).ravel())
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    np.concatenate(data).tofile(fin)
    args = [DRIVER, "run", str(dim), *map(str, n), *map(repr, dd), str(cfg.get("mhd", 1)), str(cfg.get("solver", 5)),
            str(cfg.get("limiter", 1)), str(cfg.get("ct_energy", 2)), repr(cfg.get("gamma", 5.0 / 3.0)), bc, repr(t_end), fin, fout]
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    out = np.fromfile(fout)
    nw = w.size
    return out[:nw].reshape(w.shape), (out[nw:].reshape(B.shape) if len(data) > 1 else None)


@pytest.mark.gpu
def test_cpp_api_sod_matches_oracle(tmp_path):
    need_driver()
    pb = P.sod_1d(256)
    w, _ = run_driver(pb, 0.05, "outflow", tmp_path)
    og = oracle_grid(pb)
    og.advance(0.05)
    assert np.array_equal(w[..., :5], og.get_w()[..., :5])


@pytest.mark.gpu
def test_cpp_api_orszag_tang_matches_oracle(tmp_path):
    need_driver()
    pb = P.orszag_tang_2d(32)
    w, B = run_driver(pb, 0.02, "periodic", tmp_path)
    og = oracle_grid(pb)
    og.advance(0.02)
    assert np.array_equal(w, og.get_w()) and np.array_equal(B, og.get_B())


@pytest.mark.gpu
def test_cpp_api_host_boundary_element_and_split_mode():
    """HostGhostFill in a BoundaryList (the C++ counterpart of a user GhostFill
    subclass) gives the same bits as the built-in Fixed it restates, under the
    unsplit and the split step; its exceptions reach the caller;
    Config.dimension_unsplit = 0 makes advance() split."""
    need_driver()
    r = subprocess.run([DRIVER, "hostbc"], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "hostbc checks passed" in r.stdout


@pytest.mark.gpu
def test_cpp_api_frames_and_restart(tmp_path):
    """The C++ DRAGONHOARD mirror (include/dragon_hoard.hpp): a restart from a
    frame written by writeToFile continues the run bit for bit (the driver
    checks that), and the frame holds the state the grid had, in the
    reference's file order, as the Python reader reads it."""
    need_driver()
    from dragon_amd import hoard as H
    r = subprocess.run([DRIVER, "hoard", str(tmp_path)], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "hoard checks passed" in r.stdout
    meta = H.read_meta(str(tmp_path / "frame_00007.h5"))
    assert (meta.dim, meta.mhd, meta.nx, meta.ny, meta.nz, meta.ng, meta.cycle) == (3, 1, 12, 10, 14, 0, 7)
    assert meta.time == 0.02 and meta.dx == 1.0 / 12
    _, cells, _ = H.read_arrays(str(tmp_path / "frame_00007.h5"), meta)
    g = 3
    state = np.fromfile(str(tmp_path / "state.bin")).reshape(12 + 2 * g, 10 + 2 * g, 14 + 2 * g, 8)[g:-g, g:-g, g:-g]
    want = np.ascontiguousarray(np.transpose(state, (2, 1, 0, 3))).reshape(-1, 8).T  # x fastest
    assert np.array_equal(cells, want)
