"""Frames and restarts from device buffers (dg_snapshot / dg_restore +
libdragon_hoard.so) on the GPU: the snapshot equals the state in the
reference's file order; a frame written and read back restores the grid;
a run restarted from a frame continues bit for bit like the uninterrupted
run; frames written by the background writer while the grid steps equal
frames written synchronously at the same points."""
import os

import numpy as np
import pytest

from dragon_amd import hoard as H
from dragon_amd import problems as P

from helpers import gpu_grid

pytestmark = pytest.mark.gpu


def file_order(a, dim, g, ng, faces=False):
    """Host restatement of the file order (x fastest) of the box [-ng, n+ng)."""
    sl = tuple(slice(g - ng, a.shape[q] - (g - ng)) for q in range(dim))
    b = a[sl]
    return np.ascontiguousarray(np.transpose(b, tuple(reversed(range(dim))) + (dim,))).reshape(-1, a.shape[-1]).T


CASES = [("sod", lambda: P.sod_1d(64)), ("kh", lambda: P.kelvin_helmholtz_2d(40)), ("ot", lambda: P.orszag_tang_2d(48)),
         ("turb", lambda: P.turbulence_3d(12)), ("rect3d", None)]


def rect3d():
    from test_gpu_parity import rect_mhd_3d
    return rect_mhd_3d((9, 70, 13))


@pytest.mark.parametrize("name,make", CASES)
def test_snapshot_is_file_order(name, make):
    pb = make() if make else rect3d()
    G = gpu_grid(pb)
    G.advance(1e9, True, 3)
    w = G.get_w()
    B = G.get_B() if G.has_B else None
    for ng in (0, G.g):
        f = H.frame_meta(G, ng=ng)
        nc, nf = H.array_sizes(f)
        cells = np.empty((8, nc))
        faces = np.empty((3, nf)) if nf else None
        H._snapshot(G, ng, cells, faces)
        assert H.lib().dg_snapshot_wait(G._h) == 0
        assert np.array_equal(cells, file_order(w, G.dim, G.g, ng)), (name, ng)
        if nf:
            assert np.array_equal(faces, file_order(B, G.dim, G.g, ng, faces=True)), (name, ng)


@pytest.mark.parametrize("name,make", CASES[:4])
def test_write_load_round_trip(tmp_path, name, make):
    """writeToFile then loadFromFile into a fresh grid (IO_Tests.cpp:86-150):
    t, cycle, dx kept; rho, p, v exact; cell B float-rounded; face B exact."""
    pb = make()
    G = gpu_grid(pb)
    G.advance(1e9, True, 4)
    old = H.output_dir
    try:
        H.output_dir = str(tmp_path)
        H.writeToFile(G, 0.666, 666, "frame_" + H.cycle_string(666))
        pb2 = list(pb)
        pb2[1] = tuple(0.5 * x for x in pb[1])  # spacing replaced by the file's
        G2 = gpu_grid(tuple(pb2))
        t, cyc = H.loadFromFile(G2, "frame_00666")
    finally:
        H.output_dir = old
    assert t == 0.666 and cyc == 666 and G2.d == G.d
    g, sh = G.g, pb[0]
    sl = tuple(slice(g, g + s) for s in sh)
    a, b = G.get_w()[sl], G2.get_w()[sl]
    assert np.array_equal(a[..., :5], b[..., :5])
    if G.config.mhd:
        assert np.array_equal(b[..., 5:], a[..., 5:].astype(np.float32).astype(np.float64))
        if G.has_B:
            fl = tuple(slice(g, g + s + 1) for s in sh)
            assert np.array_equal(G.get_B()[fl], G2.get_B()[fl])


@pytest.mark.parametrize("name,make", [("kh", lambda: P.kelvin_helmholtz_2d(64)), ("ot", lambda: P.orszag_tang_2d(64)),
                                       ("turb", lambda: P.turbulence_3d(16))])
def test_restart_continues_bit_for_bit(tmp_path, name, make):
    """main.cpp:71-97, 133-150: a run restarted from a frame (cell B written as
    double, HDF5_WRITE_DOUBLE) continues exactly like the run that never
    stopped: 20 steps, frame, 20 more steps."""
    pb = make()
    A = gpu_grid(pb)
    A.advance(1e9, True, 20)
    old = H.output_dir
    try:
        H.output_dir = str(tmp_path)
        H.writeToFile(A, 0.0, 1, "frame_00001", redundant=H.HDF5_WRITE_DOUBLE)
        R = gpu_grid(pb)
        H.loadFromFile(R, "frame_00001")
    finally:
        H.output_dir = old
    A.advance(1e9, True, 20)
    R.advance(1e9, True, 20)
    g, sh = A.g, pb[0]
    sl = tuple(slice(g, g + s) for s in sh)
    assert np.array_equal(A.get_w()[sl], R.get_w()[sl])
    if A.has_B:
        fl = tuple(slice(g, g + s + 1) for s in sh)
        assert np.array_equal(A.get_B()[fl], R.get_B()[fl])


def test_background_writer_matches_synchronous_frames(tmp_path):
    """FrameWriter: frames snapshotted and written in a background thread
    while the grid keeps stepping equal frames written synchronously from a
    second grid advanced alike."""
    pb = P.orszag_tang_2d(96)
    A, B = gpu_grid(pb), gpu_grid(pb)
    old = H.output_dir
    try:
        H.output_dir = str(tmp_path)
        with H.FrameWriter(A, compression=1) as fw:
            for c in range(1, 5):
                A.advance(1e9, True, 7)
                fw.submit(0.1 * c, c, f"async_{H.cycle_string(c)}")
        for c in range(1, 5):
            B.advance(1e9, True, 7)
            H.writeToFile(B, 0.1 * c, c, f"sync_{H.cycle_string(c)}", compression=1)
    finally:
        H.output_dir = old
    for c in range(1, 5):
        want = H.frame_meta(A, ng=A.g)
        ga, ca, fa = H.read_arrays(str(tmp_path / f"async_{H.cycle_string(c)}.h5"), want)
        gs, cs, fs = H.read_arrays(str(tmp_path / f"sync_{H.cycle_string(c)}.h5"), want)
        assert ga.cycle == gs.cycle == c and ga.time == gs.time
        assert np.array_equal(ca, cs) and np.array_equal(fa, fs), c
