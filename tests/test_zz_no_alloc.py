"""dg_step_noalloc, the variant of dg_step that allocates nothing (scratch in
a static device array), against dg_step on the same input: bit-identical
next state, on two calls in a row with the second grid smaller than the
first, in both precisions; sizes beyond the static array are refused."""
import ctypes as C

import pytest

import dragon_amd as D
from dragon_amd import problems as P

from helpers import gpu_grid


def noalloc_lib():
    return D.lib()


def desc_of(G):
    d = D._Desc()
    n = list(G.shape) + [1] * (3 - G.dim)
    d.dim, d.nx, d.ny, d.nz = G.dim, *n
    d.dx, d.dy, d.dz = list(G.d) + [1.0] * (3 - G.dim)
    d.ghosts = G.g
    for f in D.fields(D.Config):
        setattr(d, f.name, getattr(G.config, f.name))
    return d


def one_case(pb, precision, torch):
    """dt from the CFL of the boundary-filled state; dg_step on a grid
    handle vs dg_step_noalloc on the same state copied into caller buffers."""
    G = gpu_grid(pb, precision)
    G.advance(1e9, True, 2)  # a state with structure
    G.apply_boundary()
    dt = G.cfl_time()
    nc, nfc = D.lib().dg_cell_count(G._h), D.lib().dg_face_count(G._h)
    dev = torch.device("cuda:0")
    w_in = torch.empty(8 * nc, dtype=torch.float64, device=dev)
    B_in = torch.empty(max(1, 3 * nfc), dtype=torch.float64, device=dev)
    G.download_device(w_in.data_ptr(), B_in.data_ptr() if nfc else None)
    w_out = torch.full_like(w_in, float("nan"))
    B_out = torch.full_like(B_in, float("nan"))
    G.unsplit_step(dt)
    w_ref = torch.empty_like(w_in)
    B_ref = torch.empty_like(B_in)
    G.download_device(w_ref.data_ptr(), B_ref.data_ptr() if nfc else None)
    L = noalloc_lib()
    d = desc_of(G)
    st = D._Status()
    torch.cuda.synchronize()
    rc = L.dg_step_noalloc(C.byref(d), w_in.data_ptr(), B_in.data_ptr() if nfc else None, w_out.data_ptr(),
                           B_out.data_ptr() if nfc else None, dt, None, C.byref(st))
    assert rc == 0, (rc, L.dg_noalloc_last_error())
    torch.cuda.synchronize()
    nf = 8 if G.config.mhd else 5
    same_w = torch.equal(w_out[: nf * nc], w_ref[: nf * nc])
    same_B = torch.equal(B_out[: 3 * nfc], B_ref[: 3 * nfc]) if nfc else True
    print(f"{pb[0]} precision {precision}: scratch {L.dg_noalloc_scratch_bytes(C.byref(d))} bytes, cells equal {same_w}, "
          f"faces equal {same_B}", flush=True)
    assert same_w and same_B


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [D.FP64, D.MIXED])
def test_noalloc_matches_step_twice_second_smaller(precision):
    import torch
    one_case(P.turbulence_3d(16), precision, torch)     # 3D MHD CTU + CT
    one_case(P.orszag_tang_2d(40), precision, torch)    # smaller: 2D MHD
    one_case(P.kelvin_helmholtz_2d(24), precision, torch)  # 2D hydro CTU
    one_case(P.sod_1d(64), precision, torch)            # 1D


def test_noalloc_refuses_grids_beyond_the_array():
    """On the CPU: the argument checks come before any device call, so a grid
    too large for the static array, or outputs aliasing the inputs, are
    refused without touching a pointer."""
    L = noalloc_lib()
    d = D._Desc()
    d.dim, d.nx, d.ny, d.nz = 2, 4096, 4096, 1
    d.dx = d.dy = d.dz = 1.0 / 4096
    for f in D.fields(D.Config):
        setattr(d, f.name, getattr(D.Config(), f.name))
    need = L.dg_noalloc_scratch_bytes(C.byref(d))
    assert need > L.dg_noalloc_scratch_limit() == 64 << 20
    d16 = D._Desc.from_buffer_copy(d)
    d16.nx = d16.ny = 16
    assert 0 < L.dg_noalloc_scratch_bytes(C.byref(d16)) < L.dg_noalloc_scratch_limit()
    st = D._Status()
    rc = L.dg_step_noalloc(C.byref(d), 8, 16, 24, 32, 1e-3, None, C.byref(st))
    assert rc == D.DG_ERR_ARG and b"bytes of scratch" in L.dg_noalloc_last_error()
    rc = L.dg_step_noalloc(C.byref(d16), 8, 16, 8, 32, 1e-3, None, C.byref(st))
    assert rc == D.DG_ERR_ARG and b"alias" in L.dg_noalloc_last_error()
