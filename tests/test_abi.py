"""The C ABI without a GPU: libdragon_gpu.so loads and exports every function
include/dragon_gpu.h declares, the ctypes mirror's structures match the C
layout (checked against gcc's own sizeof/offsetof), and descriptor
validation rejects bad arguments before any HIP call."""
import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "dragon_gpu.h")


def declared_functions():
    src = open(HDR).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"^[A-Za-z_][\w \*]*?\b(dg_\w+)\s*\(", src, flags=re.M)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("dg_create", "dg_destroy", "dg_step", "dg_advance", "dg_cfl_dt", "dg_set_boundary",
                 "dg_apply_boundary", "dg_upload", "dg_download", "dg_init_B"):
        assert must in names


def test_library_exports_every_declared_symbol():
    import dragon_amd
    L = C.CDLL(dragon_amd.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_ctypes_mirror_binds_every_symbol():
    import dragon_amd
    L = dragon_amd.lib()
    for n in declared_functions():
        assert getattr(L, n).argtypes is not None, f"{n} has no ctypes signature"


def test_oracle_library_loads():
    import oracle
    assert oracle.lib() is not None


def _c_layout(work):
    prog = r"""
#include <stdio.h>
#include <stddef.h>
#include "dragon_gpu.h"
#define O(T, f) printf(#T "." #f " %zu\n", offsetof(T, f))
int main(void) {
  printf("dg_desc %zu\ndg_bc %zu\ndg_status %zu\ndg_stats %zu\n", sizeof(dg_desc), sizeof(dg_bc), sizeof(dg_status), sizeof(dg_stats));
  O(dg_desc, dx); O(dg_desc, ghosts); O(dg_desc, gamma); O(dg_desc, ct_beta); O(dg_desc, cfl_coeff);
  O(dg_desc, fallback_param); O(dg_desc, timestep_tol); O(dg_desc, precision);
  O(dg_bc, state); O(dg_status, msg); O(dg_stats, t_advanced);
  return 0;
}
"""
    src = os.path.join(work, "layout.c")
    exe = os.path.join(work, "layout")
    with open(src, "w") as f:
        f.write(prog)
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), src, "-o", exe], check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    return {ln.split()[0]: int(ln.split()[1]) for ln in out.splitlines()}


def test_struct_layout_matches_c(tmp_path):
    import dragon_amd as D
    lay = _c_layout(str(tmp_path))
    assert lay["dg_desc"] == C.sizeof(D._Desc)
    assert lay["dg_bc"] == C.sizeof(D._BC)
    assert lay["dg_status"] == C.sizeof(D._Status)
    assert lay["dg_stats"] == C.sizeof(D._Stats)
    for key, off in lay.items():
        if "." in key:
            t, f = key.split(".")
            cls = {"dg_desc": D._Desc, "dg_bc": D._BC, "dg_status": D._Status, "dg_stats": D._Stats}[t]
            assert getattr(cls, f).offset == off, key


@pytest.mark.parametrize("bad", [
    dict(dim=0), dict(dim=4), dict(nx=0), dict(dim=2, ny=0), dict(dim=3, nz=0),
    dict(dx=0.0), dict(dim=2, dy=-1.0), dict(muscl=0, ctu=1),
    dict(nx=(1 << 29) - 7),                                   # 1D face index space of 2^29 elements (32-bit indexing)
    dict(dim=2, nx=30000, ny=20000), dict(dim=3, nx=900, ny=900, nz=900),
])
def test_create_rejects_bad_descriptor(bad):
    """GridND constructors and CTU.cpp:26-30 reject these before any device
    work; the ABI returns DG_ERR_ARG without touching HIP."""
    import dragon_amd as D
    d = D._Desc(dim=1, nx=8, ny=1, nz=1, dx=0.1, dy=0.1, dz=0.1, ghosts=1, mhd=0, gamma=5 / 3, solver=5,
                limiter=1, muscl=1, ctu=1, ct_energy=2, ct_beta=0.1, cfl_mode=1, cfl_coeff=0.3,
                verify_fallback=-1, fallback_try_hlle=0, fallback_param=1.0, hlld_safety=1,
                timestep_tol=1e-14, precision=0)
    for k, v in bad.items():
        setattr(d, k, v)
    out = C.c_void_p()
    assert D.lib().dg_create(C.byref(d), C.byref(out)) == D.DG_ERR_ARG
    assert not out.value
    assert D.lib().dg_create(None, C.byref(out)) == D.DG_ERR_ARG


def test_product_path_fails_loudly_without_the_library(tmp_path):
    """No CPU fallback: with libdragon_gpu.so absent, creating a grid raises
    ImportError instead of computing anything elsewhere."""
    import subprocess
    import sys
    code = ("import dragon_amd as D\n"
            f"D.LIB_PATH = {str(tmp_path / 'absent.so')!r}\n"
            "try:\n"
            "    D.Grid2D(8, 8, 0.1, 0.1, 1, D.Config(mhd=0))\n"
            "except ImportError as e:\n"
            "    print('ImportError', e)\n"
            "else:\n"
            "    print('no error')\n")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120,
                       cwd=os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    assert "ImportError" in r.stdout and "absent.so" in r.stdout, r.stdout + r.stderr
