"""dragon_amd -- MI355X-native unsplit Godunov step of DRAGON.

Python mirror of the reference's Grid / Boundary / Config.h API over the
C ABI in ``include/dragon_gpu.h`` (``libdragon_gpu.so``, HIP kernels for
gfx950).  There is no CPU fallback: importing the package without the built
library, or stepping without a GPU, raises.

Reference names kept (paths relative to the reference's DRAGON/ tree):
``Grid1D/2D/3D`` (Hydro/Grid.hpp:40-144) with ``advance``,
``advance_unsplit``, ``unsplit_step``, ``on_step_fail``,
``initialize_B_fields`` and a ``boundary`` list; boundary elements
``Outflow`` (+``Outflow.Gated``), ``Fixed``, ``Reflective``, ``Periodic``,
``Ignore`` composed into a ``BoundaryList`` (Boundary/Boundary.hpp:160-229,
BoundaryList.hpp); ``face_mask`` (Boundary/GhostFill.cpp:14-49).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, fields

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdragon_gpu.so")

# Config.h enums
RIEMANN_EXACT, RIEMANN_HLL, RIEMANN_HLLC, RIEMANN_HLLD, RIEMANN_HLLE, RIEMANN_HLLX = 0, 1, 2, 3, 4, 5
LIMITER_MINMOD, LIMITER_MC, LIMITER_VANLEER, LIMITER_SUPERBEE, LIMITER_VANALBADA = 0, 1, 2, 3, 4
CT_CONSV_TOTAL_E, CT_CONSV_THERMAL, CT_CONSV_BETA_GATED = 0, 1, 2
CFL_MAX, CFL_ADD = 0, 1
FP64, MIXED = 0, 1

# status codes (dragon_gpu.h)
DG_OK, DG_ERR_UNPHYSICAL, DG_ERR_NAN, DG_ERR_RIEMANN_INPUT, DG_ERR_RIEMANN_SOLUTION = 0, 1, 2, 3, 4
DG_ERR_CFL, DG_ERR_DT_UNDERFLOW, DG_ERR_ARG, DG_ERR_HIP, DG_ERR_UNSUPPORTED = 5, 6, 7, 8, 9

# Face bits (Boundary.hpp:139-149)
X_negative, X_positive, Y_negative, Y_positive, Z_negative, Z_positive = 1, 2, 4, 8, 16, 32
X, Y, Z = 3, 12, 48


class DragonError(RuntimeError):
    """A failed call; ``code`` is a DG_* status, ``where`` the cell if any."""

    def __init__(self, code, msg, where=None):
        super().__init__(msg)
        self.code = code
        self.where = where


# --------------------------------------------------------------- C types
class _Desc(C.Structure):
    _fields_ = [("dim", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double), ("ghosts", C.c_int),
                ("mhd", C.c_int), ("gamma", C.c_double), ("solver", C.c_int), ("limiter", C.c_int),
                ("muscl", C.c_int), ("ctu", C.c_int), ("ct_energy", C.c_int), ("ct_beta", C.c_double),
                ("cfl_mode", C.c_int), ("cfl_coeff", C.c_double), ("verify_fallback", C.c_int),
                ("fallback_try_hlle", C.c_int), ("fallback_param", C.c_double), ("hlld_safety", C.c_int),
                ("timestep_tol", C.c_double), ("precision", C.c_int)]


class _BC(C.Structure):
    _fields_ = [("kind", C.c_int), ("faces", C.c_int), ("corners", C.c_int), ("gated", C.c_int),
                ("state", C.c_double * 8)]


class _Status(C.Structure):
    _fields_ = [("code", C.c_int), ("i", C.c_int), ("j", C.c_int), ("k", C.c_int), ("msg", C.c_char * 192)]


class _Stats(C.Structure):
    _fields_ = [("steps", C.c_long), ("failures", C.c_long), ("t_advanced", C.c_double)]


_dp = C.POINTER(C.c_double)
_BC_CALLBACK = C.CFUNCTYPE(C.c_int, C.c_void_p)
_lib = None


def lib():
    """The loaded HIP library; raises if it was not built (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"dragon_amd: {LIB_PATH} is missing -- build it (python -c 'import __graft_entry__ as g; g.build()')")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        sig = {
            "dg_create": (C.c_int, [C.POINTER(_Desc), C.POINTER(vp)]),
            "dg_destroy": (C.c_int, [vp]),
            "dg_ghosts": (C.c_int, [vp]),
            "dg_cell_count": (C.c_long, [vp]),
            "dg_face_count": (C.c_long, [vp]),
            "dg_upload": (C.c_int, [vp, _dp, C.c_int, _dp]),
            "dg_download": (C.c_int, [vp, _dp, C.c_int, _dp]),
            "dg_upload_device": (C.c_int, [vp, vp, vp]),
            "dg_download_device": (C.c_int, [vp, vp, vp]),
            "dg_set_boundary": (C.c_int, [vp, C.POINTER(_BC), C.c_int]),
            "dg_apply_boundary": (C.c_int, [vp]),
            "dg_apply_boundary_element": (C.c_int, [vp, C.POINTER(_BC)]),
            "dg_set_boundary_callback": (C.c_int, [vp, vp, vp]),
            "dg_cfl_dt": (C.c_int, [vp, _dp, C.POINTER(_Status)]),
            "dg_step": (C.c_int, [vp, C.c_double, C.POINTER(_Status)]),
            "dg_step_noalloc": (C.c_int, [C.POINTER(_Desc), vp, vp, vp, vp, C.c_double, vp, C.POINTER(_Status)]),
            "dg_noalloc_scratch_bytes": (C.c_long, [C.POINTER(_Desc)]),
            "dg_noalloc_scratch_limit": (C.c_long, []),
            "dg_noalloc_last_error": (C.c_char_p, []),
            "dg_advance": (C.c_int, [vp, C.c_double, C.c_int, C.c_long, C.POINTER(_Stats), C.POINTER(_Status)]),
            "dg_split_step": (C.c_int, [vp, C.c_double, C.POINTER(_Status)]),
            "dg_advance_split": (C.c_int, [vp, C.c_double, C.c_int, C.c_long, C.POINTER(_Stats), C.POINTER(_Status)]),
            "dg_init_B": (C.c_int, [vp]),
            "dg_last_error": (C.c_char_p, [vp]),
            "dg_stream": (vp, [vp]),
            "dg_divb": (C.c_int, [vp, _dp]),
            "dg_timing": (C.c_int, [vp, C.c_int, _dp, C.POINTER(C.c_long)]),
            "dg_kernel_time": (C.c_int, [vp, C.c_int, C.POINTER(C.c_char_p), _dp, C.POINTER(C.c_long)]),
            "dg_snapshot": (C.c_int, [vp, C.c_int, vp, vp]),
            "dg_snapshot_wait": (C.c_int, [vp]),
            "dg_restore": (C.c_int, [vp, C.c_int, vp, vp]),
            "dg_set_spacing": (C.c_int, [vp, C.c_double, C.c_double, C.c_double]),
            "dg_host_alloc": (C.c_int, [C.c_long, C.POINTER(vp)]),
            "dg_host_free": (C.c_int, [vp]),
            "dg_batch_create": (C.c_int, [C.POINTER(_Desc), C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), _dp, _dp, C.c_int,
                                          C.POINTER(vp)]),
            "dg_batch_destroy": (C.c_int, [vp]),
            "dg_batch_zones": (C.c_long, [vp]),
            "dg_batch_ghosts": (C.c_int, [vp]),
            "dg_batch_upload": (C.c_int, [vp, C.c_int, _dp, C.c_int, _dp]),
            "dg_batch_download": (C.c_int, [vp, C.c_int, _dp, C.c_int, _dp]),
            "dg_batch_init_B": (C.c_int, [vp]),
            "dg_batch_run": (C.c_int, [vp, C.c_int, C.c_int, _dp]),
            "dg_batch_sync": (C.c_int, [vp, C.POINTER(_Status)]),
            "dg_batch_patch_stats": (C.c_int, [vp, C.c_int, _dp, C.POINTER(C.c_long), C.POINTER(C.c_long), C.POINTER(C.c_int)]),
            "dg_batch_set_tiles": (C.c_int, [vp, C.POINTER(C.c_int), C.c_int]),
            "dg_batch_get_tiles": (C.c_int, [vp, C.POINTER(C.c_int), C.c_int]),
            "dg_batch_autotune": (C.c_int, [vp, C.c_int, C.POINTER(C.c_int), _dp]),
            "dg_batch_set_ysweep": (C.c_int, [vp, C.POINTER(C.c_int), C.c_int]),
            "dg_batch_get_ysweep": (C.c_int, [vp, C.POINTER(C.c_int), C.c_int]),
            "dg_batch_last_error": (C.c_char_p, [vp]),
            "dg_batch_stream": (vp, [vp]),
        }
        for n, (r, a) in sig.items():
            f = getattr(L, n)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


# ------------------------------------------------------------------ config
@dataclass
class Config:
    """Config.h switches as run-time values (production defaults, Config.h:27-90)."""
    mhd: int = 1
    gamma: float = 5.0 / 3.0
    solver: int = RIEMANN_HLLX
    limiter: int = LIMITER_MC
    muscl: int = 1
    ctu: int = 1
    ct_energy: int = CT_CONSV_BETA_GATED
    ct_beta: float = 0.1
    cfl_mode: int = CFL_ADD
    cfl_coeff: float = 0.3
    verify_fallback: int = -1
    fallback_try_hlle: int = 0
    fallback_param: float = 1.0
    hlld_safety: int = 1
    timestep_tol: float = 1e-14
    precision: int = FP64
    # DIMENSION_UNSPLIT: Grid.advance runs advance_unsplit when set, advance_split
    # (Strang splitting, hydro only) when not (CFL.cpp:151-157)
    dimension_unsplit: int = 1


# ---------------------------------------------------------------- boundary
def face_mask(s: str) -> int:
    """String face mask, e.g. "XY+" or "LR" (GhostFill.cpp:14-49)."""
    mask = 0
    s = s + " "
    for a, b in zip(s, s[1:]):
        u = a.upper()
        if u in "XYZ":
            sh = 2 * "XYZ".index(u)
            mask |= (2 << sh) if b == "+" else (1 << sh) if b == "-" else (3 << sh)
        elif u == "L":
            mask |= X_negative
        elif u == "R":
            mask |= X_positive
        elif u == "U":
            mask |= Y_positive
        elif u == "D":
            mask |= Y_negative
    return mask


def _faces(f) -> int:
    return face_mask(f) if isinstance(f, str) else int(f)


class GhostFill:
    """A boundary element (Boundary/GhostFill.hpp:16-30).  The built-in kinds
    run on the GPU.  A user subclass that overrides apply(grid) -- like the
    reference's Jet (Examples/CollidingFlows/Jets/Jets.hpp:12-39) -- runs on
    the host: the grid's state is downloaded into grid.w (cells, ghosts
    included, 8 fields) and grid.B (faces), apply() edits their ghost entries,
    and the state is uploaded again.  Lists mixing both keep the reference's
    order (implicit Outflow on uncovered faces first, then the elements)."""
    kind = -1

    def __init__(self, faces=X | Y | Z, corner_ghosts=True):
        self.faces = _faces(faces)
        self.corners = bool(corner_ghosts)
        self.gated = False
        self.state = np.zeros(8)

    def get_faces(self):
        return self.faces

    def apply(self, grid):
        """Fill this element's ghost cells of grid.w / grid.B (host arrays)."""
        raise NotImplementedError("GhostFill subclasses evaluated on the host override apply(grid)")

    @property
    def on_host(self) -> bool:
        return type(self).apply is not GhostFill.apply

    def _c(self) -> _BC:
        b = _BC()
        b.kind, b.faces, b.corners, b.gated = self.kind, self.faces, int(self.corners), int(self.gated)
        for q in range(8):
            b.state[q] = float(self.state[q]) if q < len(self.state) else 0.0
        return b

    def __add__(self, other):
        return BoundaryList(self) + other


class Outflow(GhostFill):
    kind = 0

    def __init__(self, faces=X | Y | Z, corner_ghosts=True, gated=False):
        super().__init__(faces, corner_ghosts)
        self.gated = bool(gated)

    @staticmethod
    def Gated(faces=X | Y | Z, corner_ghosts=True):
        return Outflow(faces, corner_ghosts, True)


class Periodic(GhostFill):
    kind = 1

    def __init__(self, faces=X | Y | Z, corner_ghosts=True):
        f = _faces(faces)
        f = (X if f & X else 0) | (Y if f & Y else 0) | (Z if f & Z else 0)  # Periodic.cpp:14-15
        super().__init__(f, corner_ghosts)


class Reflective(GhostFill):
    kind = 2


class Fixed(GhostFill):
    kind = 3

    def __init__(self, state, faces=X | Y | Z, corner_ghosts=True):
        super().__init__(faces, corner_ghosts)
        self.state = np.zeros(8)
        st = np.asarray(state, dtype=float)
        self.state[: len(st)] = st


class Ignore(GhostFill):
    kind = 4


class BoundaryList:
    """Ordered boundary elements; uncovered faces get an implicit Outflow first."""

    def __init__(self, *elements):
        self.elements = []
        for e in elements:
            self.append(e)

    def append(self, e):
        if isinstance(e, BoundaryList):
            self.elements.extend(e.elements)
        else:
            self.elements.append(e)
        return self

    def prepend(self, e):
        if isinstance(e, BoundaryList):
            self.elements[:0] = e.elements
        else:
            self.elements.insert(0, e)
        return self

    def clear(self):
        self.elements = []

    def __add__(self, other):
        out = BoundaryList(*self.elements)
        return out.append(other)

    def __iadd__(self, other):
        return self.append(other)


# -------------------------------------------------------------------- grid
class Grid:
    """Device-resident grid with the reference's Grid API (Hydro/Grid.hpp:19-37)."""

    def __init__(self, shape, d, ghosts, config: Config):
        self.config = config
        self.dim = len(shape)
        self.shape = tuple(int(s) for s in shape)
        desc = _Desc()
        n = list(self.shape) + [1] * (3 - self.dim)
        dd = list(d) + [1.0] * (3 - len(d))
        desc.dim, desc.nx, desc.ny, desc.nz = self.dim, n[0], n[1], n[2]
        desc.dx, desc.dy, desc.dz = dd
        desc.ghosts = int(ghosts)
        for f in fields(Config):
            setattr(desc, f.name, getattr(config, f.name))
        h = C.c_void_p()
        rc = lib().dg_create(C.byref(desc), C.byref(h))
        if rc:
            raise DragonError(rc, f"dg_create failed ({rc})")
        self._h = h
        self.g = lib().dg_ghosts(h)
        self.d = tuple(dd[: self.dim])
        self.dx = dd[0]
        if self.dim >= 2:
            self.dy = dd[1]
        if self.dim >= 3:
            self.dz = dd[2]
        self.has_B = bool(config.mhd) and self.dim >= 2
        self._boundary = BoundaryList()
        self._bc_pushed = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().dg_destroy(h)
            self._h = None

    # -- geometry
    def getGhosts(self):
        return self.g

    @property
    def wshape(self):
        return tuple(s + 2 * self.g for s in self.shape) + (8,)

    @property
    def bshape(self):
        return tuple(s + 1 + 2 * self.g for s in self.shape) + (3,)

    @property
    def stream(self):
        return lib().dg_stream(self._h)

    # -- state
    def get_w(self) -> np.ndarray:
        a = np.zeros(self.wshape)
        self._chk(lib().dg_download(self._h, a.ctypes.data_as(_dp), 8, None))
        return a

    def get_B(self) -> np.ndarray:
        a = np.zeros(self.bshape)
        if self.has_B:
            w = np.zeros(self.wshape)
            self._chk(lib().dg_download(self._h, w.ctypes.data_as(_dp), 8, a.ctypes.data_as(_dp)))
        return a

    def set_state(self, w: np.ndarray, B: np.ndarray | None = None):
        w = np.ascontiguousarray(w, dtype=np.float64)
        assert w.shape == self.wshape, (w.shape, self.wshape)
        bp = None
        if B is not None and self.has_B:
            B = np.ascontiguousarray(B, dtype=np.float64)
            assert B.shape == self.bshape
            bp = B.ctypes.data_as(_dp)
        self._chk(lib().dg_upload(self._h, w.ctypes.data_as(_dp), 8, bp))

    def upload_device(self, w_ptr: int, B_ptr: int | None = None):
        self._chk(lib().dg_upload_device(self._h, C.c_void_p(w_ptr), C.c_void_p(B_ptr) if B_ptr else None))

    def download_device(self, w_ptr: int, B_ptr: int | None = None):
        self._chk(lib().dg_download_device(self._h, C.c_void_p(w_ptr), C.c_void_p(B_ptr) if B_ptr else None))

    # -- boundary (Grid::boundary)
    @property
    def boundary(self) -> BoundaryList:
        return self._boundary

    @boundary.setter
    def boundary(self, b):
        self._boundary = b if isinstance(b, BoundaryList) else BoundaryList(b)

    def _host_bc(self) -> bool:
        return any(e.on_host for e in self._boundary.elements)

    def _push_bc(self):
        # with host elements the device list is one Ignore on every face and a
        # boundary callback is set: every boundary pass of the library (the
        # advance loops, the pass before each split sweep, dg_init_B) calls
        # back here, and _apply_list sequences the list
        host = self._host_bc()
        els = [Ignore(X | Y | Z)._c()] if host else [e._c() for e in self._boundary.elements]
        key = (host, [(b.kind, b.faces, b.corners, b.gated, tuple(b.state)) for b in els])
        if key == self._bc_pushed:
            return
        arr = (_BC * max(1, len(els)))(*els)
        self._chk(lib().dg_set_boundary(self._h, arr, len(els)))
        if host:
            self._cb = _BC_CALLBACK(self._bc_callback)
            self._chk(lib().dg_set_boundary_callback(self._h, C.cast(self._cb, C.c_void_p), None))
        else:
            self._chk(lib().dg_set_boundary_callback(self._h, None, None))
            self._cb = None
        self._bc_pushed = key

    def _bc_callback(self, user):
        try:
            self._apply_list()
            return 0
        except DragonError as e:
            self._cb_error = e
            return e.code or DG_ERR_ARG
        except Exception as e:  # a user GhostFill raised: reported after the library call returns
            self._cb_error = e
            return DG_ERR_ARG

    def _raise_cb(self):
        e, self._cb_error = getattr(self, "_cb_error", None), None
        if e is not None:
            raise e

    def apply_boundary(self):
        """BoundaryList::apply (BoundaryList.hpp:79-104)."""
        self._push_bc()
        try:
            self._chk(lib().dg_apply_boundary(self._h))
        finally:
            self._raise_cb()

    def _apply_list(self):
        """The list with host elements, in the reference's order: the implicit
        Outflow on uncovered faces, then each element; device elements run
        through dg_apply_boundary_element, host ones on a downloaded copy that
        is uploaded before the next device element (or at the end)."""
        els = self._boundary.elements
        covered = 0
        for e in els:
            covered |= e.faces
        fresh = dirty = False  # host arrays current / edited since the download

        def device(e):
            nonlocal fresh, dirty
            if dirty:
                self.set_state(self.w, self.B if self.has_B else None)
                dirty = False
            fresh = False
            b = e._c()
            self._chk(lib().dg_apply_boundary_element(self._h, C.byref(b)))

        device(Outflow(63 & ~covered))  # implicit Outflow on the uncovered faces first
        for e in els:
            if e.on_host:
                if not fresh:
                    self.w = self.get_w()
                    self.B = self.get_B() if self.has_B else None
                    fresh = True
                e.apply(self)
                dirty = True
            else:
                device(e)
        if dirty:
            self.set_state(self.w, self.B if self.has_B else None)

    # -- stepping
    def _chk(self, rc, st: _Status | None = None):
        if rc:
            msg = st.msg.decode() if st is not None and st.msg else (lib().dg_last_error(self._h) or b"").decode()
            where = (st.i, st.j, st.k) if st is not None else None
            raise DragonError(rc, msg, where)

    def cfl_time(self) -> float:
        """CFL::cfl_time (CFL.cpp:113-162)."""
        out = np.zeros(1)
        st = _Status()
        self._chk(lib().dg_cfl_dt(self._h, out.ctypes.data_as(_dp), C.byref(st)), st)
        return float(out[0])

    def unsplit_step(self, dt: float):
        """One unsplit step; raises DragonError and commits nothing on failure."""
        st = _Status()
        self._chk(lib().dg_step(self._h, float(dt), C.byref(st)), st)

    def on_step_fail(self, e) -> bool:  # Unsplit.cpp:21-23 (root grid: retry)
        return True

    def advance_unsplit(self, dt: float, check_cfl=True, max_steps=-1):
        """Grid::advance_unsplit (CFL.cpp:213-238); returns (steps, failed attempts)."""
        self._push_bc()
        st, S = _Status(), _Stats()
        try:
            self._chk(lib().dg_advance(self._h, float(dt), int(check_cfl), int(max_steps), C.byref(S), C.byref(st)), st)
        finally:
            self._raise_cb()
        self.t_advanced = float(S.t_advanced)  # simulated time of this call
        return int(S.steps), int(S.failures)

    def split_step(self, dt: float):
        """Grid2D/3D::split_step (Splitting.cpp:118-216), hydro; raises
        DragonError and commits nothing on failure."""
        self._push_bc()
        st = _Status()
        try:
            self._chk(lib().dg_split_step(self._h, float(dt), C.byref(st)), st)
        finally:
            self._raise_cb()

    def advance_split(self, dt: float, check_cfl=True, max_steps=-1):
        """Grid::advance_split (CFL.cpp:160-184); returns (steps, failed attempts)."""
        self._push_bc()
        st, S = _Status(), _Stats()
        try:
            self._chk(lib().dg_advance_split(self._h, float(dt), int(check_cfl), int(max_steps), C.byref(S), C.byref(st)), st)
        finally:
            self._raise_cb()
        self.t_advanced = float(S.t_advanced)
        return int(S.steps), int(S.failures)

    def advance(self, dt: float, check_cfl=True, max_steps=-1):
        """Grid::advance (CFL.cpp:151-157): unsplit or split by Config.dimension_unsplit."""
        if self.config.dimension_unsplit:
            return self.advance_unsplit(dt, check_cfl, max_steps)
        return self.advance_split(dt, check_cfl, max_steps)

    def initialize_B_fields(self):
        """initialize_B_fields (B_fields.cpp:117-149): the boundary pass, then
        the cell-centred B from the faces."""
        self._push_bc()
        try:
            self._chk(lib().dg_init_B(self._h))
        finally:
            self._raise_cb()

    def timing(self, enable: int = -1):
        """Start (1) / stop (0) / read (-1) live HIP-event timing of the step
        kernels; returns (summed step ms, timed steps) before the change."""
        ms = np.zeros(1)
        n = C.c_long(0)
        self._chk(lib().dg_timing(self._h, int(enable), ms.ctypes.data_as(_dp), C.byref(n)))
        return float(ms[0]), int(n.value)

    def kernel_times(self) -> dict:
        """Per-kernel-class device time recorded while timing is on:
        {name: (total ms, launches)}; classes never launched are omitted."""
        out = {}
        k = 0
        while True:
            name = C.c_char_p()
            ms = np.zeros(1)
            n = C.c_long(0)
            if lib().dg_kernel_time(self._h, k, C.byref(name), ms.ctypes.data_as(_dp), C.byref(n)) != DG_OK:
                break
            if n.value:
                out[name.value.decode()] = (float(ms[0]), int(n.value))
            k += 1
        return out

    def divb(self) -> float:
        out = np.zeros(1)
        self._chk(lib().dg_divb(self._h, out.ctypes.data_as(_dp)))
        return float(out[0])


class Grid1D(Grid):
    def __init__(self, size, dx, ghosts=1, config: Config | None = None):
        super().__init__((size,), (dx,), ghosts, config or Config(mhd=0))


class Grid2D(Grid):
    def __init__(self, nx, ny, dx, dy, ghosts=1, config: Config | None = None):
        super().__init__((nx, ny), (dx, dy), ghosts, config or Config())


class Grid3D(Grid):
    def __init__(self, nx, ny, nz, dx, dy, dz, ghosts=2, config: Config | None = None):
        super().__init__((nx, ny, nz), (dx, dy, dz), ghosts, config or Config())


# --------------------------------------------------------------- batches
BATCH_CLASSES = 4


class Batch:
    """Many independent 2D MHD patches (CTU + CT), each its own Grid2D with
    3 ghosts and outflow (or periodic) faces, advanced together: a step is
    a fixed sequence of launches over every patch, each patch attempting one
    step at its own CFL dt.  A failed attempt keeps the patch's state and
    halves its dt for the next step.  It replaces the per-child dispatch of
    DistGrid<T>::step (Refinement/DistGrid.cpp:291-303) for patches that
    share no dt."""

    def __init__(self, sizes, d, config: Config | None = None, bc: str = "outflow"):
        cfg = config or Config()
        sizes = np.asarray(sizes, dtype=np.int64).reshape(-1, 2)
        d = np.asarray(d, dtype=np.float64).reshape(-1, 2)
        assert len(sizes) == len(d)
        self.n = len(sizes)
        self.sizes = [tuple(int(v) for v in r) for r in sizes]
        self.config = cfg
        desc = _Desc()
        for f in fields(Config):
            setattr(desc, f.name, getattr(cfg, f.name))
        desc.dim = 2
        nx = (C.c_int * self.n)(*[r[0] for r in self.sizes])
        ny = (C.c_int * self.n)(*[r[1] for r in self.sizes])
        dx = np.ascontiguousarray(d[:, 0])
        dy = np.ascontiguousarray(d[:, 1])
        kind = {"outflow": Outflow.kind, "periodic": Periodic.kind}[bc]
        h = C.c_void_p()
        rc = lib().dg_batch_create(C.byref(desc), self.n, nx, ny, dx.ctypes.data_as(_dp), dy.ctypes.data_as(_dp), kind, C.byref(h))
        if rc:
            raise DragonError(rc, f"dg_batch_create failed ({rc})")
        self._h = h
        self.g = lib().dg_batch_ghosts(h)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().dg_batch_destroy(h)
            self._h = None

    def _chk(self, rc, st: _Status | None = None):
        if rc:
            msg = st.msg.decode() if st is not None and st.msg else (lib().dg_batch_last_error(self._h) or b"").decode()
            raise DragonError(rc, msg, (st.i,) if st is not None else None)

    @property
    def zones(self) -> int:
        return int(lib().dg_batch_zones(self._h))

    @property
    def stream(self):
        return lib().dg_batch_stream(self._h)

    def wshape(self, p):
        return tuple(s + 2 * self.g for s in self.sizes[p]) + (8,)

    def bshape(self, p):
        return tuple(s + 1 + 2 * self.g for s in self.sizes[p]) + (3,)

    def set_state(self, p: int, w: np.ndarray, B: np.ndarray):
        w = np.ascontiguousarray(w, dtype=np.float64)
        B = np.ascontiguousarray(B, dtype=np.float64)
        assert w.shape == self.wshape(p) and B.shape == self.bshape(p)
        self._chk(lib().dg_batch_upload(self._h, p, w.ctypes.data_as(_dp), 8, B.ctypes.data_as(_dp)))

    def get_state(self, p: int):
        w = np.zeros(self.wshape(p))
        B = np.zeros(self.bshape(p))
        self._chk(lib().dg_batch_download(self._h, p, w.ctypes.data_as(_dp), 8, B.ctypes.data_as(_dp)))
        return w, B

    def initialize_B_fields(self):
        self._chk(lib().dg_batch_init_B(self._h))

    def run(self, nsteps: int = 1, graph: bool = False, timed: bool = False):
        """nsteps steps of every patch (asynchronous unless timed); returns
        the device ms of the nsteps when timed."""
        ms = np.zeros(1)
        self._chk(lib().dg_batch_run(self._h, int(nsteps), int(graph), ms.ctypes.data_as(_dp) if timed else None))
        return float(ms[0]) if timed else None

    def sync(self):
        """Wait; raise DragonError for the first patch whose step failed."""
        st = _Status()
        self._chk(lib().dg_batch_sync(self._h, C.byref(st)), st)

    def stats(self, p: int) -> dict:
        dt = np.zeros(1)
        steps, fails, status = C.c_long(0), C.c_long(0), C.c_int(0)
        self._chk(lib().dg_batch_patch_stats(self._h, p, dt.ctypes.data_as(_dp), C.byref(steps), C.byref(fails), C.byref(status)))
        return {"dt": float(dt[0]), "steps": steps.value, "fails": fails.value, "status": status.value}

    def tiles(self):
        out = (C.c_int * BATCH_CLASSES)()
        self._chk(lib().dg_batch_get_tiles(self._h, out, BATCH_CLASSES))
        return list(out)

    def set_tiles(self, chunks):
        arr = (C.c_int * len(chunks))(*chunks)
        self._chk(lib().dg_batch_set_tiles(self._h, arr, len(chunks)))

    def ysweep(self):
        """The y sweep's segment width (lanes) per size class; 0: each
        patch's columns streamed through its wavefronts."""
        out = (C.c_int * BATCH_CLASSES)()
        self._chk(lib().dg_batch_get_ysweep(self._h, out, BATCH_CLASSES))
        return list(out)

    def set_ysweep(self, widths):
        arr = (C.c_int * len(widths))(*widths)
        self._chk(lib().dg_batch_set_ysweep(self._h, arr, len(widths)))

    def autotune(self, reps: int = 3):
        """Pick the x-sweep tile and then the y-sweep segment width per size
        class by timing; returns (tiles, ms per class launch); ysweep() reads
        the widths chosen."""
        out = (C.c_int * BATCH_CLASSES)()
        ms = np.zeros(BATCH_CLASSES)
        self._chk(lib().dg_batch_autotune(self._h, int(reps), out, ms.ctypes.data_as(_dp)))
        return list(out), ms.tolist()

