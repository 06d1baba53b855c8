"""CPU oracle for DRAGON's unsplit Godunov step -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg may import this package, and only as the checker (or the timed CPU
baseline).  The product path (``dragon_amd``) never loads it.

``liboracle.so`` is our C++ restatement of the reference (oracle/*.cpp).
``_ref/libref_<variant>.so`` (built by ``oracle/build_ref.sh`` only where the
reference tree is present) wraps the reference's own per-function sources and
is used to pin the restatement.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

# Config.h enums
EXACT, HLL, HLLC, HLLD, HLLE, HLLX = 0, 1, 2, 3, 4, 5
MINMOD, MC, VANLEER, SUPERBEE, VANALBADA = 0, 1, 2, 3, 4
CT_TOTAL_E, CT_THERMAL, CT_BETA_GATED = 0, 1, 2
CFL_MAX, CFL_ADD = 0, 1
# Boundary kinds and face bits (Boundary.hpp:139-149)
OUTFLOW, PERIODIC, REFLECTIVE, FIXED, IGNORE = 0, 1, 2, 3, 4
XN, XP, YN, YP, ZN, ZP = 1, 2, 4, 8, 16, 32
FX, FY, FZ, FALL = 3, 12, 48, 63


class _CCfg(C.Structure):
    _fields_ = [("mhd", C.c_int), ("gamma", C.c_double), ("solver", C.c_int), ("limiter", C.c_int),
                ("muscl", C.c_int), ("ctu", C.c_int), ("ct_energy", C.c_int), ("ct_beta", C.c_double),
                ("cfl_mode", C.c_int), ("cfl_coeff", C.c_double), ("verify_fallback", C.c_int),
                ("fallback_try_hlle", C.c_int), ("fallback_param", C.c_double), ("hlld_safety", C.c_int),
                ("threads", C.c_int)]


@dataclass
class Cfg:
    """Runtime form of the reference's Config.h switches (production defaults)."""
    mhd: int = 1
    gamma: float = 5.0 / 3.0
    solver: int = HLLX
    limiter: int = MC
    muscl: int = 1
    ctu: int = 1
    ct_energy: int = CT_BETA_GATED
    ct_beta: float = 0.1
    cfl_mode: int = CFL_ADD
    cfl_coeff: float = 0.3
    verify_fallback: int = -1      # -1: on iff hydro (Config.h:57-61)
    fallback_try_hlle: int = 0
    fallback_param: float = 1.0
    hlld_safety: int = 1
    threads: int = 1

    def c(self) -> _CCfg:
        return _CCfg(*[getattr(self, f[0]) for f in _CCfg._fields_])


_dp = C.POINTER(C.c_double)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _load():
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
    lib = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    cp = C.POINTER(_CCfg)
    sig = {
        "oc_grid_create": (vp, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, cp]),
        "oc_grid_destroy": (None, [vp]),
        "oc_grid_ghosts": (C.c_int, [vp]),
        "oc_grid_set_threads": (None, [vp, C.c_int]),
        "oc_grid_get_w": (None, [vp, _dp]), "oc_grid_set_w": (None, [vp, _dp]),
        "oc_grid_get_B": (None, [vp, _dp]), "oc_grid_set_B": (None, [vp, _dp]),
        "oc_grid_clear_bc": (None, [vp]),
        "oc_grid_add_bc": (None, [vp, C.c_int, C.c_int, C.c_int, C.c_int, _dp]),
        "oc_grid_apply_bc": (None, [vp]), "oc_grid_init_B": (None, [vp]),
        "oc_grid_cfl": (C.c_int, [vp, _dp, C.c_char_p, C.c_int]),
        "oc_grid_step": (C.c_int, [vp, C.c_double, C.c_char_p, C.c_int]),
        "oc_grid_advance": (C.c_long, [vp, C.c_double, C.c_int, C.c_long, C.POINTER(C.c_long), C.c_char_p, C.c_int]),
        "oc_grid_split_step": (C.c_int, [vp, C.c_double, C.c_char_p, C.c_int]),
        "oc_grid_advance_split": (C.c_long, [vp, C.c_double, C.c_int, C.c_long, C.POINTER(C.c_long), C.c_char_p, C.c_int]),
        "oc_grid_faraday_init": (None, [vp, _dp]),
        "oc_riemann_flux": (C.c_int, [cp, _dp, _dp, C.c_double, C.c_int, _dp, C.c_char_p, C.c_int]),
        "oc_riemann_solver": (None, [cp, _dp, _dp, C.c_int, C.c_double, C.c_double, _dp]),
        "oc_exact_star": (None, [cp, _dp, _dp, _dp]),
        "oc_exact_sample": (None, [cp, _dp, _dp, C.c_double, _dp]),
        "oc_exact_f": (C.c_double, [cp, C.c_double, _dp]),
        "oc_muscl": (None, [cp, _dp, _dp, _dp, C.c_double, _dp, _dp, _dp]),
        "oc_limit": (C.c_double, [cp, C.c_double, C.c_double]),
        "oc_to_cons": (None, [cp, _dp, _dp]), "oc_to_prim": (None, [cp, _dp, _dp]),
        "oc_flux_prim": (None, [cp, _dp, _dp]), "oc_speeds": (None, [cp, _dp, _dp]),
        "oc_is_physical": (C.c_int, [cp, _dp]),
        "oc_upwind_corr": (C.c_double, [C.c_double] * 5),
        "oc_protect_thermal": (C.c_int, [cp, _dp]),
        "oc_face_mask": (C.c_int, [C.c_char_p]),
        "oc_ct_electric": (None, [cp] + [C.c_int] * 5 + [_dp, _dp, _dp, _dp, C.c_int, _dp]),
        "oc_ct_body_electric": (None, [cp] + [C.c_int] * 5 + [_dp, C.c_int, _dp]),
        "oc_ct_faraday": (None, [cp] + [C.c_int] * 5 + [_dp, C.c_double, C.c_double, C.c_double, C.c_int, _dp]),
        "oc_ct_body_fields": (None, [cp] + [C.c_int] * 5 + [_dp, _dp]),
        "oc_cfl_speed": (C.c_double, [cp, _dp, C.c_double, C.c_double, C.c_double]),
        "oc_correct_state": (None, [cp] + [C.c_int] * 5 + [_dp, _dp, _dp, C.c_double, C.c_int, C.c_int]),
        "oc_ctu_flux": (C.c_int, [cp] + [C.c_int] * 5 + [_dp, _dp, _dp, _dp, C.c_double, C.c_double, C.c_int, C.c_char_p,
                                                       C.c_int]),
    }
    for n, (r, a) in sig.items():
        f = getattr(lib, n)
        f.restype = r
        f.argtypes = a
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def s8(rho=1.0, p=1.0, vx=0.0, vy=0.0, vz=0.0, Bx=0.0, By=0.0, Bz=0.0) -> np.ndarray:
    """A primitive state as the 8-double record (rho, p, vx, vy, vz, Bx, By, Bz)."""
    return np.array([rho, p, vx, vy, vz, Bx, By, Bz], dtype=np.float64)


class OracleError(RuntimeError):
    pass


class Grid:
    """Oracle grid: Grid1D/2D/3D with the reference's ghost layout."""

    def __init__(self, shape, d, cfg: Cfg, ghosts: int = 0):
        shape = tuple(int(s) for s in shape)
        self.dim = len(shape)
        self.shape = shape
        n = list(shape) + [1] * (3 - self.dim)
        dd = list(d) + [1.0] * (3 - len(d))
        self.cfg = cfg
        self._c = cfg.c()
        self.h = lib().oc_grid_create(self.dim, n[0], n[1], n[2], dd[0], dd[1], dd[2], ghosts, C.byref(self._c))
        self.g = lib().oc_grid_ghosts(self.h)
        self.d = tuple(dd[: self.dim])
        self.has_B = bool(cfg.mhd) and self.dim >= 2

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            lib().oc_grid_destroy(h)
            self.h = None

    @property
    def wshape(self):
        return tuple(s + 2 * self.g for s in self.shape) + (8,)

    @property
    def bshape(self):
        return tuple(s + 1 + 2 * self.g for s in self.shape) + (3,)

    def get_w(self) -> np.ndarray:
        a = np.zeros(self.wshape)
        lib().oc_grid_get_w(self.h, _ptr(a))
        return a

    def set_w(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.shape == self.wshape
        lib().oc_grid_set_w(self.h, _ptr(a))

    def get_B(self) -> np.ndarray:
        a = np.zeros(self.bshape)
        if self.has_B:
            lib().oc_grid_get_B(self.h, _ptr(a))
        return a

    def set_B(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.shape == self.bshape
        lib().oc_grid_set_B(self.h, _ptr(a))

    def set_threads(self, t: int):
        lib().oc_grid_set_threads(self.h, int(t))

    def add_bc(self, kind, faces=FALL, corners=True, gated=False, state=None):
        st = None if state is None else _ptr(np.ascontiguousarray(state, dtype=np.float64))
        lib().oc_grid_add_bc(self.h, kind, faces, int(corners), int(gated), st)

    def clear_bc(self):
        lib().oc_grid_clear_bc(self.h)

    def apply_bc(self):
        lib().oc_grid_apply_bc(self.h)

    def init_B(self):
        lib().oc_grid_init_B(self.h)

    def faraday_init(self, A: np.ndarray):
        A = np.ascontiguousarray(A, dtype=np.float64)
        lib().oc_grid_faraday_init(self.h, _ptr(A))

    def cfl(self) -> float:
        out = np.zeros(1)
        err = C.create_string_buffer(256)
        if lib().oc_grid_cfl(self.h, _ptr(out), err, 256):
            raise OracleError(err.value.decode())
        return float(out[0])

    def step(self, dt: float):
        err = C.create_string_buffer(256)
        if lib().oc_grid_step(self.h, dt, err, 256):
            raise OracleError(err.value.decode())

    def advance(self, dt: float, check_cfl=True, max_steps=-1):
        """Grid::advance_unsplit; returns (accepted steps, failed attempts)."""
        err = C.create_string_buffer(256)
        fails = C.c_long(0)
        n = lib().oc_grid_advance(self.h, dt, int(check_cfl), int(max_steps), C.byref(fails), err, 256)
        if n < 0:
            raise OracleError(err.value.decode())
        return int(n), int(fails.value)

    def split_step(self, dt: float):
        err = C.create_string_buffer(256)
        if lib().oc_grid_split_step(self.h, dt, err, 256):
            raise OracleError(err.value.decode())

    def advance_split(self, dt: float, check_cfl=True, max_steps=-1):
        """Grid::advance_split (Strang splitting, hydro); returns (accepted steps, failed attempts)."""
        err = C.create_string_buffer(256)
        fails = C.c_long(0)
        n = lib().oc_grid_advance_split(self.h, dt, int(check_cfl), int(max_steps), C.byref(fails), err, 256)
        if n < 0:
            raise OracleError(err.value.decode())
        return int(n), int(fails.value)


# ------------------------------------------------------------ per-function
def riemann_flux(cfg: Cfg, L, R, dt_dx=0.0, direction=0):
    F = np.zeros(8)
    err = C.create_string_buffer(256)
    c = cfg.c()
    rc = lib().oc_riemann_flux(C.byref(c), _ptr(np.asarray(L, float)), _ptr(np.asarray(R, float)), dt_dx, direction,
                               _ptr(F), err, 256)
    if rc:
        raise OracleError(err.value.decode())
    return F


SOLVER_IDS = {"HLL": 0, "HLLE": 1, "HLLC": 2, "HLLD": 3, "EXACT": 4, "HLL_S": 5, "HLLC_S": 6}


def riemann_solver(cfg: Cfg, L, R, which: str, sl=0.0, sr=0.0):
    F = np.zeros(8)
    c = cfg.c()
    lib().oc_riemann_solver(C.byref(c), _ptr(np.asarray(L, float)), _ptr(np.asarray(R, float)), SOLVER_IDS[which],
                            sl, sr, _ptr(F))
    return F


def exact_star(cfg: Cfg, L, R):
    out = np.zeros(5)
    c = cfg.c()
    lib().oc_exact_star(C.byref(c), _ptr(np.asarray(L, float)), _ptr(np.asarray(R, float)), _ptr(out))
    return out


def exact_sample(cfg: Cfg, L, R, x_t):
    out = np.zeros(8)
    c = cfg.c()
    lib().oc_exact_sample(C.byref(c), _ptr(np.asarray(L, float)), _ptr(np.asarray(R, float)), x_t, _ptr(out))
    return out


def muscl(cfg: Cfg, wL, wC, wR, dt_dL, dB=(0.0, 0.0, 0.0)):
    Lo, Ro = np.zeros(8), np.zeros(8)
    c = cfg.c()
    lib().oc_muscl(C.byref(c), _ptr(np.asarray(wL, float)), _ptr(np.asarray(wC, float)), _ptr(np.asarray(wR, float)),
                   dt_dL, _ptr(np.asarray(dB, float)), _ptr(Lo), _ptr(Ro))
    return Lo, Ro


def limit(cfg: Cfg, a, b):
    c = cfg.c()
    return lib().oc_limit(C.byref(c), a, b)


def to_cons(cfg: Cfg, w):
    u = np.zeros(8)
    c = cfg.c()
    lib().oc_to_cons(C.byref(c), _ptr(np.asarray(w, float)), _ptr(u))
    return u


def to_prim(cfg: Cfg, u):
    w = np.zeros(8)
    c = cfg.c()
    lib().oc_to_prim(C.byref(c), _ptr(np.asarray(u, float)), _ptr(w))
    return w


def flux_prim(cfg: Cfg, w):
    F = np.zeros(8)
    c = cfg.c()
    lib().oc_flux_prim(C.byref(c), _ptr(np.asarray(w, float)), _ptr(F))
    return F


def speeds(cfg: Cfg, w):
    """(cs, c_alfven, c_fast_x, c_fast_max, enthalpy, energy)."""
    out = np.zeros(6)
    c = cfg.c()
    lib().oc_speeds(C.byref(c), _ptr(np.asarray(w, float)), _ptr(out))
    return out


def is_physical(cfg: Cfg, w) -> bool:
    c = cfg.c()
    return bool(lib().oc_is_physical(C.byref(c), _ptr(np.asarray(w, float))))


def upwind_corr(m, fm, fp, rm, rp):
    return lib().oc_upwind_corr(m, fm, fp, rm, rp)


def protect_thermal(cfg: Cfg, w) -> bool:
    c = cfg.c()
    return bool(lib().oc_protect_thermal(C.byref(c), _ptr(np.asarray(w, float))))


def face_mask(s: str) -> int:
    return lib().oc_face_mask(s.encode())


def cfl_speed(cfg: Cfg, w, dx, dy=0.0, dz=0.0):
    c = cfg.c()
    return lib().oc_cfl_speed(C.byref(c), _ptr(np.asarray(w, float)), dx, dy, dz)


def _shape5(shp):
    s = list(shp) + [1] * (3 - len(shp))
    return len(shp), s[0], s[1], s[2]


def ct_electric(cfg: Cfg, shp, g, FX, FY, FZ, ref, gg, E):
    """computeElectric (+ upwindElectric if ref is not None) into E (in place)."""
    c = cfg.c()
    d, nx, ny, nz = _shape5(shp)
    FZ = FY if FZ is None else FZ
    lib().oc_ct_electric(C.byref(c), d, nx, ny, nz, g, _ptr(FX), _ptr(FY), _ptr(FZ), None if ref is None else _ptr(ref), gg,
                         _ptr(E))


def ct_body_electric(cfg: Cfg, shp, g, W, gg, E):
    c = cfg.c()
    d, nx, ny, nz = _shape5(shp)
    lib().oc_ct_body_electric(C.byref(c), d, nx, ny, nz, g, _ptr(W), gg, _ptr(E))


def ct_faraday(cfg: Cfg, shp, g, E, dts, gg, B):
    c = cfg.c()
    d, nx, ny, nz = _shape5(shp)
    dts = list(dts) + [0.0] * (3 - len(dts))
    lib().oc_ct_faraday(C.byref(c), d, nx, ny, nz, g, _ptr(E), dts[0], dts[1], dts[2], gg, _ptr(B))


def ct_body_fields(cfg: Cfg, shp, g, B, W):
    c = cfg.c()
    d, nx, ny, nz = _shape5(shp)
    lib().oc_ct_body_fields(C.byref(c), d, nx, ny, nz, g, _ptr(B), _ptr(W))



def correct_state(cfg: Cfg, shp, g, L, R, F, dt_dL, sdim, fdim):
    """correctState (CTU.cpp:264-301): L and R updated in place.  2D takes the
    flux direction fdim only (sdim is ignored)."""
    c = cfg.c()
    d, nx, ny, nz = _shape5(shp)
    lib().oc_correct_state(C.byref(c), d, nx, ny, nz, g, _ptr(L), _ptr(R), _ptr(F), dt_dL, sdim, fdim)


def ctu_flux(cfg: Cfg, axis, shp, g, L, R, T, F, dt_d, dt_t, dimt):
    """computeCTUFlux_X/Y/Z (CTU.cpp:306-370): F updated in place."""
    c = cfg.c()
    _, nx, ny, nz = _shape5(shp)
    err = C.create_string_buffer(256)
    if lib().oc_ctu_flux(C.byref(c), axis, nx, ny, nz, g, _ptr(L), _ptr(R), _ptr(T), _ptr(F), dt_d, dt_t, dimt, err, 256):
        raise OracleError(err.value.decode())


# ------------------------------------------------------------ reference
REF_DIR = os.path.join(HERE, "_ref")


class RefLib:
    """The reference's own per-function code (oracle/_ref), if it was built."""

    def __init__(self, variant: str):
        path = os.path.join(REF_DIR, f"libref_{variant}.so")
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.lib = L = C.CDLL(path)
        sig = {
            "rf_is_mhd": (C.c_int, []), "rf_is_testmode": (C.c_int, []), "rf_gamma": (C.c_double, []),
            "rf_set": (None, [C.c_int] * 4),
            "rf_riemann_flux": (C.c_int, [_dp, _dp, C.c_double, C.c_int, _dp, C.c_char_p, C.c_int]),
            "rf_riemann_solver": (None, [_dp, _dp, C.c_int, C.c_double, C.c_double, _dp]),
            "rf_exact_star": (None, [_dp, _dp, _dp]),
            "rf_exact_sample": (None, [_dp, _dp, C.c_double, _dp]),
            "rf_exact_f": (C.c_double, [C.c_double, _dp]),
            "rf_muscl": (None, [_dp, _dp, _dp, C.c_double, _dp, _dp, _dp]),
            "rf_limit": (C.c_double, [C.c_double, C.c_double]),
            "rf_to_cons": (None, [_dp, _dp]), "rf_to_prim": (None, [_dp, _dp]),
            "rf_flux_prim": (None, [_dp, _dp]), "rf_speeds": (None, [_dp, _dp]),
            "rf_is_physical": (C.c_int, [_dp]),
            "rf_cfl_speed": (C.c_double, [_dp, C.c_double, C.c_double, C.c_double]),
        }
        if L.rf_is_mhd():
            sig.update({
                "rf_electric2D": (None, [C.c_int] * 3 + [_dp, _dp, _dp, C.c_int, _dp]),
                "rf_electric3D": (None, [C.c_int] * 4 + [_dp, _dp, _dp, _dp, C.c_int, _dp]),
                "rf_body_electric2D": (None, [C.c_int] * 3 + [_dp, C.c_int, _dp]),
                "rf_body_electric3D": (None, [C.c_int] * 4 + [_dp, C.c_int, _dp]),
                "rf_faraday2D": (None, [C.c_int] * 3 + [_dp, C.c_double, C.c_double, C.c_int, _dp]),
                "rf_faraday3D": (None, [C.c_int] * 4 + [_dp, C.c_double, C.c_double, C.c_double, C.c_int, _dp]),
                "rf_body_fields2D": (None, [C.c_int] * 3 + [_dp, _dp]),
                "rf_body_fields3D": (None, [C.c_int] * 4 + [_dp, _dp]),
            })
        else:
            sig.update({
                "rf_correct_state2D": (None, [C.c_int] * 3 + [_dp, _dp, _dp, C.c_double, C.c_int]),
                "rf_correct_state3D": (None, [C.c_int] * 4 + [_dp, _dp, _dp, C.c_double, C.c_int, C.c_int]),
                "rf_ctu_flux3D": (C.c_int, [C.c_int] * 5 + [_dp, _dp, _dp, _dp, C.c_double, C.c_double, C.c_int,
                                                             C.c_char_p, C.c_int]),
            })
        for n, (r, a) in sig.items():
            f = getattr(L, n)
            f.restype = r
            f.argtypes = a
        self.mhd = bool(L.rf_is_mhd())
        self.testmode = bool(L.rf_is_testmode())
        self.gamma = L.rf_gamma()

    def set(self, riemann=HLLX, limiter=MINMOD, cfl=CFL_ADD, ct=CT_TOTAL_E):
        self.lib.rf_set(riemann, limiter, cfl, ct)

    def riemann_flux(self, L, R, dt_dx=0.0, direction=0):
        F = np.zeros(8)
        err = C.create_string_buffer(256)
        if self.lib.rf_riemann_flux(_ptr(np.asarray(L, float)), _ptr(np.asarray(R, float)), dt_dx, direction,
                                    _ptr(F), err, 256):
            raise OracleError(err.value.decode())
        return F

    def riemann_solver(self, L, R, which: str, sl=0.0, sr=0.0):
        F = np.zeros(8)
        self.lib.rf_riemann_solver(_ptr(np.asarray(L, float)), _ptr(np.asarray(R, float)), SOLVER_IDS[which], sl, sr,
                                   _ptr(F))
        return F

    def exact_star(self, L, R):
        out = np.zeros(4)
        self.lib.rf_exact_star(_ptr(np.asarray(L, float)), _ptr(np.asarray(R, float)), _ptr(out))
        return out

    def exact_sample(self, L, R, x_t):
        out = np.zeros(8)
        self.lib.rf_exact_sample(_ptr(np.asarray(L, float)), _ptr(np.asarray(R, float)), x_t, _ptr(out))
        return out

    def muscl(self, wL, wC, wR, dt_dL, dB=(0.0, 0.0, 0.0)):
        Lo, Ro = np.zeros(8), np.zeros(8)
        self.lib.rf_muscl(_ptr(np.asarray(wL, float)), _ptr(np.asarray(wC, float)), _ptr(np.asarray(wR, float)),
                          dt_dL, _ptr(np.asarray(dB, float)), _ptr(Lo), _ptr(Ro))
        return Lo, Ro

    def limit(self, a, b):
        return self.lib.rf_limit(a, b)

    def to_cons(self, w):
        u = np.zeros(8)
        self.lib.rf_to_cons(_ptr(np.asarray(w, float)), _ptr(u))
        return u

    def to_prim(self, u):
        w = np.zeros(8)
        self.lib.rf_to_prim(_ptr(np.asarray(u, float)), _ptr(w))
        return w

    def flux_prim(self, w):
        F = np.zeros(8)
        self.lib.rf_flux_prim(_ptr(np.asarray(w, float)), _ptr(F))
        return F

    def speeds(self, w):
        out = np.zeros(6)
        self.lib.rf_speeds(_ptr(np.asarray(w, float)), _ptr(out))
        return out

    def is_physical(self, w):
        return bool(self.lib.rf_is_physical(_ptr(np.asarray(w, float))))

    def cfl_speed(self, w, dx, dy=0.0, dz=0.0):
        return self.lib.rf_cfl_speed(_ptr(np.asarray(w, float)), dx, dy, dz)


def ref_available(variant: str) -> bool:
    return os.path.exists(os.path.join(REF_DIR, f"libref_{variant}.so"))
