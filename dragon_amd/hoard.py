"""Frames and restarts of device-resident grids in the reference's HDF5
format -- the DRAGONHOARD module (DRAGONHOARD/HDF5Output.cpp,
HDF5Restart.cpp, OutputDir.cpp; switches of DRAGONHOARD_Config.h).

``writeToFile(grid, t, cycle, filename)`` and ``loadFromFile(grid, filename)``
keep the reference's names, file layout, keys and types.  The state goes
through ``dg_snapshot`` (a device transpose into file order, then an
asynchronous device-to-host copy) and ``dg_restore``; the HDF5 work is host
code in ``libdragon_hoard.so`` (``include/dragon_hoard.h``).
``FrameWriter`` writes frames in a background thread while the grid keeps
stepping.
"""
from __future__ import annotations

import ctypes as C
import os
import re
import threading

import numpy as np

from . import HERE, DragonError, lib

HOARD_PATH = os.path.join(HERE, "libdragon_hoard.so")

# DRAGONHOARD_Config.h:16-37
output_base_name = "frame"
output_dir = ""
RESTART_FROM_FILE = True
RESTART_FRAME = 0
HDF5_COMPRESSION_LEVEL = 4
HDF5_WRITE_PRIMITIVE, HDF5_WRITE_CONSERVATIVE = 1, 6
HDF5_WRITE_PRIMITIVE_AND_ENERGY, HDF5_WRITE_PRIMITIVE_AND_CONSERVATIVE = 3, 7
HDF5_WRITE_OPTION = HDF5_WRITE_PRIMITIVE_AND_ENERGY
HDF5_WRITE_OMIT, HDF5_WRITE_FLOAT, HDF5_WRITE_DOUBLE = 0, 1, 2
HDF5_REDUNDANT_VALS_OPTION = HDF5_WRITE_FLOAT
WRITE_GHOSTS_TO_FILE = False
file_ext = ".h5"  # HDF5_Attrs.hpp:59


class _Frame(C.Structure):
    _fields_ = [("dim", C.c_int), ("mhd", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("ng", C.c_int),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double), ("gamma", C.c_double),
                ("time", C.c_double), ("cycle", C.c_int), ("write_option", C.c_int), ("redundant", C.c_int),
                ("compression", C.c_int)]


_hl = None


def hoard_lib():
    """libdragon_hoard.so (HDF5 C library of the image); raises if not built."""
    global _hl
    if _hl is None:
        if not os.path.exists(HOARD_PATH):
            raise ImportError(f"dragon_amd.hoard: {HOARD_PATH} is missing -- build it (make -C dragon_amd/csrc)")
        L = C.CDLL(HOARD_PATH)
        fp = C.POINTER(_Frame)
        sig = {
            "dgh_cell_count": (C.c_long, [fp]), "dgh_face_count": (C.c_long, [fp]),
            "dgh_write": (C.c_int, [C.c_char_p, fp, C.c_void_p, C.c_void_p, C.c_char_p, C.c_int]),
            "dgh_read_meta": (C.c_int, [C.c_char_p, fp, C.c_char_p, C.c_int]),
            "dgh_read": (C.c_int, [C.c_char_p, fp, fp, C.c_void_p, C.c_void_p, C.c_char_p, C.c_int]),
        }
        for n, (r, a) in sig.items():
            f = getattr(L, n)
            f.restype = r
            f.argtypes = a
        _hl = L
    return _hl


class HoardError(RuntimeError):
    """A frame could not be written or read; the message is the reference's where it has one."""


# ---------------------------------------------------------------- names
def cycle_string(n: int) -> str:
    """HDF5Output.cpp:33-36: the cycle number zero-padded to five digits."""
    zeroes = "0000" if n < 10 else ("000" if n < 100 else ("00" if n < 1000 else ("0" if n < 10000 else "")))
    return zeroes + str(n)


def frame_path(filename: str) -> str:
    """output_dir + "/" + filename + ".h5" (HDF5Output.cpp:134, 246, 389)."""
    return output_dir + "/" + filename + file_ext


def check_extension(filename: str) -> str:
    """HDF5Restart.cpp:54-58."""
    if len(filename) <= len(file_ext):
        return filename + file_ext
    return filename if filename.endswith(file_ext) else filename + file_ext


def verifyOutputDirectory():
    """OutputDir.cpp:18-27."""
    d = output_dir
    if os.path.exists(d):
        if not os.path.isdir(d):
            raise HoardError(f"{d} exists but is not a directory")
    else:
        os.makedirs(d)


def _frame_number(name: str) -> int:
    """OutputDir.cpp:34-46: the digits after the last '_' before the first '.'."""
    end = name.find(".")
    if end < 0:
        end = len(name)
    pos = name.rfind("_", 0, end)
    if pos < 0 or pos + 1 >= end or not re.fullmatch(r"[0-9]+", name[pos + 1:end]):
        return -1
    return int(name[pos + 1:end])


def restartFileName() -> str:
    """OutputDir.cpp:49-72: the file of frame RESTART_FRAME in output_dir, or
    (RESTART_FRAME < 0) the highest-numbered frame; "" if none."""
    best, best_n = "", -1
    for name in sorted(os.listdir(output_dir)):
        if not os.path.isfile(os.path.join(output_dir, name)):
            continue
        dot = name.rfind(".")
        if (name[dot:] if dot >= 0 else "") != file_ext:
            continue
        n = _frame_number(name)
        if RESTART_FRAME < 0:
            if n > best_n:
                best, best_n = name, n
        elif n == RESTART_FRAME:
            return name
    return best


# ---------------------------------------------------------------- arrays
def frame_meta(grid, t=0.0, cycle=0, ng=None, write_option=None, redundant=None, compression=None) -> _Frame:
    f = _Frame()
    f.dim = grid.dim
    f.mhd = int(bool(grid.config.mhd))
    n = list(grid.shape) + [1] * (3 - grid.dim)
    f.nx, f.ny, f.nz = n
    f.ng = (grid.g if WRITE_GHOSTS_TO_FILE else 0) if ng is None else int(ng)
    d = list(grid.d) + [1.0] * (3 - grid.dim)
    f.dx, f.dy, f.dz = d
    f.gamma = grid.config.gamma
    f.time, f.cycle = float(t), int(cycle)
    f.write_option = HDF5_WRITE_OPTION if write_option is None else write_option
    f.redundant = HDF5_REDUNDANT_VALS_OPTION if redundant is None else redundant
    f.compression = HDF5_COMPRESSION_LEVEL if compression is None else compression
    return f


def array_sizes(f: _Frame):
    """(cell count, face count) of the x-fastest arrays of a frame."""
    L = hoard_lib()
    faces = L.dgh_face_count(C.byref(f)) if (f.mhd and f.dim >= 2) else 0
    return L.dgh_cell_count(C.byref(f)), faces


def write_arrays(path: str, f: _Frame, cells: np.ndarray, faces: np.ndarray | None = None):
    """dgh_write: cells (8, count) and faces (3, count) in file order (x fastest)."""
    cells = np.ascontiguousarray(cells, dtype=np.float64)
    fa = None if faces is None else np.ascontiguousarray(faces, dtype=np.float64)
    err = C.create_string_buffer(512)
    rc = hoard_lib().dgh_write(path.encode(), C.byref(f), cells.ctypes.data, None if fa is None else fa.ctypes.data, err, 512)
    if rc:
        raise HoardError(err.value.decode())


def read_meta(path: str) -> _Frame:
    f = _Frame()
    err = C.create_string_buffer(512)
    if hoard_lib().dgh_read_meta(path.encode(), C.byref(f), err, 512):
        raise HoardError(err.value.decode())
    return f


def read_arrays(path: str, want: _Frame):
    """dgh_read for a receiving grid described by `want` (its ng = the grid's
    ghosts): (file metadata, cells (8, count), faces (3, count) or None)."""
    got = _Frame()
    err = C.create_string_buffer(512)
    # sizes: the clamped ng is at most want.ng; read the metadata first
    meta = read_meta(path)
    probe = _Frame.from_buffer_copy(want)
    probe.ng = min(meta.ng, want.ng)
    nc, nf = array_sizes(probe)
    cells = np.zeros((8, nc))
    faces = np.zeros((3, nf)) if nf else None
    rc = hoard_lib().dgh_read(path.encode(), C.byref(want), C.byref(got), cells.ctypes.data,
                              None if faces is None else faces.ctypes.data, err, 512)
    if rc:
        raise HoardError(err.value.decode())
    return got, cells, faces


class _Pinned:
    """Pinned host buffer (dg_host_alloc) viewed as float64."""

    def __init__(self, count):
        p = C.c_void_p()
        if lib().dg_host_alloc(max(1, count) * 8, C.byref(p)):
            raise DragonError(8, "dg_host_alloc failed")
        self.p = p
        self.a = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_double)), shape=(max(1, count),))

    def __del__(self):
        p = getattr(self, "p", None)
        if p:
            lib().dg_host_free(p)
            self.p = None


def _snapshot(grid, ng, cells: np.ndarray, faces: np.ndarray | None):
    rc = lib().dg_snapshot(grid._h, int(ng), cells.ctypes.data, None if faces is None else faces.ctypes.data)
    if rc:
        raise DragonError(rc, lib().dg_last_error(grid._h).decode())


# ------------------------------------------------------- writeToFile / loadFromFile
def writeToFile(grid, t: float, cycle: int, filename: str, **options):
    """DRAGONHOARD::writeToFile (HDF5Output.cpp:103-525): the grid's cells
    (and face B) to output_dir/filename.h5; options override the switches of
    DRAGONHOARD_Config.h (write_option, redundant, compression, ng)."""
    f = frame_meta(grid, t, cycle, **options)
    nc, nf = array_sizes(f)
    cells = np.empty((8, nc))
    faces = np.empty((3, nf)) if nf else None
    _snapshot(grid, f.ng, cells, faces)
    rc = lib().dg_snapshot_wait(grid._h)
    if rc:
        raise DragonError(rc, lib().dg_last_error(grid._h).decode())
    write_arrays(frame_path(filename), f, cells, faces)


def loadFromFile(grid, filename: str):
    """DRAGONHOARD::loadFromFile (HDF5Restart.cpp:61-374): returns (t, cycle);
    the grid takes the file's cells (and face B) and its dx/dy/dz."""
    path = output_dir + "/" + check_extension(filename)
    want = frame_meta(grid, ng=grid.g)
    got, cells, faces = read_arrays(path, want)
    d = [got.dx, got.dy, got.dz]
    rc = lib().dg_set_spacing(grid._h, *d)
    if rc:
        raise DragonError(rc, "dg_set_spacing failed")
    grid.d = tuple(d[: grid.dim])
    grid.dx = got.dx
    if grid.dim >= 2:
        grid.dy = got.dy
    if grid.dim >= 3:
        grid.dz = got.dz
    rc = lib().dg_restore(grid._h, int(got.ng), cells.ctypes.data, None if faces is None else faces.ctypes.data)
    if rc:
        raise DragonError(rc, lib().dg_last_error(grid._h).decode())
    return got.time, got.cycle


class FrameWriter:
    """Frames written while the grid keeps stepping: ``submit`` starts the
    device snapshot into one of two pinned buffers and returns; a background
    thread waits for the copy and writes the file (the HDF5 call releases the
    GIL; the image's HDF5 library is built thread-safe, so the two writer
    threads serialise inside it).  ``close`` (or leaving a ``with`` block)
    waits for the last file."""

    def __init__(self, grid, **options):
        self.grid = grid
        self.options = options
        f = frame_meta(grid, **options)
        self.nc, self.nf = array_sizes(f)
        self.bufs = [(_Pinned(8 * self.nc), _Pinned(3 * self.nf) if self.nf else None) for _ in range(2)]
        self.turn = 0
        self.threads = [None, None]
        self.errors = []

    def submit(self, t: float, cycle: int, filename: str):
        k = self.turn
        self.turn ^= 1
        if self.threads[k] is not None:  # the buffer's previous file must be written first
            self.threads[k].join()
        self._raise()
        cb, fb = self.bufs[k]
        cells = cb.a[: 8 * self.nc].reshape(8, self.nc)
        faces = fb.a[: 3 * self.nf].reshape(3, self.nf) if fb is not None else None
        f = frame_meta(self.grid, t, cycle, **self.options)
        _snapshot(self.grid, f.ng, cells, faces)
        path = frame_path(filename)

        def work():
            try:
                rc = lib().dg_snapshot_wait(self.grid._h)
                if rc:
                    raise DragonError(rc, lib().dg_last_error(self.grid._h).decode())
                write_arrays(path, f, cells, faces)
            except Exception as e:  # reported by the next submit or close
                self.errors.append(e)

        th = threading.Thread(target=work, daemon=True)
        th.start()
        self.threads[k] = th

    def _raise(self):
        if self.errors:
            raise self.errors.pop(0)

    def close(self):
        for th in self.threads:
            if th is not None:
                th.join()
        self.threads = [None, None]
        self._raise()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
