"""File formats of magudi's workflows over the library's host-side readers / writers.

* PLOT3D whole-format grid (.xyz), solution (.q) and function (.f) files
  (src/PLOT3DFormat.c, src/PLOT3DHelperImpl.f90), either byte order, multi block.
* Actuator control-forcing / gradient files (.dat, src/ActuatorPatchImpl.f90:226-458)
  and ZAXPY (bin/ZAXPY.f90).
* bc.dat boundary-condition tables (src/RegionImpl.f90:175-337) -> cases.Patch rows.

Arrays use the library's layout: (components, grid points) with i fastest.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import capi
from .capi import check, dptr
from .cases import PATCH_TYPE_NAMES, Patch

GRID, SOLUTION, FUNCTION = 0, 1, 2


def _b(path: str) -> bytes:
    return os.fsencode(path)


def _ip(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_int))


def _lib():
    L = capi.lib()
    if not getattr(L, "_plot3d_sigs", False):
        ip, dp, cp = C.POINTER(C.c_int), C.POINTER(C.c_double), C.c_char_p
        sigs = {
            "mgd_plot3d_detect": [cp, C.c_int, ip, ip, ip, ip, ip, ip, C.c_int],
            "mgd_plot3d_read_grid": [cp, C.c_int, dp, ip],
            "mgd_plot3d_read_solution": [cp, C.c_int, C.c_int, dp, dp],
            "mgd_plot3d_read_function": [cp, C.c_int, C.c_int, dp],
            "mgd_plot3d_write_grid": [cp, C.c_int, ip, dp, ip],
            "mgd_plot3d_write_solution": [cp, ip, C.c_int, dp, dp],
            "mgd_plot3d_write_function": [cp, ip, C.c_int, dp],
            "mgd_control_write": [cp, C.c_int, C.c_int, dp],
            "mgd_control_read": [cp, C.c_int, C.c_int, dp],
            "mgd_control_zaxpy": [cp, C.c_double, cp, cp],
            "mgd_plot3d_set_record_markers": [C.c_int],
        }
        for name, args in sigs.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int
        L._plot3d_sigs = True
    return L


def set_record_markers(nbytes: int):
    """Record-marker width the writers use: 4 (default) or 8 bytes (USE_EXTENDED_PLOT3D)."""
    check(_lib().mgd_plot3d_set_record_markers(nbytes))


def detect(path: str, include_function: bool = True, max_grids: int = 64) -> Dict:
    """plot3dDetectFormat: number of grids, file type, sizes, endianness, iblank."""
    v = [C.c_int(0) for _ in range(5)]
    sizes = np.zeros(3 * max_grids, dtype=np.intc)
    check(_lib().mgd_plot3d_detect(_b(path), int(include_function), *[C.byref(x) for x in v],
                                   _ip(sizes), max_grids))
    n_grids, ftype, nscal, iblank, native = (x.value for x in v)
    return {"n_grids": n_grids, "file_type": ftype, "n_scalars": nscal,
            "has_iblank": bool(iblank), "native_endian": bool(native),
            "sizes": [tuple(int(s) for s in sizes[3 * g:3 * g + 3]) for g in range(min(n_grids, max_grids))]}


def _npts(path, block):
    info = detect(path)
    nx, ny, nz = info["sizes"][block]
    return info, nx * ny * nz


def read_grid(path: str, block: int = 0):
    """(coordinates (3, N), iblank (N,))."""
    _, N = _npts(path, block)
    xyz = np.zeros((3, N))
    ib = np.zeros(N, dtype=np.intc)
    check(_lib().mgd_plot3d_read_grid(_b(path), block, dptr(xyz), _ip(ib)))
    return xyz, ib


def read_solution(path: str, nv: int, block: int = 0):
    """(state (nv, N), auxiliary data [timestep, -, -, time])."""
    _, N = _npts(path, block)
    q = np.zeros((nv, N))
    aux = np.zeros(4)
    check(_lib().mgd_plot3d_read_solution(_b(path), block, nv, dptr(q), dptr(aux)))
    return q, aux


def read_function(path: str, block: int = 0):
    info, N = _npts(path, block)
    f = np.zeros((info["n_scalars"], N))
    check(_lib().mgd_plot3d_read_function(_b(path), block, info["n_scalars"], dptr(f)))
    return f


def write_grid(path: str, n: Sequence[int], coords: np.ndarray, iblank: Optional[np.ndarray] = None):
    nn = np.array((list(n) + [1, 1, 1])[:3], dtype=np.intc)
    c = np.ascontiguousarray(coords, dtype=np.float64)
    ib = None if iblank is None else np.ascontiguousarray(iblank, dtype=np.intc)
    check(_lib().mgd_plot3d_write_grid(_b(path), c.shape[0], _ip(nn), dptr(c), _ip(ib)))


def write_solution(path: str, n: Sequence[int], state: np.ndarray, timestep: int = 0,
                   time: float = 0.0):
    nn = np.array((list(n) + [1, 1, 1])[:3], dtype=np.intc)
    q = np.ascontiguousarray(state, dtype=np.float64)
    aux = np.array([float(timestep), 0.0, 0.0, time])
    check(_lib().mgd_plot3d_write_solution(_b(path), _ip(nn), q.shape[0], dptr(q), dptr(aux)))


def write_function(path: str, n: Sequence[int], data: np.ndarray):
    nn = np.array((list(n) + [1, 1, 1])[:3], dtype=np.intc)
    f = np.ascontiguousarray(np.atleast_2d(data), dtype=np.float64)
    check(_lib().mgd_plot3d_write_function(_b(path), _ip(nn), f.shape[0], dptr(f)))


def write_control(path: str, forward_order: np.ndarray):
    """Gradient / control forcing (stages x patch points, forward stage order) to a .dat."""
    a = np.ascontiguousarray(forward_order, dtype=np.float64)
    check(_lib().mgd_control_write(_b(path), a.shape[0], a.shape[1], dptr(a)))


def read_control(path: str, npoints: int) -> np.ndarray:
    nst = os.path.getsize(path) // (8 * npoints)
    a = np.zeros((nst, npoints))
    check(_lib().mgd_control_read(_b(path), nst, npoints, dptr(a)))
    return a


def zaxpy(z: str, a: float, x: str, y: Optional[str] = None):
    check(_lib().mgd_control_zaxpy(_b(z), a, _b(x), None if y is None else _b(y)))


def read_bc(path: str, params: Optional[Dict[str, Sequence[float]]] = None) -> List[Patch]:
    """bc.dat rows (Name Type Grid normDir iMin iMax jMin jMax kMin kMax) of grid 1 as
    Patch records; `params` gives the patch options by name (the patches/<name>/...
    entries of magudi.inp).  Unknown types raise."""
    out = []
    for line in open(path):
        s = line.split("#", 1)[0].split()
        if not s:
            continue
        if len(s) != 10:
            raise ValueError(f"bc.dat row needs 10 fields: {line.rstrip()}")
        name, typ, grid = s[0], s[1], int(s[2])
        if grid != 1:
            continue
        if typ not in PATCH_TYPE_NAMES:
            raise ValueError(f"unsupported patch type {typ}")
        p = tuple((params or {}).get(name, (1.0, 0.0, 0.0, 0.0)))
        out.append(Patch(name, PATCH_TYPE_NAMES[typ], int(s[3]), tuple(int(v) for v in s[4:10]),
                         (tuple(p) + (0.0,) * 4)[:4]))
    return out
