"""DEN and NRRD files for volumes, projections and P-matrices (io/den, io/nrrd of the reference),
through the library's C ABI (ctlproj_den_*, ctlproj_nrrd_*).

Mapping (io/den/denfileio.h, io/basetypeio.tpp):
  volume (X, Y, Z)                      DEN cols=X rows=Y count=Z;   NRRD sizes X Y Z
  projections (views, modules, R, C)    DEN cols=C rows=R count=M*V; NRRD sizes C R M V
  geometry [V][M][3][4]                 DEN cols=4 rows=3 count=V*M (FP64)
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib
from ._lib import check, lib
from .projector import ProjectionData, VoxelVolume

_DTYPES = {1: np.uint8, 2: np.uint16, 4: np.float32, 8: np.float64}
_ELEM = {np.dtype(np.uint8): 1, np.dtype(np.uint16): 2, np.dtype(np.float32): 4, np.dtype(np.float64): 8}
VOLUME_LABELS = '"num vox x" "num vox y" "num vox z"'
PROJECTION_LABELS = '"num channel" "num row" "num det module" "num proj"'


def _path(p) -> bytes:
    return os.fsencode(p)


# ---- DEN ----------------------------------------------------------------------------------------

def den_info(path) -> tuple[int, int, int, type]:
    """(rows, cols, count, numpy dtype) of a DEN file."""
    r, c, n, eb = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint32()
    check(lib().ctlproj_den_info(_path(path), C.byref(r), C.byref(c), C.byref(n), C.byref(eb)))
    return r.value, c.value, n.value, _DTYPES[eb.value]


def load_den(path) -> np.ndarray:
    """The data as an array of shape (count, rows, cols)."""
    r, c, n, dt = den_info(path)
    out = np.empty((n, r, c), dtype=dt)
    check(lib().ctlproj_den_read(_path(path), out.ctypes.data_as(C.c_void_p), out.nbytes))
    return out


def save_den(path, data: np.ndarray) -> None:
    """data of shape (count, rows, cols) or (rows, cols); uint8, uint16, float32 or float64."""
    a = np.ascontiguousarray(data)
    if a.dtype not in _ELEM:
        raise TypeError(f"DEN stores uint8, uint16, float32 or float64, not {a.dtype}")
    if a.ndim == 2:
        a = a[None]
    if a.ndim != 3:
        raise ValueError("DEN data must be 2-D or 3-D")
    n, r, c = a.shape
    check(lib().ctlproj_den_write(_path(path), a.ctypes.data_as(C.c_void_p), _ELEM[a.dtype], r, c, n))


def save_volume_den(path, volume: VoxelVolume) -> None:
    save_den(path, volume.data)  # (Z, Y, X): count=Z rows=Y cols=X


def load_volume_den(path, voxel_size=(1.0, 1.0, 1.0), offset=(0.0, 0.0, 0.0)) -> VoxelVolume:
    return VoxelVolume(load_den(path).astype(np.float32, copy=False), voxel_size, offset)


def save_projections_den(path, proj: ProjectionData) -> None:
    v, m, r, c = proj.data.shape
    save_den(path, proj.data.reshape(v * m, r, c))


def load_projections_den(path, modules: int = 1) -> ProjectionData:
    a = load_den(path).astype(np.float32, copy=False)
    n, r, c = a.shape
    if n % modules:
        raise ValueError(f"{n} matrices do not split into views of {modules} modules")
    return ProjectionData(a.reshape(n // modules, modules, r, c))


def save_geometry_den(path, pmats: np.ndarray) -> None:
    """P-matrices [V][M][3][4] (FP64), as BaseTypeIO::write(FullGeometry) stores them."""
    p = np.asarray(pmats, np.float64)
    save_den(path, p.reshape(-1, 3, 4))


def load_geometry_den(path, modules: int = 1) -> np.ndarray:
    """BaseTypeIO::readFullGeometry(fileName, nbModules): [V][M][3][4] FP64."""
    a = load_den(path)
    if a.shape[1:] != (3, 4):
        raise ValueError(f"a geometry file holds 3x4 matrices, not {a.shape[1:]}")
    if a.shape[0] % modules:
        raise ValueError(f"{a.shape[0]} matrices do not split into views of {modules} modules")
    return a.astype(np.float64).reshape(-1, modules, 3, 4)


# ---- NRRD ---------------------------------------------------------------------------------------

def nrrd_info(path) -> tuple[_lib.NrrdInfo, dict]:
    """Header fields and the key:=value pairs."""
    info = _lib.NrrdInfo()
    buf = C.create_string_buffer(1 << 16)
    check(lib().ctlproj_nrrd_read_header(_path(path), C.byref(info), buf, len(buf)))
    kv = {}
    for line in buf.value.decode().splitlines():
        k, _, v = line.partition(":=")
        kv[k] = v
    return info, kv


def load_nrrd(path) -> np.ndarray:
    """The data as an array with the NRRD sizes reversed (first size fastest)."""
    info, _ = nrrd_info(path)
    shape = tuple(int(info.sizes[k]) for k in range(info.ndim))[::-1]
    out = np.empty(shape, dtype=_DTYPES[info.elem_bytes])
    check(lib().ctlproj_nrrd_read(_path(path), out.ctypes.data_as(C.c_void_p), out.nbytes))
    return out


def save_nrrd(path, data: np.ndarray, labels: str | None = None, spacing=None, origin=None,
              keyvals: dict | None = None) -> None:
    a = np.ascontiguousarray(data)
    if a.dtype not in _ELEM:
        raise TypeError(f"NRRD output supports uint8, uint16, float32 or float64, not {a.dtype}")
    if not 1 <= a.ndim <= 4:
        raise ValueError("NRRD data must have 1 to 4 dimensions")
    sizes = (C.c_uint64 * 4)(*a.shape[::-1])
    sp = (C.c_double * 3)(*spacing) if spacing is not None else None
    org = (C.c_double * 3)(*origin) if origin is not None else None
    kv = "".join(f"{k}:={v}\n" for k, v in (keyvals or {}).items()).encode() or None
    check(lib().ctlproj_nrrd_write(_path(path), a.ctypes.data_as(C.c_void_p), _ELEM[a.dtype], a.ndim, sizes,
                                   labels.encode() if labels else None, sp, org, kv))


def save_volume_nrrd(path, volume: VoxelVolume) -> None:
    save_nrrd(path, volume.data, VOLUME_LABELS, volume.voxel_size, volume.offset)


def load_volume_nrrd(path) -> VoxelVolume:
    info, _ = nrrd_info(path)
    if info.ndim != 3:
        raise ValueError(f"a volume has 3 dimensions, the file {info.ndim}")
    vs = tuple(info.spacing) if info.has_spacing else (1.0, 1.0, 1.0)
    off = tuple(info.origin) if info.has_origin else (0.0, 0.0, 0.0)
    return VoxelVolume(load_nrrd(path).astype(np.float32, copy=False), vs, off)


def save_projections_nrrd(path, proj: ProjectionData) -> None:
    save_nrrd(path, proj.data, PROJECTION_LABELS)  # (V, M, R, C): sizes C R M V


def load_projections_nrrd(path) -> ProjectionData:
    a = load_nrrd(path).astype(np.float32, copy=False)
    if a.ndim == 2:
        a = a[None, None]
    elif a.ndim == 3:
        a = a[:, None]
    return ProjectionData(a)
