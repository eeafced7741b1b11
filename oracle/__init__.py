"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE, NOT PRODUCT CODE: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module, as the checker or as the timed CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"


class OracleSettings(C.Structure):
    _fields_ = [
        ("rays_per_pixel", C.c_uint32 * 2),
        ("ray_sampling", C.c_float),
        ("volume_upsampling", C.c_uint32),
        ("interpolate", C.c_int32),
    ]


_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise ImportError(f"{LIB} is missing: run `make -C oracle`")
        L = C.CDLL(str(LIB))
        d = C.POINTER(C.c_double)
        f = C.POINTER(C.c_float)
        u = C.POINTER(C.c_uint32)
        L.oracle_qr.argtypes = [d, d, d]
        L.oracle_qr.restype = C.c_int
        L.oracle_decompose_m.argtypes = [d, d]
        L.oracle_source_position.argtypes = [d, d]
        L.oracle_direction64.argtypes = [d, C.c_double, C.c_double, d]
        L.oracle_project.argtypes = [f, u, f, f, d, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                     C.POINTER(OracleSettings), C.c_int, C.c_int, u, C.c_uint32, f, d,
                                     C.POINTER(C.c_uint64)]
        L.oracle_project.restype = C.c_int
        L.oracle_sample_linear32.argtypes = [f, u, C.c_float, C.c_float, C.c_float]
        L.oracle_sample_linear32.restype = C.c_float
        L.oracle_sample_linear64.argtypes = [f, u, C.c_double, C.c_double, C.c_double]
        L.oracle_sample_linear64.restype = C.c_double
        L.oracle_clip32.argtypes = [f, f, f, f, C.c_int, f]
        L.oracle_cpu_interpolated_read.argtypes = [f, u, C.c_double, C.c_double, C.c_double]
        L.oracle_cpu_interpolated_read.restype = C.c_float
        L.oracle_cpu_nearest_read.argtypes = [f, u, C.c_double, C.c_double, C.c_double]
        L.oracle_cpu_nearest_read.restype = C.c_float
        L.oracle_slice.argtypes = [f, u, f, f, d, C.c_double, C.c_uint32, C.c_uint32, C.c_float,
                                   C.c_int, f]
        L.oracle_slice.restype = C.c_int
        L.oracle_resample.argtypes = [f, u, f, f, C.c_uint32, f, C.c_uint32, f, C.c_uint32, C.c_int, f]
        L.oracle_resample.restype = C.c_int
        L.oracle_sample_points.argtypes = [f, u, f, f, C.c_size_t, C.c_int, f]
        L.oracle_sample_points.restype = C.c_int
        L.oracle_plane_integrals.argtypes = [f, u, f, f, d, C.c_uint32, f, C.c_uint32, C.c_uint32,
                                             C.c_uint32, C.c_float, C.c_int, f]
        L.oracle_plane_integrals.restype = C.c_int
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def qr(A):
    A = np.ascontiguousarray(A, np.float64).reshape(9)
    Q = np.zeros(9)
    R = np.zeros(9)
    lib().oracle_qr(_d(A), _d(Q), _d(R))
    return Q.reshape(3, 3), R.reshape(3, 3)


def decompose_m(P):
    P = np.ascontiguousarray(P, np.float64).reshape(12)
    out = np.zeros(16)
    lib().oracle_decompose_m(_d(P), _d(out))
    return out


def source_position(P):
    P = np.ascontiguousarray(P, np.float64).reshape(12)
    c = np.zeros(3)
    lib().oracle_source_position(_d(P), _d(c))
    return c


def direction64(qr16, x, y):
    q = np.ascontiguousarray(qr16, np.float64)
    d = np.zeros(3)
    lib().oracle_direction64(_d(q), x, y, _d(d))
    return d


def sample_linear32(vol, px, py, pz):
    v = np.ascontiguousarray(vol, np.float32)
    z, y, x = v.shape
    dims = (C.c_uint32 * 3)(x, y, z)
    return lib().oracle_sample_linear32(_f(v), dims, px, py, pz)


def sample_linear64(vol, px, py, pz):
    v = np.ascontiguousarray(vol, np.float32)
    z, y, x = v.shape
    dims = (C.c_uint32 * 3)(x, y, z)
    return lib().oracle_sample_linear64(_f(v), dims, px, py, pz)


def cpu_read(vol, px, py, pz, interpolate=True):
    """RayCasterProjectorCPU's interpolatedRead / nonInterpolatedRead at an FP64 position."""
    v = np.ascontiguousarray(vol, np.float32)
    z, y, x = v.shape
    dims = (C.c_uint32 * 3)(x, y, z)
    fn = lib().oracle_cpu_interpolated_read if interpolate else lib().oracle_cpu_nearest_read
    return fn(_f(v), dims, px, py, pz)


def _vol_args(vol):
    v = np.ascontiguousarray(vol, np.float32)
    z, y, x = v.shape
    return v, (C.c_uint32 * 3)(x, y, z)


def slice_plane(vol, voxel_mm, offset_mm, normal, distance, width, height, resolution, fp64=False):
    """VolumeSlicer::slice; vol is a (Z, Y, X) array, returns (height, width)."""
    v, dims = _vol_args(vol)
    out = np.zeros((height, width), np.float32)
    vs, off = np.asarray(voxel_mm, np.float32), np.asarray(offset_mm, np.float32)
    n = np.asarray(normal, np.float64)
    st = lib().oracle_slice(_f(v), dims, _f(vs), _f(off), _d(n), float(distance), width, height,
                            float(resolution), int(fp64), _f(out))
    assert st == 0
    return out


def resample(vol, ranges, p1, p2, p3, fp64=False):
    """VolumeResampler::resample; returns (len(p3), len(p2), len(p1))."""
    v, dims = _vol_args(vol)
    r = np.asarray(ranges, np.float32).reshape(6)
    a = [np.ascontiguousarray(p, np.float32) for p in (p1, p2, p3)]
    out = np.zeros((len(a[2]), len(a[1]), len(a[0])), np.float32)
    st = lib().oracle_resample(_f(v), dims, _f(r), _f(a[0]), len(a[0]), _f(a[1]), len(a[1]),
                               _f(a[2]), len(a[2]), int(fp64), _f(out))
    assert st == 0
    return out


def sample_points(vol, ranges, xyz, fp64=False):
    """VolumeResampler::sample; xyz is (n, 3), returns (n,)."""
    v, dims = _vol_args(vol)
    r = np.asarray(ranges, np.float32).reshape(6)
    pts = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
    out = np.zeros(len(pts), np.float32)
    st = lib().oracle_sample_points(_f(v), dims, _f(r), _f(pts), len(pts), int(fp64), _f(out))
    assert st == 0
    return out


def plane_integrals(vol, voxel_mm, offset_mm, normals, distances, width, height, resolution,
                    fp64=False):
    """RadonTransform3D plane integrals; normals (n_orient, 3), returns (n_dist, n_orient)."""
    v, dims = _vol_args(vol)
    vs, off = np.asarray(voxel_mm, np.float32), np.asarray(offset_mm, np.float32)
    n = np.ascontiguousarray(normals, np.float64).reshape(-1, 3)
    dist = np.ascontiguousarray(distances, np.float32).reshape(-1)
    out = np.zeros((len(dist), len(n)), np.float32)
    st = lib().oracle_plane_integrals(_f(v), dims, _f(vs), _f(off), _d(n), len(n), _f(dist), len(dist),
                                      width, height, float(resolution), int(fp64), _f(out))
    assert st == 0
    return out


def clip32(source, direction, vol_size, vol_corner, interpolate=True):
    a = [np.ascontiguousarray(v, np.float32) for v in (source, direction, vol_size, vol_corner)]
    mm = np.zeros(2, np.float32)
    lib().oracle_clip32(*[_f(v) for v in a], int(interpolate), _f(mm))
    return mm


def project(volume, voxel_mm, offset_mm, pmats, channels, rows, ray_sampling=0.3,
            rays_per_pixel=(1, 1), upsampling=1, interpolate=True, fp64=False, threads=None,
            views=None, want64=False, cpu_projector=False):
    """Runs ref32 (fp64=False) or ref64 (fp64=True), or with cpu_projector=True the restatement
    of RayCasterProjectorCPU (refcpu; no upsampling). volume: (Z, Y, X) float32;
    pmats: (V, M, 3, 4). Returns (out[V', M, R, C] float32, n_samples[, out64])."""
    vol = np.ascontiguousarray(volume, np.float32)
    z, y, x = vol.shape
    pm = np.ascontiguousarray(pmats, np.float64)
    if pm.ndim == 3:
        pm = pm[:, None]
    V, M = pm.shape[:2]
    s = OracleSettings()
    s.rays_per_pixel[:] = rays_per_pixel
    s.ray_sampling = ray_sampling
    s.volume_upsampling = upsampling
    s.interpolate = 1 if interpolate else 0
    if views is not None:
        vs = np.ascontiguousarray(views, np.uint32)
        vptr, nv = vs.ctypes.data_as(C.POINTER(C.c_uint32)), len(vs)
    else:
        vptr, nv = None, V
    out = np.zeros((nv, M, rows, channels), np.float32)
    out64 = np.zeros(out.shape, np.float64) if (want64 and fp64 and not cpu_projector) else None
    n = C.c_uint64()
    threads = threads or os.cpu_count() or 1
    rc = lib().oracle_project(_f(vol), (C.c_uint32 * 3)(x, y, z), (C.c_float * 3)(*voxel_mm),
                              (C.c_float * 3)(*offset_mm), _d(pm), V, M, channels, rows,
                              C.byref(s), 2 if cpu_projector else int(bool(fp64)), int(threads), vptr, nv, _f(out),
                              _d(out64) if out64 is not None else None, C.byref(n))
    if rc != 0:
        raise ValueError(f"oracle_project failed ({rc})")
    if out64 is not None:
        return out, n.value, out64
    return out, n.value


def spectral_extinction(material_projs, mu, intensity, response=None, total_intensity=None):
    """SpectralEffectsExtension::projectLinear (spectraleffectsextension.cpp:232-300, 330-352) in
    FP64: material_projs (n_mat, V, ...) projections of the densities, mu (n_bins, n_mat),
    intensity (n_bins, V); per pixel ln(I0_v / sum_b r_b i_bv exp(-sum_m mu_bm p_m)). Bins whose
    intensities sum to zero over the views are skipped (:337-341); I0 defaults to the sum over
    bins."""
    p = np.asarray(material_projs, np.float64)
    mu = np.asarray(mu, np.float64)
    inten = np.asarray(intensity, np.float64)
    n_bins, V = inten.shape
    r = np.ones(n_bins) if response is None else np.asarray(response, np.float64)
    i0 = inten.sum(axis=0) if total_intensity is None else np.asarray(total_intensity, np.float64)
    shape = (V,) + (1,) * (p.ndim - 2)
    s = np.zeros(p.shape[1:])
    for b in range(n_bins):
        if abs(inten[b].sum()) <= 1e-12:
            continue
        e = np.tensordot(mu[b], p, axes=(0, 0))
        s += inten[b].reshape(shape) * np.exp(-e) * r[b]
    return np.log(i0.reshape(shape) / s)


def rel_l2(ref, x) -> float:
    """RelativeL2Norm (processing/errormetrics.cpp:69-85): ||ref - x|| / ||ref||, ref first."""
    r = np.asarray(ref, np.float64).ravel()
    d = r - np.asarray(x, np.float64).ravel()
    return float(np.sqrt(np.sum(d * d)) / np.sqrt(np.sum(r * r)))


def max_abs_rel(ref, x, floor_frac=1e-3) -> float:
    """max over pixels of |x - ref| / max(|ref|, floor_frac * max|ref|)."""
    r = np.asarray(ref, np.float64).ravel()
    d = np.abs(np.asarray(x, np.float64).ravel() - r)
    den = np.maximum(np.abs(r), floor_frac * np.max(np.abs(r)))
    return float(np.max(d / den))
