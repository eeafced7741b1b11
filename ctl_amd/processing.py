"""VolumeSlicer, VolumeResampler and RadonTransform3D on resident volumes.

Mirrors CTL's OpenCL processing classes (modules/src/processing/volumeslicer.h,
volumeresampler.h, radontransform3d.h): the same constructors, method names, argument meaning
and default slice sizes. Instead of uploading the volume into an OpenCL image, each instance
keeps a ResidentVolume in HBM (FP32, FP16 or BF16 storage) and every call runs the HIP kernels
behind the C ABI (ctlproj_slice, ctlproj_resample, ctlproj_sample,
ctlproj_radon3d_plane_integrals). Angles are in radians, lengths in mm.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib
from .projector import Context, ResidentVolume, VoxelVolume


def _resident(volume, storage: str, ctx: Context | None) -> ResidentVolume:
    if isinstance(volume, ResidentVolume):
        return volume
    if not isinstance(volume, VoxelVolume):
        volume = VoxelVolume(volume)
    return ResidentVolume(volume, storage, ctx)


def _unit_normal(azimuth: float, polar: float) -> np.ndarray:
    """(sin(polar) cos(azimuth), sin(polar) sin(azimuth), cos(polar)) in double precision."""
    a, p = float(azimuth), float(polar)
    return np.array([np.sin(p) * np.cos(a), np.sin(p) * np.sin(a), np.cos(p)])


class SamplingRange:
    """Range<float> (processing/coordinates.h:40-88): start, end, linspace, spacing, center."""

    def __init__(self, start: float, end: float):
        self._d = np.array([start, end], np.float32)

    def start(self) -> float:
        return float(self._d[0])

    def end(self) -> float:
        return float(self._d[1])

    def width(self) -> float:
        return float(self._d[1] - self._d[0])

    def spacing(self, n: int) -> float:
        return float((self._d[1] - self._d[0]) / np.float32(n - 1)) if n > 1 else 0.0

    def center(self) -> float:
        return float((self._d[0] + self._d[1]) / np.float32(2))

    def linspace(self, n: int) -> np.ndarray:
        inc = (self._d[1] - self._d[0]) / np.float32(n - 1) if n > 1 else np.float32(0)
        return (self._d[0] + np.arange(n, dtype=np.float32) * np.float32(inc)).astype(np.float32)


class VolumeSlicer:
    """VolumeSlicer (volumeslicer.h:15-60, volumeslicer.cpp:20-229)."""

    def __init__(self, volume, storage: str = "f32", ctx: Context | None = None):
        self._vol = _resident(volume, storage, ctx)
        side = C.c_uint32()
        check(lib().ctlproj_slicer_default_side(_lib.u32arr(*self._vol.dims), C.byref(side)))
        self._dim = (side.value, side.value)
        self._reso = float(min(self._vol.voxel_size))

    def setSliceDimensions(self, dimensions) -> None:
        self._dim = (int(dimensions[0]), int(dimensions[1]))

    def setSliceResolution(self, pixelResolution: float) -> None:
        self._reso = float(pixelResolution)

    def sliceDimensions(self) -> tuple[int, int]:
        return self._dim

    def sliceResolution(self) -> float:
        return self._reso

    def volDim(self):
        return self._vol.dims

    def volOffset(self):
        return self._vol.offset

    def volVoxSize(self):
        return self._vol.voxel_size

    def slice(self, *args) -> np.ndarray:
        """slice(planeUnitNormal, distance) or slice(azimuth, polar, distance): a (height, width)
        array (Chunk2D, x fastest)."""
        if len(args) == 2:
            normal, dist = np.asarray(args[0], np.float64).reshape(3), float(args[1])
        elif len(args) == 3:
            normal, dist = _unit_normal(args[0], args[1]), float(args[2])
        else:
            raise TypeError("slice(normal, distance) or slice(azimuth, polar, distance)")
        return self.slices(normal[None], [dist])[0]

    def slices(self, normals, distances) -> np.ndarray:
        """Many planes in one launch: (n, height, width)."""
        n = np.ascontiguousarray(normals, np.float64).reshape(-1, 3)
        d = np.ascontiguousarray(distances, np.float64).reshape(-1)
        if len(d) != len(n):
            raise ValueError("one distance per normal")
        w, h = self._dim
        out = np.zeros((len(n), h, w), np.float32)
        check(lib().ctlproj_slice(self._vol.h, _lib.dptr(n), _lib.dptr(d), len(n), w, h, self._reso,
                                  _lib.fptr(out), out.size))
        return out


class VolumeResampler:
    """VolumeResampler (volumeresampler.h:11-56, volumeresampler.cpp:14-297)."""

    def __init__(self, volume, rangeDim1=None, rangeDim2=None, rangeDim3=None, storage: str = "f32",
                 ctx: Context | None = None):
        self._vol = _resident(volume, storage, ctx)
        self._host = volume if isinstance(volume, VoxelVolume) else getattr(self._vol, "volume", None)
        if rangeDim1 is None:
            r = (C.c_float * 6)()
            check(lib().ctlproj_resampler_default_ranges(self._vol.h, r))
            self._ranges = np.array(r[:], np.float32)
        else:
            self.setSamplingRanges(rangeDim1, rangeDim2, rangeDim3)

    @staticmethod
    def _pair(r) -> tuple[float, float]:
        if isinstance(r, SamplingRange):
            return r.start(), r.end()
        return float(r[0]), float(r[1])

    def setSamplingRanges(self, rangeDim1, rangeDim2, rangeDim3) -> None:
        self._ranges = np.array([*self._pair(rangeDim1), *self._pair(rangeDim2), *self._pair(rangeDim3)], np.float32)

    def rangeDim1(self) -> SamplingRange:
        return SamplingRange(*self._ranges[0:2])

    def rangeDim2(self) -> SamplingRange:
        return SamplingRange(*self._ranges[2:4])

    def rangeDim3(self) -> SamplingRange:
        return SamplingRange(*self._ranges[4:6])

    def volDim(self):
        return self._vol.dims

    def volOffset(self):
        return (self.rangeDim1().center(), self.rangeDim2().center(), self.rangeDim3().center())

    def volVoxSize(self):
        d = self._vol.dims
        return (self.rangeDim1().spacing(d[0]), self.rangeDim2().spacing(d[1]), self.rangeDim3().spacing(d[2]))

    def volume(self) -> VoxelVolume:
        """The managed volume with the voxel size and offset the sampling ranges define."""
        if self._host is None:
            raise RuntimeError("volume() needs the host copy the resampler was created from")
        return VoxelVolume(self._host.data, self.volVoxSize(), self.volOffset())

    def resample(self, samplingPtsDim1, samplingPtsDim2, samplingPtsDim3) -> VoxelVolume:
        p = [np.ascontiguousarray(v, np.float32).reshape(-1) for v in (samplingPtsDim1, samplingPtsDim2, samplingPtsDim3)]
        out = np.zeros((len(p[2]), len(p[1]), len(p[0])), np.float32)
        check(lib().ctlproj_resample(self._vol.h, _lib.fptr(self._ranges), _lib.fptr(p[0]), len(p[0]),
                                     _lib.fptr(p[1]), len(p[1]), _lib.fptr(p[2]), len(p[2]), _lib.fptr(out),
                                     out.size))
        return VoxelVolume(out)

    def sample(self, samplingPts) -> np.ndarray:
        """Values at a list of points ((n, 3) array of Generic3DCoord)."""
        pts = np.ascontiguousarray(samplingPts, np.float32).reshape(-1, 3)
        out = np.zeros(len(pts), np.float32)
        check(lib().ctlproj_sample(self._vol.h, _lib.fptr(self._ranges), _lib.fptr(pts), len(pts), _lib.fptr(out)))
        return out

    def sample_device(self, device_xyz: int, n: int, device_out: int) -> None:
        """The cl::Buffer overload (volumeresampler.cpp:215-255): points and results in device memory."""
        check(lib().ctlproj_sample_device(self._vol.h, _lib.fptr(self._ranges), C.c_void_p(device_xyz), n,
                                          C.c_void_p(device_out)))


class RadonTransform3D:
    """RadonTransform3D (radontransform3d.h:92-173, radontransform3d.cpp:18-301)."""

    def __init__(self, volume, storage: str = "f32", ctx: Context | None = None):
        self._vol = _resident(volume, storage, ctx)
        side, reso = C.c_uint32(), C.c_float()
        check(lib().ctlproj_radon3d_default_slice(_lib.u32arr(*self._vol.dims), _lib.f32arr(*self._vol.voxel_size),
                                                  C.byref(side), C.byref(reso)))
        self._side, self._reso = side.value, reso.value

    def setSliceResolution(self, pixelResolution: float) -> None:
        side = C.c_uint32()
        check(lib().ctlproj_radon3d_side_for_resolution(self._side, self._reso, float(pixelResolution),
                                                        C.byref(side)))
        self._side, self._reso = side.value, float(np.float32(pixelResolution))

    def sliceDimensions(self) -> tuple[int, int]:
        return (self._side, self._side)

    def sliceResolution(self) -> float:
        return self._reso

    def volDim(self):
        return self._vol.dims

    def volOffset(self):
        return self._vol.offset

    def volVoxSize(self):
        return self._vol.voxel_size

    def planeIntegrals(self, normals, distances) -> np.ndarray:
        """All (distance, orientation) pairs: (n_dist, n_orient) array."""
        n = np.ascontiguousarray(normals, np.float64).reshape(-1, 3)
        d = np.ascontiguousarray(distances, np.float32).reshape(-1)
        out = np.zeros((len(d), len(n)), np.float32)
        check(lib().ctlproj_radon3d_plane_integrals(self._vol.h, _lib.dptr(n), len(n), _lib.fptr(d), len(d),
                                                    self._side, self._reso, _lib.fptr(out), out.size))
        return out

    def planeIntegral(self, *args) -> float:
        """planeIntegral(planeUnitNormal, distance) or planeIntegral(azimuth, polar, distance)."""
        if len(args) == 2:
            normal, dist = np.asarray(args[0], np.float64).reshape(3), args[1]
        elif len(args) == 3:
            normal, dist = _unit_normal(args[0], args[1]), args[2]
        else:
            raise TypeError("planeIntegral(normal, distance) or planeIntegral(azimuth, polar, distance)")
        return float(self.planeIntegrals(normal[None], [dist])[0, 0])

    def sampleTransform(self, *args) -> VoxelVolume:
        """sampleTransform(azimuths, polars, distances) or sampleTransform(aziRange, nAzi, polRange,
        nPol, distRange, nDist); returns a VoxelVolume of (nAzi, nPol, nDist) samples."""
        if len(args) == 3:
            azi, pol, dist = (np.ascontiguousarray(a, np.float32).reshape(-1) for a in args)
            ranges = None
        elif len(args) == 6:
            ranges = [a if isinstance(a, SamplingRange) else SamplingRange(*a) for a in args[0::2]]
            counts = [int(c) for c in args[1::2]]
            azi, pol, dist = (r.linspace(c) for r, c in zip(ranges, counts))
        else:
            raise TypeError("sampleTransform(azi, pol, dist) or sampleTransform(range, n, range, n, range, n)")
        out = np.zeros((len(dist), len(pol), len(azi)), np.float32)
        check(lib().ctlproj_radon3d_sample_transform(self._vol.h, _lib.fptr(azi), len(azi), _lib.fptr(pol), len(pol),
                                                     _lib.fptr(dist), len(dist), self._side, self._reso,
                                                     _lib.fptr(out), out.size))
        if ranges is None:
            return VoxelVolume(out)
        return VoxelVolume(out, tuple(r.spacing(c) for r, c in zip(ranges, counts)),
                           tuple(r.center() for r in ranges))
