"""Python mirror of CTL's projector interface over the C ABI.

Names follow the reference: ``RayCasterProjector`` with ``settings()``, ``configure(setup)`` and
``project(volume)`` (modules/src/projectors/raycasterprojector.h:20-58), ``VoxelVolume`` /
``VolumeData`` (img/voxelvolume.h), ``ProjectionData`` (img/projectiondata.h),
``AcquisitionSetup`` with systems and trajectory protocols (acquisition/), and
``GeometryEncoder.encodeFullGeometry`` (acquisition/geometryencoder.cpp:83-98).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import check, lib

DEG = np.pi / 180.0

# ----------------------------------------------------------------------------------------------
# containers


class VoxelVolume:
    """Float volume, x fastest (voxelvolume.tpp:310-332). ``data`` has shape (Z, Y, X)."""

    def __init__(self, data, voxel_size=(1.0, 1.0, 1.0), offset=(0.0, 0.0, 0.0)):
        self.data = np.ascontiguousarray(data, dtype=np.float32)
        if self.data.ndim != 3:
            raise ValueError("volume data must be 3-D (Z, Y, X)")
        self.voxel_size = tuple(float(v) for v in voxel_size)
        self.offset = tuple(float(v) for v in offset)

    @property
    def dims(self) -> tuple[int, int, int]:
        z, y, x = self.data.shape
        return (x, y, z)

    def hasData(self) -> bool:
        return self.data.size > 0

    def smallestVoxelSize(self) -> float:
        return min(self.voxel_size)

    @staticmethod
    def cube(n: int, voxel_size: float, value: float) -> "VoxelVolume":
        return VoxelVolume(np.full((n, n, n), value, np.float32), (voxel_size,) * 3)


VolumeData = VoxelVolume


class ProjectionData:
    """Projections as an array (views, modules, rows, channels) (projectiondata.cpp:134-228)."""

    def __init__(self, data: np.ndarray):
        self.data = data

    def nbViews(self) -> int:
        return self.data.shape[0]

    def view(self, v: int) -> np.ndarray:
        return self.data[v]

    def module(self, v: int, m: int) -> np.ndarray:
        """Module m of view v as a (rows, channels) array: pixel (x, y) is [y, x]."""
        return self.data[v, m]

    def toVector(self) -> np.ndarray:
        return self.data.reshape(-1)

    @classmethod
    def fromVector(cls, values, nbChannels: int, nbRows: int, nbModules: int) -> "ProjectionData":
        """ProjectionData::setDataFromVector (projectiondata.cpp:134-170): the views follow one
        another, each module after module, each row after row, channels fastest."""
        a = np.ascontiguousarray(values, np.float32).reshape(-1)
        per_view = nbChannels * nbRows * nbModules
        if per_view == 0 or a.size % per_view:
            raise ValueError("the vector does not hold a whole number of views")
        return cls(a.reshape(a.size // per_view, nbModules, nbRows, nbChannels))

    @property
    def dimensions(self):
        v, m, r, c = self.data.shape
        return dict(nbChannels=c, nbRows=r, nbModules=m, nbViews=v)


# ----------------------------------------------------------------------------------------------
# system components and protocols


@dataclass
class FlatPanelDetector:
    nbPixels: tuple[int, int]  # (channels, rows)
    pixelDimensions: tuple[float, float]  # mm

    def _c(self) -> _lib.Detector:
        d = _lib.Detector()
        d.pixels[:] = self.nbPixels
        d.pixel_mm[:] = self.pixelDimensions
        d.modules = 1
        d.cylindrical = 0
        return d


@dataclass
class CylindricalDetector:
    nbPixelPerModule: tuple[int, int]
    pixelDimensions: tuple[float, float]
    nbDetectorModules: int
    angulationPerModule: float  # rad
    moduleSpacing: float = 0.0

    def _c(self) -> _lib.Detector:
        d = _lib.Detector()
        d.pixels[:] = self.nbPixelPerModule
        d.pixel_mm[:] = self.pixelDimensions
        d.modules = self.nbDetectorModules
        d.cylindrical = 1
        d.angulation = self.angulationPerModule
        d.module_spacing = self.moduleSpacing
        return d


@dataclass
class CarmGantry:
    cArmSpan: float = 1000.0


@dataclass
class TubularGantry:
    sourceToDetectorDistance: float
    sourceToIsoCenterDistance: float
    rotationAngle: float = 0.0
    pitchPosition: float = 0.0
    tiltAngle: float = 0.0


def _gantry(g) -> _lib.Gantry:
    out = _lib.Gantry()
    if isinstance(g, CarmGantry):
        out.type = _lib.CARM
        out.carm_span = g.cArmSpan
    else:
        out.type = _lib.TUBULAR
        out.sdd = g.sourceToDetectorDistance
        out.sid = g.sourceToIsoCenterDistance
        out.rotation = g.rotationAngle
        out.pitch = g.pitchPosition
        out.tilt = g.tiltAngle
    return out


@dataclass
class CTSystem:
    gantry: object
    detector: object
    focalSpotSize: tuple = (0.0, 0.0)  # mm (width, height) of the X-ray source's focal spot


def _traj(kind: int, *p: float) -> _lib.Trajectory:
    t = _lib.Trajectory()
    t.type = kind
    for i, v in enumerate(p):
        t.p[i] = v
    return t


class ShortScanTrajectory:  # trajectories.cpp:177-205
    def __init__(self, sourceToIsocenter: float, startAngle: float = 0.0, angleSpan: float = -1.0):
        self.c = _traj(_lib.SHORT_SCAN, sourceToIsocenter, startAngle, angleSpan)


class WobbleTrajectory:  # trajectories.cpp:71-96
    def __init__(self, angleSpan, sourceToIsocenter, startAngle=0.0, wobbleAngle=15.0 * DEG,
                 wobbleFreq=1.0):
        self.c = _traj(_lib.WOBBLE, angleSpan, sourceToIsocenter, startAngle, wobbleAngle, wobbleFreq)


class CirclePlusLineTrajectory:  # trajectories.cpp:125-175
    def __init__(self, angleSpan, sourceToIsocenter, lineLength, fractionOfViewsForLine,
                 startAngle=0.0):
        self.c = _traj(_lib.CIRCLE_PLUS_LINE, angleSpan, sourceToIsocenter, lineLength,
                       fractionOfViewsForLine, startAngle)


class HelicalTrajectory:  # trajectories.cpp:33-50
    def __init__(self, angleIncrement, pitchIncrement=0.0, startPitch=0.0, startAngle=0.0):
        self.c = _traj(_lib.HELICAL, angleIncrement, pitchIncrement, startPitch, startAngle)


class AxialScanTrajectory:  # trajectories.cpp:239-256
    def __init__(self, startAngle=0.0):
        self.c = _traj(_lib.AXIAL, startAngle)


class AcquisitionSetup:
    """System + number of views + preparation protocol (acquisitionsetup.h:182-245)."""

    def __init__(self, system: CTSystem, nbViews: int = 0):
        self.system = system
        self._nb_views = 0
        self._time_stamps: list[float] = []
        self._protocol = None
        self._poses = None
        self.setNbViews(nbViews)

    def setNbViews(self, n: int) -> None:
        """Removes excess views, or appends views whose time stamps continue from the last view
        with the last increment (1.0 with fewer than two views), acquisitionsetup.cpp:534-551."""
        n = int(n)
        ts = self._time_stamps
        if n <= len(ts):
            del ts[n:]
        else:
            last = ts[-1] if ts else -1.0
            inc = ts[-1] - ts[-2] if len(ts) > 1 else 1.0
            ts.extend(last + (k + 1) * inc for k in range(n - len(ts)))
        self._nb_views = n

    def nbViews(self) -> int:
        return self._nb_views

    def timeStamp(self, view: int) -> float:
        """AcquisitionSetup::View::timeStamp (acquisitionsetup.cpp:36)."""
        return self._time_stamps[view]

    def setTimeStamp(self, view: int, t: float) -> None:
        self._time_stamps[view] = float(t)

    def applyPreparationProtocol(self, protocol) -> None:
        self._protocol = protocol
        self._poses = None

    def setCarmPoses(self, positions, rotations) -> None:
        """Per-view C-arm source locations (prepare::CarmGantryParam::setLocation)."""
        self._poses = (np.ascontiguousarray(positions, np.float64),
                       np.ascontiguousarray(rotations, np.float64))
        self.setNbViews(len(self._poses[0]))
        self._protocol = None


def _module_pixels(setup: AcquisitionSetup) -> tuple[int, int]:
    """(channels, rows) of one detector module."""
    det = setup.system.detector
    px = det.nbPixels if isinstance(det, FlatPanelDetector) else det.nbPixelPerModule
    return int(px[0]), int(px[1])


class GeometryEncoder:
    @staticmethod
    def encodeFullGeometry(setup: AcquisitionSetup, source_shift=None) -> np.ndarray:
        """FP64 projection matrices, shape (views, modules, 3, 4) (geometryencoder.cpp:83-98).
        source_shift: displacement (mm) of the source in its own frame, as
        prepare::GantryDisplacementParam's source displacement."""
        det = setup.system.detector._c()
        modules = det.modules if det.cylindrical else 1
        V = setup.nbViews()
        out = np.zeros((V, modules, 3, 4), np.float64)
        shift = _lib.dptr(np.asarray(source_shift, np.float64)) if source_shift is not None else None
        if setup._poses is not None:
            pos, rot = setup._poses
            check(lib().ctlproj_encode_carm_poses_shifted(setup.system.gantry.cArmSpan, C.byref(det),
                                                          _lib.dptr(pos), _lib.dptr(rot), V, shift,
                                                          _lib.dptr(out), out.size))
        else:
            if setup._protocol is None:
                raise ValueError("no preparation protocol applied")
            g = _gantry(setup.system.gantry)
            check(lib().ctlproj_encode_trajectory_shifted(C.byref(g), C.byref(det), C.byref(setup._protocol.c),
                                                          V, shift, _lib.dptr(out), out.size))
        return out


class DecodedGeometry:
    """Physical geometry behind P-matrices, as GeometryDecoder builds it for a GenericGantry and
    GenericDetector: per view the source location, pixel size and skew, per view and module the
    module location (centre position, rotation world -> module)."""

    def __init__(self, channels, rows, source_positions, source_rotations, pixel_sizes, skew,
                 module_positions, module_rotations):
        self.channels, self.rows = channels, rows
        self.source_positions = source_positions    # (V, 3)
        self.source_rotations = source_rotations    # (V, 3, 3), columns r1, r2, source->detector
        self.pixel_sizes = pixel_sizes              # (V, 2) width, height
        self.skew = skew                            # (V,)
        self.module_positions = module_positions    # (V, M, 3)
        self.module_rotations = module_rotations    # (V, M, 3, 3)

    def nbViews(self) -> int:
        return self.source_positions.shape[0]

    def encode(self) -> np.ndarray:
        """GeometryEncoder::encodeFullGeometry of the decoded setup: (V, M, 3, 4)."""
        V, M = self.module_positions.shape[:2]
        out = np.zeros((V, M, 3, 4), np.float64)
        f = lambda a: _lib.dptr(np.ascontiguousarray(a, np.float64))
        check(lib().ctlproj_encode_generic(f(self.source_positions), f(self.module_positions),
                                           f(self.module_rotations), f(self.pixel_sizes), f(self.skew),
                                           V, M, self.channels, self.rows, _lib.dptr(out), out.size))
        return out


class GeometryDecoder:
    """GeometryDecoder (geometrydecoder.h:25-74): decodes P-matrices into the physical geometry,
    given the pixels per module and either the pixel size or one physical length."""
    PixelWidth, PixelHeight, SourceDetectorDistance = 1, 2, 3  # PhysicalDimension

    def __init__(self, pixelPerModule, pixelDimensions=None, physicalDimension=None, mm=None):
        self.pixelPerModule = tuple(int(v) for v in pixelPerModule)
        if pixelDimensions is not None:
            self._ref, self._a, self._b = 0, float(pixelDimensions[0]), float(pixelDimensions[1])
        elif physicalDimension is not None and mm is not None:
            self._ref, self._a, self._b = int(physicalDimension), float(mm), 0.0
        else:
            raise ValueError("give the pixel dimensions or a physical dimension with its length")

    def decodeFullGeometry(self, pmats) -> DecodedGeometry:
        P = np.ascontiguousarray(pmats, np.float64)
        if P.ndim == 3:
            P = P[:, None]
        V, M = P.shape[:2]
        C_, R_ = self.pixelPerModule
        src = np.zeros((V, 3)); srot = np.zeros((V, 3, 3)); px = np.zeros((V, 2)); sk = np.zeros(V)
        mpos = np.zeros((V, M, 3)); mrot = np.zeros((V, M, 3, 3))
        d = _lib.dptr
        check(lib().ctlproj_decode_geometry(d(P), V, M, C_, R_, self._ref, self._a, self._b, d(src), d(srot),
                                            d(px), d(sk), d(mpos), d(mrot)))
        return DecodedGeometry(C_, R_, src, srot, px, sk, mpos, mrot)


# ----------------------------------------------------------------------------------------------
# environment and projector


class Context:
    """A device context (replaces OpenCLConfig); one per process and device is enough."""

    def __init__(self, device: int = 0):
        self.h = C.c_void_p()
        check(lib().ctlproj_context_create(device, C.byref(self.h)))

    def close(self) -> None:
        if self.h:
            lib().ctlproj_context_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx: Context | None = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


STORAGE = {"f32": _lib.F32, "f16": _lib.F16, "bf16": _lib.BF16}


class ResidentVolume:
    """A volume kept in HBM (FP32, FP16 or BF16 storage) across projections."""

    def __init__(self, volume: VoxelVolume, storage: str = "f32", ctx: Context | None = None):
        self.ctx = ctx or default_context()
        self.h = C.c_void_p()
        self.volume = volume
        self.storage = storage
        self.dims, self.voxel_size, self.offset = volume.dims, volume.voxel_size, volume.offset
        check(lib().ctlproj_volume_create(
            self.ctx.h, _lib.fptr(volume.data), _lib.u32arr(*volume.dims),
            _lib.f32arr(*volume.voxel_size), _lib.f32arr(*volume.offset), STORAGE[storage],
            C.byref(self.h)))

    @classmethod
    def from_device(cls, ptr: int, dims, voxel_size, offset=(0.0, 0.0, 0.0), storage="f32",
                    ctx: Context | None = None) -> "ResidentVolume":
        self = cls.__new__(cls)
        self.ctx = ctx or default_context()
        self.h = C.c_void_p()
        self.volume = None
        self.storage = storage
        self.dims = tuple(int(v) for v in dims)
        self.voxel_size = tuple(float(v) for v in voxel_size)
        self.offset = tuple(float(v) for v in offset)
        check(lib().ctlproj_volume_create_from_device(
            self.ctx.h, C.c_void_p(ptr), _lib.u32arr(*dims), _lib.f32arr(*voxel_size),
            _lib.f32arr(*offset), STORAGE[storage], C.byref(self.h)))
        return self

    def nbytes(self) -> int:
        n = C.c_size_t()
        check(lib().ctlproj_volume_bytes(self.h, C.byref(n)))
        return n.value

    def close(self) -> None:
        if self.h:
            lib().ctlproj_volume_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class Settings:
    """RayCasterProjector::Settings (raycasterprojector.h:27-37) plus tuning fields."""
    raysPerPixel: list = field(default_factory=lambda: [1, 1])
    raySampling: float = 0.3
    volumeUpSampling: int = 1
    interpolate: bool = True
    tile: tuple | None = None  # (wave_w, wave_h), 0 = automatic
    unroll: int = 0
    blocks_per_cu: int = 0  # cap on resident workgroups per CU (0 = automatic)
    lds_chunk: int = 0  # LDS staging, samples per chunk (0 = automatic, -1 = off)
    lds_texels: int = 0  # LDS box capacity in texels (0 = automatic)

    def _c(self) -> _lib.Settings:
        s = _lib.Settings()
        s.rays_per_pixel[:] = self.raysPerPixel
        s.ray_sampling = self.raySampling
        s.volume_upsampling = self.volumeUpSampling
        s.interpolate = 1 if self.interpolate else 0
        if self.tile:
            s.tile_w, s.tile_h = self.tile
        s.unroll = self.unroll
        s.blocks_per_cu = self.blocks_per_cu
        s.lds_chunk = self.lds_chunk
        s.lds_texels = self.lds_texels
        return s

    @staticmethod
    def optimizedFor(volume: "VoxelVolume", detector) -> "Settings":
        """Settings::optimizedFor (raycasterprojector.cpp:482-521): rays per pixel and volume
        upsampling from the detector's pixel size and the volume's smallest voxel size."""
        c = _lib.Settings()
        vox = (C.c_float * 3)(*volume.voxel_size)
        w, h = detector.pixelDimensions
        check(lib().ctlproj_settings_optimized_for(vox, float(w), float(h), C.byref(c)))
        return Settings(raysPerPixel=list(c.rays_per_pixel), raySampling=c.ray_sampling,
                        volumeUpSampling=c.volume_upsampling, interpolate=bool(c.interpolate))


class RayCasterProjector:
    """Drop-in for CTL::OCL::RayCasterProjector (raycasterprojector.h:20-58)."""

    def __init__(self, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        self._settings = Settings()
        self.h = C.c_void_p()
        s = self._settings._c()
        check(lib().ctlproj_projector_create(self.ctx.h, C.byref(s), C.byref(self.h)))
        self._dims = None
        self.progress = None  # callable(view) -- ProjectorNotifier::projectionFinished

    def settings(self) -> Settings:
        return self._settings

    # SerializationInterface (raycasterprojector.cpp:60-91, abstractprojector.h:114-123,
    # serializationinterface.h:83-88): the same keys and type id as the reference
    TYPE_ID = 1  # CTL_TYPE_ID(1), raycasterprojector.h:22

    def parameter(self) -> dict:
        s = self._settings
        return {"Rays per pixel X": int(s.raysPerPixel[0]), "Rays per pixel Y": int(s.raysPerPixel[1]),
                "Ray sampling step length": float(s.raySampling),
                "Volume upsampling factor": int(s.volumeUpSampling), "Interpolate": bool(s.interpolate)}

    def setParameter(self, parameter: dict) -> None:
        s = self._settings
        s.raysPerPixel = [int(parameter.get("Rays per pixel X", 1)), int(parameter.get("Rays per pixel Y", 1))]
        s.raySampling = float(parameter.get("Ray sampling step length", 0.3))
        s.volumeUpSampling = int(parameter.get("Volume upsampling factor", 1))
        s.interpolate = bool(parameter.get("Interpolate", True))

    def toVariant(self) -> dict:
        return {"type-id": self.TYPE_ID, "parameters": self.parameter(), "#": "RayCasterProjector"}

    def fromVariant(self, variant: dict) -> None:
        if variant.get("type-id") != self.TYPE_ID:
            raise ValueError("variant does not describe a RayCasterProjector")
        self.setParameter(variant.get("parameters", {}))

    def _push_settings(self) -> None:
        s = self._settings._c()
        check(lib().ctlproj_projector_set_settings(self.h, C.byref(s)))

    def configure(self, setup) -> None:
        """configure(AcquisitionSetup) or configure((pmats, channels, rows))."""
        if isinstance(setup, AcquisitionSetup):
            pm = GeometryEncoder.encodeFullGeometry(setup)
            self.configure_matrices(pm, *_module_pixels(setup))
        else:
            pm, c, r = setup
            self.configure_matrices(pm, c, r)

    def configure_matrices(self, pmats: np.ndarray, channels: int, rows: int) -> None:
        pm = np.ascontiguousarray(pmats, np.float64)
        if pm.ndim == 3:
            pm = pm[:, None]
        V, M = pm.shape[:2]
        check(lib().ctlproj_projector_configure(self.h, _lib.dptr(pm), V, M, channels, rows))
        self._dims = (V, M, rows, channels)
        self._pmats = pm

    def _out(self) -> np.ndarray:
        if self._dims is None:
            raise _lib.CtlProjError(_lib.ERR_NOT_CONFIGURED, "project() called before configure()")
        return np.empty(self._dims, np.float32)

    def project(self, volume) -> ProjectionData:
        """project(VolumeData) -> ProjectionData; `volume` may also be a ResidentVolume."""
        self._push_settings()
        out = self._out()
        if isinstance(volume, ResidentVolume):
            cb = _lib.PROGRESS_FN(lambda v, u: self.progress(v)) if self.progress else _lib.PROGRESS_FN()
            check(lib().ctlproj_project(self.h, volume.h, _lib.fptr(out), out.size, cb, None))
        else:
            if not volume.hasData():
                return ProjectionData(np.zeros((0,) + out.shape[1:], np.float32))
            check(lib().ctlproj_project_host_volume(
                self.h, _lib.fptr(volume.data), _lib.u32arr(*volume.dims),
                _lib.f32arr(*volume.voxel_size), _lib.f32arr(*volume.offset), _lib.fptr(out),
                out.size))
        return ProjectionData(out)

    def project_device(self, volume: ResidentVolume, out_ptr: int, count: int) -> None:
        self._push_settings()
        check(lib().ctlproj_project_device(self.h, volume.h, C.c_void_p(out_ptr), count))

    def project_device_noalloc(self, volume: ResidentVolume, out_ptr: int, count: int) -> None:
        """project_device without any device allocation (static per-device scratch; at most
        CTLPROJ_NOALLOC_MAX_VIEW_MODULES (view, module) pairs, upsampling factor 1). Calls that
        overlap on one device race."""
        self._push_settings()
        check(lib().ctlproj_project_device_noalloc(self.h, volume.h, C.c_void_p(out_ptr), count))

    def synchronize(self) -> None:
        check(lib().ctlproj_projector_synchronize(self.h))

    def last_kernel_ms(self) -> float:
        ms = C.c_float()
        check(lib().ctlproj_projector_last_kernel_ms(self.h, C.byref(ms)))
        return ms.value

    def autotune(self, volume: ResidentVolume) -> float:
        """Picks the fastest launch shape for this workload class (see ctlproj_projector_autotune)."""
        self._push_settings()
        ms = C.c_float()
        check(lib().ctlproj_projector_autotune(self.h, volume.h, C.byref(ms)))
        return ms.value

    def count_samples(self, volume: ResidentVolume) -> int:
        self._push_settings()
        n = C.c_uint64()
        check(lib().ctlproj_projector_count_samples(self.h, volume.h, C.byref(n)))
        return n.value

    def count_taken(self, volume: ResidentVolume, device_out: int, out_count: int) -> int:
        """Projects into device_out and returns the number of samples the kernels took
        (ctlproj_projector_count_taken): the reference's count less the samples of chunks that
        read only zero voxels, which are skipped."""
        self._push_settings()
        n = C.c_uint64()
        check(lib().ctlproj_projector_count_taken(self.h, volume.h, C.c_void_p(device_out), out_count,
                                                  C.byref(n)))
        return n.value

    def gather_probe(self, volume: ResidentVolume, mode: str, device_out: int, out_count: int):
        """Gather-ceiling diagnostic (ctlproj_diag_gather_probe): mode "gather" issues this
        projection's texel loads without the interpolation, mode "stream" reads a 16 KiB
        L1-resident window with coalesced 16-byte loads, mode "lds" reads a 16 KiB LDS window
        with conflict-free 16-byte reads, mode "lds_pairs" the same bytes as two adjacent 8-byte
        reads per lane. Returns (milliseconds, bytes loaded)."""
        self._push_settings()
        code = {"gather": 0, "stream": 1, "lds": 2, "lds_pairs": 3}[mode]
        ms, nbytes = C.c_float(), C.c_double()
        check(lib().ctlproj_diag_gather_probe(self.h, volume.h, code, C.c_void_p(device_out),
                                              out_count, C.byref(ms), C.byref(nbytes)))
        return ms.value, nbytes.value

    def isLinear(self) -> bool:  # abstractprojector.h:97
        return True

    def projectComposite(self, volumes) -> ProjectionData:
        """Sum of the projections of several volumes (abstractprojector.h:99-112), accumulated on
        the device (ctlproj_project_composite). Host volumes are made resident first."""
        res = [v if isinstance(v, ResidentVolume) else ResidentVolume(v) for v in volumes]
        self._push_settings()
        out = self._out()
        handles = (C.c_void_p * len(res))(*[r.h.value for r in res])
        check(lib().ctlproj_project_composite(self.h, handles, len(res), _lib.fptr(out), out.size))
        return ProjectionData(out)

    def project_composite_device(self, volumes, out_ptr: int, count: int) -> None:
        self._push_settings()
        handles = (C.c_void_p * len(volumes))(*[r.h.value for r in volumes])
        check(lib().ctlproj_project_composite_device(self.h, handles, len(volumes), C.c_void_p(out_ptr), count))

    def close(self) -> None:
        if self.h:
            lib().ctlproj_projector_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ArealFocalSpotExtension:
    """ArealFocalSpotExtension (arealfocalspotextension.cpp:82-166, 257-292): the focal spot of
    size w x h (CTSystem.focalSpotSize) is sampled on a nx x ny grid over [-1/2, 1/2]^2; each
    sample shifts the source by (x w, y h, 0) in its own frame. All samples are projected in one
    launch (the projector configured with the samples' geometries one after the other) and
    averaged on the device: in the intensity domain (-ln mean exp(-p)), or, with the low-extinction
    approximation, as extinctions."""

    def __init__(self, projector: "RayCasterProjector | None" = None, discretization=(1, 1),
                 lowExtinctionApproximation: bool = False):
        self._nested = projector or RayCasterProjector()
        self._disc = tuple(int(v) for v in discretization)
        self._low = bool(lowExtinctionApproximation)
        self._setup = None

    def use(self, projector: "RayCasterProjector") -> None:
        self._nested = projector

    def setDiscretization(self, discretization) -> None:
        self._disc = tuple(int(v) for v in discretization)

    def enableLowExtinctionApproximation(self, enable: bool = True) -> None:
        self._low = bool(enable)

    def isLinear(self) -> bool:
        return self._low

    def discretizationGrid(self) -> list:
        w, h = self._disc
        xs = [0.0] if w <= 1 else [-0.5 + i / (w - 1) for i in range(w)]
        ys = [0.0] if h <= 1 else [-0.5 + j / (h - 1) for j in range(h)]
        return [(x, y) for x in xs for y in ys]  # x outer, as the reference enumerates

    def configure(self, setup: AcquisitionSetup) -> None:
        self._setup = setup
        fw, fh = setup.system.focalSpotSize
        pm = [GeometryEncoder.encodeFullGeometry(setup, (x * fw, y * fh, 0.0))
              for x, y in self.discretizationGrid()]
        self._k = len(pm)
        self._views = pm[0].shape[0]
        self._nested.configure((np.concatenate(pm, axis=0), *_module_pixels(setup)))

    def project(self, volume) -> ProjectionData:
        import torch  # device buffers (the library computes; torch only allocates)
        vol = volume if isinstance(volume, ResidentVolume) else ResidentVolume(volume)
        V, M, R, C_ = self._nested._dims
        n = (V // self._k) * M * R * C_
        buf = torch.empty(self._k * n, dtype=torch.float32, device="cuda")
        res = torch.empty(n, dtype=torch.float32, device="cuda")
        self._nested.project_device(vol, buf.data_ptr(), buf.numel())
        self._nested.synchronize()
        check(lib().ctlproj_average_projections(self._nested.ctx.h, C.c_void_p(buf.data_ptr()), self._k, n,
                                                0 if self._low else 1, C.c_void_p(res.data_ptr())))
        return ProjectionData(res.cpu().numpy().reshape(V // self._k, M, R, C_))


class Batch:
    """Many projections (own volume, geometry, detector) in one persistent, graph-captured launch.

    Outputs are device arrays owned by the caller (e.g. torch tensors): pass their data_ptr().
    """

    def __init__(self, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        self.h = C.c_void_p()
        check(lib().ctlproj_batch_create(self.ctx.h, C.byref(self.h)))
        self._keep = []

    def add(self, volume: ResidentVolume, pmats: np.ndarray, channels: int, rows: int,
            out_ptr: int, settings: Settings | None = None) -> int:
        pm = np.ascontiguousarray(pmats, np.float64)
        if pm.ndim == 3:
            pm = pm[:, None]
        V, M = pm.shape[:2]
        s = (settings or Settings())._c()
        idx = C.c_uint32()
        check(lib().ctlproj_batch_add(self.h, volume.h, _lib.dptr(pm), V, M, channels, rows,
                                      C.byref(s), C.c_void_p(out_ptr), V * M * rows * channels,
                                      C.byref(idx)))
        self._keep.append((volume, pm))
        return idx.value

    def finalize(self, autotune: bool = True) -> None:
        check(lib().ctlproj_batch_finalize(self.h, 1 if autotune else 0))

    def launch(self) -> None:
        check(lib().ctlproj_batch_launch(self.h))

    def synchronize(self) -> None:
        check(lib().ctlproj_batch_synchronize(self.h))

    def last_ms(self) -> float:
        ms = C.c_float()
        check(lib().ctlproj_batch_last_ms(self.h, C.byref(ms)))
        return ms.value

    def info(self) -> dict:
        v = [C.c_uint32() for _ in range(4)]
        t = C.c_uint64()
        check(lib().ctlproj_batch_info(self.h, *[C.byref(x) for x in v], C.byref(t)))
        return dict(tile_w=v[0].value, tile_h=v[1].value, unroll=v[2].value, grid=v[3].value,
                    tiles=t.value)

    def count_samples(self) -> int:
        n = C.c_uint64()
        check(lib().ctlproj_batch_count_samples(self.h, C.byref(n)))
        return n.value

    def shape_settings(self, raySampling: float) -> Settings:
        """Settings whose tuning fields are the finalized batch's launch shape: a single
        projection with them runs the batch's per-pixel code (same bits)."""
        c = _lib.Settings()
        check(lib().ctlproj_batch_shape(self.h, C.byref(c)))
        return Settings(raySampling=raySampling, tile=(c.tile_w, c.tile_h), unroll=c.unroll,
                        blocks_per_cu=c.blocks_per_cu, lds_chunk=c.lds_chunk, lds_texels=c.lds_texels)

    def close(self) -> None:
        if self.h:
            lib().ctlproj_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Request loop over resident volumes (ctlproj_server_*): admission by HBM footprint,
    continuous batching across volumes and detector sizes, p50/p99 latency."""

    def __init__(self, hbm_budget: int, max_launch_pixels: int = 0, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        self.h = C.c_void_p()
        check(lib().ctlproj_server_create(self.ctx.h, int(hbm_budget), int(max_launch_pixels),
                                          C.byref(self.h)))
        self._keep = {}

    def submit(self, volume: ResidentVolume, pmats: np.ndarray, channels: int, rows: int,
               out: np.ndarray, settings: Settings | None = None) -> int:
        """Queues a projection into `out` (float32 host array of views*modules*rows*channels,
        preferably pinned); returns the request id."""
        pm = np.ascontiguousarray(pmats, np.float64)
        if pm.ndim == 3:
            pm = pm[:, None]
        V, M = pm.shape[:2]
        s = (settings or Settings())._c()
        rid = C.c_uint64()
        check(lib().ctlproj_server_submit(self.h, volume.h, _lib.dptr(pm), V, M, channels, rows,
                                          C.byref(s), C.c_void_p(out.ctypes.data), out.size,
                                          C.byref(rid)))
        self._keep[rid.value] = (volume, out)
        return rid.value

    def submit_ptr(self, volume: ResidentVolume, pmats: np.ndarray, channels: int, rows: int,
                   out_ptr: int, count: int, settings: Settings | None = None) -> int:
        pm = np.ascontiguousarray(pmats, np.float64)
        if pm.ndim == 3:
            pm = pm[:, None]
        V, M = pm.shape[:2]
        s = (settings or Settings())._c()
        rid = C.c_uint64()
        check(lib().ctlproj_server_submit(self.h, volume.h, _lib.dptr(pm), V, M, channels, rows,
                                          C.byref(s), C.c_void_p(out_ptr), count, C.byref(rid)))
        self._keep[rid.value] = (volume,)
        return rid.value

    def wait(self, rid: int, timeout: float = -1.0) -> float:
        ms = C.c_double()
        try:
            check(lib().ctlproj_server_wait(self.h, rid, timeout, C.byref(ms)))
        finally:
            if timeout < 0:
                self._keep.pop(rid, None)
        return ms.value

    def drain(self) -> None:
        check(lib().ctlproj_server_drain(self.h))
        self._keep.clear()

    def stats(self) -> dict:
        st = _lib.ServerStats()
        check(lib().ctlproj_server_get_stats(self.h, C.byref(st)))
        return {n: getattr(st, n) for n, _ in _lib.ServerStats._fields_}

    def reset_stats(self) -> None:
        check(lib().ctlproj_server_reset_stats(self.h))

    def close(self) -> None:
        if self.h:
            lib().ctlproj_server_destroy(self.h)
            self.h = C.c_void_p()
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune_table() -> str:
    n = C.c_size_t()
    check(lib().ctlproj_autotune_table(None, 0, C.byref(n)))
    buf = C.create_string_buffer(n.value)
    check(lib().ctlproj_autotune_table(buf, n.value, None))
    return buf.value.decode()


def ellipsoid_phantom(dims, voxel_mm, n_ellipsoids, seed, semi_axes=(4.0, 24.0),
                      mu=(0.005, 0.03), background=0.02) -> np.ndarray:
    """Seeded random-ellipsoid phantom, shape (Z, Y, X) float32."""
    x, y, z = dims
    out = np.empty((z, y, x), np.float32)
    check(lib().ctlproj_phantom_ellipsoids(_lib.u32arr(*dims), _lib.f32arr(*voxel_mm), n_ellipsoids,
                                           seed, semi_axes[0], semi_axes[1], mu[0], mu[1],
                                           background, _lib.fptr(out)))
    return out
