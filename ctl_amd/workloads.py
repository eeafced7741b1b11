"""Benchmark and parity workloads: seeded ellipsoid phantoms in CTL acquisition geometries.

  small    64^3 x 1 mm, 10 ellipsoids; C-arm span 1200 mm, source-isocentre 750 mm, 36 views over
           360 deg (10 deg apart), flat panel 256x128 @ 1 mm, ray step 0.1 voxel.
  real     256^3 x 0.5 mm, 20 ellipsoids; same C-arm, 360 views 1 deg apart, flat panel 1024x768
           @ 0.4 mm, ray step 0.2 voxel.
  helical  512x512x256 @ 0.7/0.7/1.0 mm, 40 ellipsoids (0-0.04 /mm); tubular gantry SDD 1085 mm,
           SID 595 mm, 2000 views over 2 turns at pitch 1.0, cylindrical detector of 40 modules of
           16x64 pixels @ 1 mm, ray step 0.25 voxel.
A full circle of n views uses ShortScanTrajectory(750, 0, (n-1)/n * 360 deg) because the protocol
spaces views by span/(n-1) (trajectories.cpp:195). Phantom seeds are 1000*config + instance.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .projector import (DEG, AcquisitionSetup, CarmGantry, CTSystem, CylindricalDetector,  # noqa: F401
                        FlatPanelDetector, GeometryEncoder, HelicalTrajectory, ShortScanTrajectory,
                        TubularGantry, ellipsoid_phantom)


@dataclass
class Workload:
    name: str
    volume: np.ndarray  # (Z, Y, X) float32
    voxel_mm: tuple
    offset_mm: tuple
    pmats: np.ndarray   # (V, M, 3, 4)
    channels: int
    rows: int
    ray_sampling: float

    @property
    def views(self) -> int:
        return self.pmats.shape[0]

    @property
    def modules(self) -> int:
        return self.pmats.shape[1]

    @property
    def dims(self):
        z, y, x = self.volume.shape
        return (x, y, z)


def carm_circle(n_views: int, pixels, pitch_mm, span=1200.0, sid=750.0) -> np.ndarray:
    setup = AcquisitionSetup(CTSystem(CarmGantry(span), FlatPanelDetector(pixels, pitch_mm)), n_views)
    setup.applyPreparationProtocol(ShortScanTrajectory(sid, 0.0, 360.0 * DEG * (n_views - 1) / n_views))
    return GeometryEncoder.encodeFullGeometry(setup)


def helical_geometry(n_views=2000, turns=2.0, pitch=1.0, sdd=1085.0, sid=595.0,
                     modules=40, module_px=(16, 64), pixel_mm=1.0) -> np.ndarray:
    # contiguous flat modules whose centres lie on a circle of radius SDD about the source:
    # (w/2) * cot(alpha/2) = SDD
    w = module_px[0] * pixel_mm
    alpha = 2.0 * math.atan(0.5 * w / sdd)
    det = CylindricalDetector(module_px, (pixel_mm, pixel_mm), modules, alpha, 0.0)
    coverage_iso = module_px[1] * pixel_mm * sid / sdd
    views_per_turn = n_views / turns
    feed = pitch * coverage_iso / views_per_turn
    setup = AcquisitionSetup(CTSystem(TubularGantry(sdd, sid), det), n_views)
    setup.applyPreparationProtocol(HelicalTrajectory(360.0 * DEG / views_per_turn, feed,
                                                     -0.5 * feed * n_views, 0.0))
    return GeometryEncoder.encodeFullGeometry(setup)


def small(views: int = 36) -> Workload:
    vol = ellipsoid_phantom((64, 64, 64), (1.0, 1.0, 1.0), 10, 1000)
    return Workload("small: 64^3 f32, 36 views, 256x128, step 0.1 vox", vol, (1.0, 1.0, 1.0),
                    (0.0, 0.0, 0.0), carm_circle(views, (256, 128), (1.0, 1.0)), 256, 128, 0.1)


def real(views: int = 360, n: int = 256, panel=(1024, 768)) -> Workload:
    vs = 0.5 * 256 / n
    vol = ellipsoid_phantom((n, n, n), (vs, vs, vs), 20, 2000)
    pitch = 0.4 * 1024 / panel[0]
    return Workload(f"real: {n}^3, {views} views, {panel[0]}x{panel[1]}, step 0.2 vox", vol,
                    (vs, vs, vs), (0.0, 0.0, 0.0), carm_circle(views, panel, (pitch, pitch)),
                    panel[0], panel[1], 0.2)


def helical(views: int = 2000, dims=(512, 512, 256)) -> Workload:
    vox = (0.7 * 512 / dims[0], 0.7 * 512 / dims[1], 1.0 * 256 / dims[2])
    vol = ellipsoid_phantom(dims, vox, 40, 3000, mu=(0.0, 0.04))
    return Workload(f"helical: {dims[0]}x{dims[1]}x{dims[2]}, {views} views, 40x16x64, step 0.25 vox",
                    vol, vox, (0.0, 0.0, 0.0), helical_geometry(views), 16, 64, 0.25)


def random_rotation(rng: np.random.Generator) -> np.ndarray:
    q = rng.normal(size=4)
    w, x, y, z = q / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def carm_arc_poses(rng: np.random.Generator, n_views: int, sid: float = 750.0,
                   span_deg=(30.0, 360.0)):
    """A C-arm pose trajectory: views on a circular arc about a random axis through the
    isocentre (random start rotation, random arc span); source at distance `sid`."""
    R0 = random_rotation(rng)
    span = np.deg2rad(rng.uniform(*span_deg))
    rots, pos = [], []
    for v in range(n_views):
        a = span * v / max(1, n_views - 1)
        c, s = np.cos(a), np.sin(a)
        Rz = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1.0]])
        R = R0 @ Rz @ _FIXED
        rots.append(R)
        pos.append(R @ np.array([0.0, 0.0, -sid]))
    return np.array(pos), np.array(rots)


# source frame of the C-arm protocols: Rz(pi/2) * Rx(-pi/2) (trajectories.cpp:187-188)
_FIXED = (np.array([[0, -1.0, 0], [1.0, 0, 0], [0, 0, 1.0]]) @
          np.array([[1.0, 0, 0], [0, 0, 1.0], [0, -1.0, 0]]))


@dataclass
class BatchInstance:
    volume: np.ndarray
    voxel_mm: tuple
    pmats: np.ndarray
    channels: int
    rows: int


def batched(n: int = 64, seed: int = 4000, edge=(64, 256), views=(1, 180), panel=(256, 768),
            fov_mm: float = 128.0, ray_sampling: float = 0.2):
    """Uneven instances: cubic phantoms of edge U{64..256} covering 128 mm, each with its own
    random C-arm arc of U{1..180} views and a square flat panel of U{256..768} pixels whose
    field of view at the isocentre is 1.2 x 128 mm (span 1200 mm, SID 750 mm)."""
    from .projector import AcquisitionSetup as _AS, CTSystem as _CS
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        N = int(rng.integers(edge[0], edge[1] + 1))
        V = int(rng.integers(views[0], views[1] + 1))
        P = int(rng.integers(panel[0], panel[1] + 1))
        vs = fov_mm / N
        vol = ellipsoid_phantom((N, N, N), (vs, vs, vs), 12, seed + k)
        pitch = 1.2 * fov_mm * (1200.0 / 750.0) / P
        pos, rot = carm_arc_poses(rng, V)
        setup = _AS(_CS(CarmGantry(1200.0), FlatPanelDetector((P, P), (pitch, pitch))))
        setup.setCarmPoses(pos, rot.reshape(V, 9))
        out.append(BatchInstance(vol, (vs, vs, vs), GeometryEncoder.encodeFullGeometry(setup), P, P))
    return out, ray_sampling


def served_drr_geometry(rng: np.random.Generator, n: int = 512, pitch_mm: float = 0.8,
                        sid: float = 750.0, span: float = 1200.0) -> np.ndarray:
    """One single-view DRR of the served workload: an n x n flat panel on a C-arm (span 1200 mm)
    whose pose is the reference pose (source on -z at `sid` from the isocentre, the C-arm
    protocols' source frame) turned by up to +-5 deg about each axis and moved by up to +-10 mm,
    as in iterative 2D/3D registration."""
    R = np.eye(3)
    for ax, a in enumerate(np.deg2rad(rng.uniform(-5, 5, 3))):
        c, s = np.cos(a), np.sin(a)
        M = np.eye(3)
        i, j = [(1, 2), (0, 2), (0, 1)][ax]
        M[i, i] = c
        M[j, j] = c
        M[i, j] = -s
        M[j, i] = s
        R = R @ M
    rot = R @ _FIXED
    pos = rot @ np.array([0, 0, -sid]) + rng.uniform(-10, 10, 3)
    setup = AcquisitionSetup(CTSystem(CarmGantry(span), FlatPanelDetector((n, n), (pitch_mm, pitch_mm))))
    setup.setCarmPoses(pos[None], rot.reshape(1, 9))
    return GeometryEncoder.encodeFullGeometry(setup)


def served_scan_geometry(views: int = 180, n: int = 768, pitch_mm: float = 0.533) -> np.ndarray:
    """One full scan of the served workload: a 180-view circle on 768 x 768 pixels."""
    return carm_circle(views, (n, n), (pitch_mm, pitch_mm))


def served_phantoms(count: int = 8, n: int = 256, voxel_mm: float = 0.5) -> list:
    """The served workload's resident phantoms (20 ellipsoids each, seeds 5000 + k)."""
    return [ellipsoid_phantom((n, n, n), (voxel_mm,) * 3, 20, 5000 + k) for k in range(count)]
