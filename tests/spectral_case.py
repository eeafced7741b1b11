"""The case of the reference's spectral-extension test (testing/src/projectortest.cpp:41-147), shared
by the oracle test and the GPU test that reproduce its stored outputs: a flat panel of 50 x 50
pixels of 1 mm on a C-arm of 1200 mm, a short scan of 5 views at 750 mm (180 degrees + fan
angle), a water ball of 40 mm radius (density 1.0) and a cortical-bone ball of 30 mm radius
(density 0.3), both of 0.5 mm voxels and centred, the ray caster at its default settings. The
spectral inputs (bins, per-view intensities, mass attenuation coefficients) are the fixture
written by golden/make_spectral_inputs.py from the reference's spectrum model and database."""
from __future__ import annotations

import json
import math
from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden"


def inputs() -> dict:
    return json.loads((GOLDEN / "spectral_extension_inputs.json").read_text())


def ball(radius: float, voxel: float, value: float) -> np.ndarray:
    """VoxelVolume<float>::ball (voxelvolume.tpp:199-211, grindBall :1047-1068): ceil(2 r / voxel)
    voxels per axis, value where the squared distance of the voxel index from the centre index
    (n - 1) / 2 is at most (r / voxel)^2 (FP32), 0 elsewhere. Returns (Z, Y, X)."""
    n = int(math.ceil(np.float32(2.0) * np.float32(radius) / np.float32(voxel)))
    c = np.float32(n - 1) / np.float32(2.0)
    i = np.arange(n, dtype=np.float32) - c
    r2 = (np.float32(radius) / np.float32(voxel)) ** 2
    d2 = i[:, None, None] ** 2 + i[None, :, None] ** 2 + i[None, None, :] ** 2
    return np.where(d2 > r2, np.float32(0.0), np.float32(value)).astype(np.float32)


def geometry() -> np.ndarray:
    from ctl_amd.projector import (AcquisitionSetup, CarmGantry, CTSystem, FlatPanelDetector, GeometryEncoder,
                                   ShortScanTrajectory)
    setup = AcquisitionSetup(CTSystem(CarmGantry(1200.0), FlatPanelDetector((50, 50), (1.0, 1.0))), 5)
    setup.applyPreparationProtocol(ShortScanTrajectory(750.0))
    return GeometryEncoder.encodeFullGeometry(setup)


def expected(name: str) -> np.ndarray:
    """The reference's stored projections, [view][row][channel]."""
    from ctl_amd import fileio
    return fileio.load_den(GOLDEN / f"spectral_{name}.den")


def stats(got, want) -> tuple[float, float, float]:
    """projectionMean / projectionVariance of got - want (projectortest.cpp:289-330: mean and
    population variance over every pixel) and the largest |difference|."""
    d = np.asarray(got, np.float64).ravel() - np.asarray(want, np.float64).ravel()
    m = float(d.mean())
    return m, float(((d - m) ** 2).mean()), float(np.abs(d).max())
