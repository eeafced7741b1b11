"""Multi-projection callers on resident volumes (SURVEY rank-2 callers of the projector):
projectComposite (abstractprojector.h:99-112) accumulated on the device, and
ArealFocalSpotExtension (arealfocalspotextension.cpp:82-166, 257-292) as one launch over all
focal-spot samples plus a device average. Checked against per-volume projections and against the
oracle projected at each shifted geometry and averaged in FP64."""
import numpy as np
import pytest

import oracle
from ctl_amd import (AcquisitionSetup, ArealFocalSpotExtension, CarmGantry, CTSystem, FlatPanelDetector,
                     GeometryEncoder, RayCasterProjector, ResidentVolume, ShortScanTrajectory, VoxelVolume)
from ctl_amd import workloads as W

pytestmark = pytest.mark.gpu


def small_setup(views=4, spot=(0.0, 0.0)):
    setup = AcquisitionSetup(CTSystem(CarmGantry(1200.0), FlatPanelDetector((256, 128), (1.0, 1.0)), spot), views)
    setup.applyPreparationProtocol(ShortScanTrajectory(750.0, 0.0, 270.0 * np.pi / 180.0))
    return setup


def test_project_composite_is_the_sum_of_the_projections():
    w = W.small(views=4)
    a = VoxelVolume(w.volume, (1.0, 1.0, 1.0))
    b = VoxelVolume(np.ascontiguousarray(w.volume[::-1, ::-1, :] * 0.5), (0.8, 0.8, 0.8), (3.0, -2.0, 1.0))
    ra, rb = ResidentVolume(a), ResidentVolume(b, "f16")
    p = RayCasterProjector()
    p.settings().raySampling = 0.1
    p.configure(small_setup())
    pa, pb = p.project(ra).data, p.project(rb).data
    comp = p.projectComposite([ra, rb]).data
    assert np.array_equal(comp, pa + pb)  # FP32 sum in volume order
    assert p.isLinear()


@pytest.mark.parametrize("low_extinction", [False, True])
def test_areal_focal_spot_against_oracle(low_extinction):
    w = W.small(views=4)
    spot = (2.0, 1.0)
    setup = small_setup(spot=spot)
    nested = RayCasterProjector()
    nested.settings().raySampling = 0.1
    ext = ArealFocalSpotExtension(nested, (2, 2), low_extinction)
    ext.configure(setup)
    got = ext.project(VoxelVolume(w.volume, (1.0, 1.0, 1.0))).data
    assert got.shape == (4, 1, 128, 256) and ext.isLinear() == low_extinction
    subs = []
    for x, y in ext.discretizationGrid():
        pm = GeometryEncoder.encodeFullGeometry(setup, (x * spot[0], y * spot[1], 0.0))
        r32, _ = oracle.project(w.volume, (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), pm, 256, 128, 0.1)
        subs.append(r32.astype(np.float64))
    subs = np.stack(subs)
    ref = subs.mean(0) if low_extinction else -np.log(np.exp(-subs).mean(0))
    e = oracle.rel_l2(ref, got)
    print(f"areal focal spot (low extinction {low_extinction}): rel L2 vs ref32 {e:.3e}")
    assert e <= 1e-4
    # the samples differ: the extension is not the projection of the centre geometry
    p0 = RayCasterProjector()
    p0.settings().raySampling = 0.1
    p0.configure(setup)
    assert oracle.rel_l2(p0.project(VoxelVolume(w.volume, (1.0, 1.0, 1.0))).data, got) > 1e-4
