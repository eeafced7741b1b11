"""Host-side rules of VolumeSlicer / RadonTransform3D / VolumeResampler; no device needed.

Expected values are worked out from the reference's formulas: the slicer's side is
ceil(1.4142 * largest dimension) (volumeslicer.cpp:224-229); the Radon transform rounds it up to
a multiple of 16 (radontransform3d.cpp:278-301) and keeps the template plane's extent when the
resolution changes (:239-251); SamplingRange (coordinates.h:40-88) spaces linspace samples
(end - start) / (n - 1) apart and centres on (start + end) / 2.
"""
import ctypes

import numpy as np
import pytest

from ctl_amd import SamplingRange, _lib


def slicer_side(dims):
    side = ctypes.c_uint32()
    assert _lib.lib().ctlproj_slicer_default_side(_lib.u32arr(*dims), ctypes.byref(side)) == _lib.OK
    return side.value


def radon_default(dims, voxel):
    side, reso = ctypes.c_uint32(), ctypes.c_float()
    st = _lib.lib().ctlproj_radon3d_default_slice(_lib.u32arr(*dims), _lib.f32arr(*voxel), ctypes.byref(side),
                                                  ctypes.byref(reso))
    assert st == _lib.OK
    return side.value, reso.value


@pytest.mark.parametrize("dims,side", [((256, 256, 256), 363), ((10, 4, 3), 15), ((1, 1, 1), 2),
                                       ((64, 100, 80), 142), ((100, 1, 1), 142)])
def test_slicer_default_side(dims, side):
    assert slicer_side(dims) == side


@pytest.mark.parametrize("dims,voxel,side,reso", [
    ((256, 256, 256), (1.0, 1.0, 1.0), 368, 1.0),   # 363 -> 368
    ((10, 4, 3), (0.5, 0.25, 2.0), 16, 0.25),        # 15 -> 16
    ((32, 8, 8), (1.0, 1.0, 1.0), 48, 1.0),          # 45.25 -> 46 -> 48
    ((16, 16, 16), (2.0, 3.0, 1.5), 32, 1.5),        # 22.63 -> 23 -> 32
])
def test_radon_default_slice(dims, voxel, side, reso):
    assert radon_default(dims, voxel) == (side, pytest.approx(reso))


@pytest.mark.parametrize("side,res,new,want", [
    (368, 1.0, 0.5, 736),    # factor 2: 736 is a multiple of 16
    (368, 1.0, 2.0, 192),    # 184 -> 192
    (48, 1.0, 3.0, 16),      # 16
    (32, 1.0, 0.7, 48),      # 45.71 -> 45 -> 48
])
def test_radon_side_for_resolution(side, res, new, want):
    out = ctypes.c_uint32()
    assert _lib.lib().ctlproj_radon3d_side_for_resolution(side, res, new, ctypes.byref(out)) == _lib.OK
    assert out.value == want


def test_radon_side_for_resolution_rejects_bad_resolution():
    out = ctypes.c_uint32()
    assert _lib.lib().ctlproj_radon3d_side_for_resolution(32, 1.0, 0.0, ctypes.byref(out)) == _lib.ERR_INVALID_ARGUMENT
    assert "resolution" in _lib.last_error()


def test_sampling_range():
    r = SamplingRange(-1.0, 2.0)
    assert r.spacing(4) == pytest.approx(1.0) and r.spacing(1) == 0.0
    assert r.center() == pytest.approx(0.5) and r.width() == pytest.approx(3.0)
    np.testing.assert_array_equal(r.linspace(4), np.float32([-1.0, 0.0, 1.0, 2.0]))
    np.testing.assert_array_equal(r.linspace(1), np.float32([-1.0]))
