"""Settings::optimizedFor (raycasterprojector.cpp:482-521) through the C ABI; no device needed.

Expected values are worked out by hand from the reference's rules: the ray grid is the preset
(1:2, 1:3, 1:4, 2:3, 3:4 or 1:1) closest to the pixel's aspect, oriented with its longer side
along the pixel's longer side; voxels larger than a sub-pixel give an upsampling factor
floor(voxel / sub-pixel); voxels smaller than a sub-pixel multiply the rays per pixel by
floor(sub-pixel / voxel).
"""
import ctypes

import pytest

from ctl_amd import _lib


def optimized(voxel, pw, ph):
    s = _lib.Settings()
    vox = (ctypes.c_float * 3)(*voxel)
    st = _lib.lib().ctlproj_settings_optimized_for(vox, pw, ph, ctypes.byref(s))
    return st, s


@pytest.mark.parametrize("voxel,pixel,rpp,up", [
    ((1.0, 1.0, 1.0), (1.0, 1.0), (1, 1), 1),     # square pixel, voxel = pixel
    ((1.0, 1.0, 1.0), (1.0, 2.0), (1, 2), 1),     # tall pixel: preset 1:2, two rays along rows
    ((1.0, 1.0, 1.0), (2.0, 1.0), (2, 1), 1),     # wide pixel: the same preset turned
    ((0.25, 0.5, 0.5), (1.0, 1.0), (4, 4), 1),    # smallest voxel a quarter pixel: 4x more rays
    ((3.0, 3.0, 3.0), (1.0, 1.0), (1, 1), 3),     # voxel three pixels: upsampling 3
    ((0.5, 0.5, 0.5), (1.0, 3.0), (2, 6), 1),     # preset 1:3, sub-pixel 1, voxel 0.5: x2
    ((1.0, 1.0, 1.0), (3.0, 4.0), (3, 4), 1),     # exact preset 3:4
    ((1.0, 1.0, 1.0), (0.6, 1.0), (2, 3), 3),     # aspect 0.6: preset 2:3 (dev 0.067); sub-pixel 0.3: x3
    ((2.0, 2.0, 2.0), (1.0, 4.0), (1, 4), 2),     # preset 1:4, sub-pixel 1: upsampling 2
])
def test_optimized_for(voxel, pixel, rpp, up):
    st, s = optimized(voxel, *pixel)
    assert st == _lib.OK
    assert tuple(s.rays_per_pixel) == rpp
    assert s.volume_upsampling == up
    # the other fields keep the reference defaults (raycasterprojector.h:31-34)
    assert s.ray_sampling == pytest.approx(0.3) and s.interpolate == 1


def test_optimized_for_singular_pixel():
    st, _ = optimized((1.0, 1.0, 1.0), 0.0, 1.0)  # the reference throws "pixel dimensions are singular"
    assert st == _lib.ERR_INVALID_ARGUMENT
    assert "singular" in _lib.last_error()
