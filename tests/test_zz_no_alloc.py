"""The allocation-free projection entry point (ctlproj_project_device_noalloc), called through
the C ABI, against ctlproj_project_device: bit-identical results on two calls in a row, the
second with a smaller geometry, volume and detector (the static scratch keeps the first call's
records, and the second must not read any of them); oversized geometries and upsampling are
refused before any launch."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from ctl_amd import RayCasterProjector, ResidentVolume, VoxelVolume, CtlProjError
from ctl_amd import _lib
from ctl_amd import workloads as W

pytestmark = pytest.mark.gpu


def both(w, storage="f32", **settings):
    import torch
    p = RayCasterProjector()
    s = p.settings()
    s.raySampling = w.ray_sampling
    for k, v in settings.items():
        setattr(s, k, v)
    p.configure((w.pmats, w.channels, w.rows))
    p._push_settings()
    vol = ResidentVolume(VoxelVolume(w.volume, w.voxel_mm, w.offset_mm), storage)
    n = w.views * w.modules * w.rows * w.channels
    ref = torch.full((n,), -1.0, dtype=torch.float32, device="cuda")
    got = torch.full((n,), -2.0, dtype=torch.float32, device="cuda")
    L = _lib.lib()
    _lib.check(L.ctlproj_project_device(p.h, vol.h, C.c_void_p(ref.data_ptr()), n))
    _lib.check(L.ctlproj_projector_synchronize(p.h))
    _lib.check(L.ctlproj_project_device_noalloc(p.h, vol.h, C.c_void_p(got.data_ptr()), n))
    _lib.check(L.ctlproj_projector_synchronize(p.h))
    return ref.cpu().numpy(), got.cpu().numpy()


def test_noalloc_equals_project_device_twice_shrinking():
    # first call: 40-module helical geometry (480 records); second: 9 views of one odd module
    pm = W.helical_geometry(n_views=12, turns=0.05)
    vol = np.ascontiguousarray(W.ellipsoid_phantom((128, 128, 64), (2.8, 2.8, 4.0), 12, 3001,
                                                   mu=(0.0, 0.04)))
    big = W.Workload("helical small", vol, (2.8, 2.8, 4.0), (0.0, 0.0, 0.0), pm, 16, 64, 0.25)
    ref, got = both(big)
    print(f"noalloc helical 12 x 40 modules: max |d| {np.max(np.abs(ref - got)):.3e}")
    assert np.array_equal(ref, got) and ref.max() > 0
    dims, vs, off = (37, 29, 41), (1.3, 0.9, 1.7), (6.0, -4.5, 9.0)
    v2 = np.ascontiguousarray(W.ellipsoid_phantom(dims, vs, 9, 3101, semi_axes=(3.0, 14.0)))
    small = W.Workload("odd", v2, vs, off, W.carm_circle(9, (77, 53), (1.1, 0.9)), 77, 53, 0.15)
    for storage in ("f32", "f16"):
        ref, got = both(small, storage)
        print(f"noalloc odd 9 views {storage}: max |d| {np.max(np.abs(ref - got)):.3e}")
        assert np.array_equal(ref, got) and ref.max() > 0
    # unstaged kernel and nearest mode take the scratch the same way
    ref, got = both(small, lds_chunk=-1)
    assert np.array_equal(ref, got)
    ref, got = both(small, interpolate=False)
    assert np.array_equal(ref, got)


def test_noalloc_refuses_what_would_allocate_or_overflow():
    import torch
    w = W.small(views=4)
    vol = ResidentVolume(VoxelVolume(w.volume, w.voxel_mm))
    out = torch.zeros(4 * 128 * 256, dtype=torch.float32, device="cuda")
    p = RayCasterProjector()
    p.settings().raySampling = w.ray_sampling
    p.settings().volumeUpSampling = 2
    p.configure((w.pmats, w.channels, w.rows))
    with pytest.raises(CtlProjError) as ei:
        p.project_device_noalloc(vol, out.data_ptr(), out.numel())
    assert ei.value.status == _lib.ERR_INVALID_ARGUMENT and "upsampling" in str(ei.value)
    # one (view, module) pair more than the scratch holds: refused on the host
    n = 131072 + 1
    q = RayCasterProjector()
    q.settings().raySampling = w.ray_sampling
    q.configure((np.repeat(w.pmats[:1], n, axis=0), 1, 1))
    big = torch.zeros(n, dtype=torch.float32, device="cuda")
    with pytest.raises(CtlProjError) as ei:
        q.project_device_noalloc(vol, big.data_ptr(), big.numel())
    assert ei.value.status == _lib.ERR_INVALID_ARGUMENT and "131072" in str(ei.value)
    # exactly the capacity works (1x1-pixel modules) and equals the allocating path
    q.configure((np.repeat(w.pmats[:1], n - 1, axis=0), 1, 1))
    ref = torch.full((n - 1,), -1.0, dtype=torch.float32, device="cuda")
    got = torch.full((n - 1,), -2.0, dtype=torch.float32, device="cuda")
    q.project_device(vol, ref.data_ptr(), ref.numel())
    q.synchronize()
    q.project_device_noalloc(vol, got.data_ptr(), got.numel())
    q.synchronize()
    assert torch.equal(ref, got)
