"""VolumeSlicer, VolumeResampler and RadonTransform3D on resident volumes against the oracle.

The oracle restates the reference's host geometry and OpenCL kernels (ctl_oracle_processing.c);
the HIP path must agree with ref32 to within FP32 rounding of the sample positions. Bars:
  * slices and resampled values: max |GPU - ref32| <= 1e-4 on volumes with values in [0, 1)
    (a sample position moved by a few ulps moves the value by at most that times the gradient);
  * plane integrals: max |GPU - ref32| <= 2e-5 * max |ref64| and the same against ref64;
  * FP16/BF16 storage: the same bars against the oracle run on the volume rounded to that type
    (the interpolation is FP32 for every storage type);
  * axis planes through voxel centres: slices bit-exact, plane integrals equal to the layer sums.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from ctl_amd import (RadonTransform3D, ResidentVolume, SamplingRange, VolumeResampler, VolumeSlicer,
                     VoxelVolume, _lib)
from ctl_amd._lib import CtlProjError

pytestmark = pytest.mark.gpu

STORAGES = ["f32", "f16", "bf16"]


def rounded(data, storage):
    """The volume as the resident storage holds it."""
    if storage == "f16":
        return data.astype(np.float16).astype(np.float32)
    if storage == "bf16":
        import torch
        return torch.from_numpy(data).to(torch.bfloat16).to(torch.float32).numpy()
    return data


def random_volume(seed, shape=(30, 34, 38), voxel=(0.9, 1.1, 0.8), offset=(1.5, -2.0, 0.7)):
    rng = np.random.default_rng(seed)
    return VoxelVolume(rng.random(shape, dtype=np.float32), voxel, offset)


def random_normals(rng, n):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


@pytest.fixture(params=["interleaved", "plain"])
def layout(request, monkeypatch):
    if request.param == "plain":
        monkeypatch.setenv("CTLPROJ_VOLUME_LAYOUT", "plain")
    else:
        monkeypatch.delenv("CTLPROJ_VOLUME_LAYOUT", raising=False)
    return request.param


@pytest.mark.parametrize("storage", STORAGES)
def test_slicer_against_oracle(storage, layout):
    vol = random_volume(1)
    ref_data = rounded(vol.data, storage)
    s = VolumeSlicer(vol, storage)
    assert s.sliceDimensions() == (54, 54) and s.sliceResolution() == pytest.approx(0.8)
    rng = np.random.default_rng(2)
    normals = np.vstack([np.eye(3), random_normals(rng, 5)])
    dists = rng.uniform(-8, 8, len(normals))
    got = s.slices(normals, dists)
    assert got.shape == (8, 54, 54)
    worst = 0.0
    for k, (n, d) in enumerate(zip(normals, dists)):
        want = oracle.slice_plane(ref_data, vol.voxel_size, vol.offset, n, d, 54, 54, 0.8)
        worst = max(worst, float(np.abs(got[k] - want).max()))
    print(f"slicer {storage} {layout}: max |GPU - ref32| {worst:.2e}")
    assert worst <= 1e-4
    # the single-plane overloads, with another slice size and resolution
    s.setSliceDimensions((31, 17))
    s.setSliceResolution(1.3)
    one = s.slice(0.4, 1.1, 2.5)
    n = np.array([np.sin(1.1) * np.cos(0.4), np.sin(1.1) * np.sin(0.4), np.cos(1.1)])
    want = oracle.slice_plane(ref_data, vol.voxel_size, vol.offset, n, 2.5, 31, 17, 1.3)
    assert one.shape == (17, 31) and np.abs(one - want).max() <= 1e-4


def test_slicer_axis_planes_are_exact():
    rng = np.random.default_rng(3)
    X, Y, Z = 6, 8, 5
    vol = VoxelVolume(rng.random((Z, Y, X), dtype=np.float32) + 1.0)
    s = VolumeSlicer(vol)
    for normal, (W, H) in [((0, 0, 1), (12, 8)), ((1, 0, 0), (9, 10)), ((0, 1, 0), (9, 8))]:
        s.setSliceDimensions((W, H))
        n_axis = int(np.argmax(normal))
        for layer in range(vol.dims[n_axis]):
            d = layer - (vol.dims[n_axis] - 1) / 2.0
            want = oracle.slice_plane(vol.data, (1, 1, 1), (0, 0, 0), normal, d, W, H, 1.0)
            assert np.array_equal(s.slice(np.array(normal, float), d), want)


@pytest.mark.parametrize("storage", STORAGES)
def test_resampler_against_oracle(storage, layout):
    vol = random_volume(4, shape=(20, 24, 28))
    ref_data = rounded(vol.data, storage)
    r = VolumeResampler(vol, storage=storage)
    rng = np.random.default_rng(5)
    lo, hi = r._ranges[0::2], r._ranges[1::2]
    p1 = rng.uniform(lo[0] - 2, hi[0] + 2, 17).astype(np.float32)
    p2 = rng.uniform(lo[1] - 2, hi[1] + 2, 13).astype(np.float32)
    p3 = rng.uniform(lo[2] - 2, hi[2] + 2, 11).astype(np.float32)
    got = r.resample(p1, p2, p3).data
    want = oracle.resample(ref_data, r._ranges, p1, p2, p3)
    assert got.shape == (11, 13, 17)
    assert np.abs(got - want).max() <= 1e-4
    pts = np.column_stack([rng.uniform(lo[k] - 3, hi[k] + 3, 5000) for k in range(3)]).astype(np.float32)
    got = r.sample(pts)
    want = oracle.sample_points(ref_data, r._ranges, pts)
    print(f"resampler {storage} {layout}: max |GPU - ref32| {np.abs(got - want).max():.2e}")
    assert np.abs(got - want).max() <= 1e-4
    # custom ranges (the registration use: a normalised coordinate system)
    r.setSamplingRanges((-1.0, 1.0), SamplingRange(0.0, 5.0), (10.0, 20.0))
    u = np.column_stack([rng.uniform(-1.2, 1.2, 3000), rng.uniform(-0.5, 5.5, 3000),
                         rng.uniform(9, 21, 3000)]).astype(np.float32)
    want = oracle.sample_points(ref_data, r._ranges, u)
    assert np.abs(r.sample(u) - want).max() <= 1e-4
    assert r.volVoxSize() == pytest.approx((2.0 / 27, 5.0 / 23, 10.0 / 19))
    assert r.volOffset() == pytest.approx((0.0, 2.5, 15.0))


def test_resampler_identity_at_voxel_centres():
    vol = random_volume(6, shape=(7, 9, 11))
    r = VolumeResampler(vol)
    centres = [np.linspace(r._ranges[2 * k], r._ranges[2 * k + 1], n).astype(np.float32)
               for k, n in enumerate(vol.dims)]
    got = r.resample(*centres).data
    np.testing.assert_allclose(got, vol.data, atol=2e-6)


def test_sample_device_matches_host():
    import torch
    vol = random_volume(7)
    r = VolumeResampler(vol)
    rng = np.random.default_rng(8)
    lo, hi = r._ranges[0::2], r._ranges[1::2]
    pts = np.column_stack([rng.uniform(lo[k], hi[k], 4096) for k in range(3)]).astype(np.float32)
    d_pts = torch.from_numpy(pts).cuda()
    d_out = torch.empty(4096, device="cuda")
    r.sample_device(d_pts.data_ptr(), 4096, d_out.data_ptr())
    assert np.array_equal(d_out.cpu().numpy(), r.sample(pts))


@pytest.mark.parametrize("storage", STORAGES)
def test_plane_integrals_against_oracle(storage, layout):
    vol = random_volume(9, shape=(22, 26, 30))
    ref_data = rounded(vol.data, storage)
    rt = RadonTransform3D(vol, storage)
    side, reso = rt.sliceDimensions()[0], rt.sliceResolution()
    assert side == 48 and reso == pytest.approx(0.8)
    rng = np.random.default_rng(10)
    normals = np.vstack([np.eye(3), random_normals(rng, 9)])
    dist = np.concatenate([rng.uniform(-14, 14, 18), [0.0, 40.0]]).astype(np.float32)  # 40 mm: outside
    got = rt.planeIntegrals(normals, dist)
    r32 = oracle.plane_integrals(ref_data, vol.voxel_size, vol.offset, normals, dist, side, side, reso)
    r64 = oracle.plane_integrals(ref_data, vol.voxel_size, vol.offset, normals, dist, side, side, reso, fp64=True)
    scale = np.abs(r64).max()
    e32, e64 = np.abs(got - r32).max() / scale, np.abs(got - r64).max() / scale
    print(f"plane integrals {storage} {layout}: max err / max |R| vs ref32 {e32:.2e}, vs ref64 {e64:.2e}")
    assert got.shape == (20, 12) and e32 <= 2e-5 and e64 <= 2e-5
    assert np.all(got[-1] == 0.0)


@pytest.mark.parametrize("storage", STORAGES)
def test_plane_integrals_brick_and_texel_reads_agree(storage, monkeypatch):
    """By default the integrals are summed brick by brick from LDS, staged from the volume's FP32
    brick copy (8-byte texel bricks for a volume this small; CTLPROJ_PLANE_READS=pairs asks for the
    16-byte pair bricks that larger volumes get): both copies hold the same texels, so the sums
    agree bit for bit. CTLPROJ_PLANE_READS=texels (and CTLPROJ_PLANE_KERNEL=tiles over the brick
    copy) run the tile kernel, which gathers the same samples through the caches and adds them in
    another order: within 1e-6 of the largest integral."""
    vol = random_volume(31, shape=(26, 22, 30))
    rng = np.random.default_rng(32)
    normals = random_normals(rng, 7)
    dist = rng.uniform(-12, 12, 10).astype(np.float32)
    bricks = RadonTransform3D(vol, storage).planeIntegrals(normals, dist)
    monkeypatch.setenv("CTLPROJ_PLANE_READS", "pairs")
    pairs = RadonTransform3D(vol, storage).planeIntegrals(normals, dist)
    monkeypatch.setenv("CTLPROJ_PLANE_READS", "texels")
    texels = RadonTransform3D(vol, storage).planeIntegrals(normals, dist)
    monkeypatch.delenv("CTLPROJ_PLANE_READS")
    monkeypatch.setenv("CTLPROJ_PLANE_KERNEL", "tiles")
    tiles = RadonTransform3D(vol, storage).planeIntegrals(normals, dist)
    assert np.array_equal(bricks, pairs)
    for got in (texels, tiles):
        assert np.abs(got - bricks).max() <= 1e-6 * np.abs(bricks).max()
    side, reso = 48, 0.8
    r32 = oracle.plane_integrals(rounded(vol.data, storage), vol.voxel_size, vol.offset, normals, dist, side, side,
                                 reso)
    for got in (bricks, tiles):
        assert np.abs(got - r32).max() <= 2e-5 * np.abs(r32).max()


def test_plane_integrals_brick_kind_switch_on_one_volume(monkeypatch):
    """One resident volume asked for pair bricks after its texel bricks were made (the copy is
    replaced) and back: every brick-kernel result is the same bit for bit, the tile kernel's within
    1e-6."""
    vol = random_volume(33, shape=(20, 24, 18))
    rng = np.random.default_rng(34)
    normals = random_normals(rng, 5)
    dist = rng.uniform(-10, 10, 9).astype(np.float32)
    rt = RadonTransform3D(vol)
    first = rt.planeIntegrals(normals, dist)
    monkeypatch.setenv("CTLPROJ_PLANE_READS", "pairs")
    second = rt.planeIntegrals(normals, dist)
    monkeypatch.delenv("CTLPROJ_PLANE_READS")
    third = rt.planeIntegrals(normals, dist)
    monkeypatch.setenv("CTLPROJ_PLANE_READS", "texels")
    texels = rt.planeIntegrals(normals, dist)
    for got in (second, third):
        assert np.array_equal(got, first)
    assert np.abs(texels - first).max() <= 1e-6 * np.abs(first).max()


def test_plane_integrals_non_finite_volume_uses_tile_kernel(monkeypatch):
    """A volume with an infinite voxel has no fixed-point scale for the brick kernel's totals: the
    tile kernel sums its planes, and the planes through the voxel are not finite, as in the
    reference's float sums."""
    data = random_volume(35, shape=(14, 16, 12)).data.copy()
    data[6, 8, 7] = np.inf
    vol = VoxelVolume(data, (1, 1, 1), (0, 0, 0))
    normals = np.eye(3)
    dist = np.array([-3.0, 0.0, 3.0], np.float32)
    got = RadonTransform3D(vol).planeIntegrals(normals, dist)
    monkeypatch.setenv("CTLPROJ_PLANE_KERNEL", "tiles")
    tiles = RadonTransform3D(vol).planeIntegrals(normals, dist)
    assert np.array_equal(got, tiles, equal_nan=True)
    assert not np.isfinite(got).all() and np.isfinite(got).any()


@pytest.mark.parametrize("storage", STORAGES)
@pytest.mark.parametrize("reads", ["bricks", "pairs", "texels", "tiles"])
def test_plane_integrals_skip_empty_space_exactly(storage, reads, monkeypatch):
    """Both plane-integral kernels skip empty space with the volume's occupancy table: the brick
    kernel its bricks that hold no non-zero voxel, the tile kernel ("texels", "tiles") the
    distances of a tile whose samples read only empty 4^3-voxel bricks. Those samples are exactly
    0: on a phantom with empty space, and on a volume of a few non-zero voxels (at the grid's
    corners too), the integrals equal the ones taken over a volume without a table
    (CTLPROJ_EMPTY_SKIP=0) bit for bit, and the oracle."""
    from ctl_amd import workloads as W
    rng = np.random.default_rng(41)
    sparse = np.zeros((27, 31, 23), np.float32)
    for x, y, z in [(0, 0, 0), (26, 30, 22), (0, 30, 0), (26, 0, 22)] + [
            tuple(int(rng.integers(0, n)) for n in sparse.shape) for _ in range(6)]:
        sparse[x, y, z] = rng.uniform(0.5, 1.0)
    phantom = W.small(views=1).volume[::2, ::2, ::2].copy()
    assert (phantom == 0).mean() > 0.3
    if reads == "tiles":
        monkeypatch.setenv("CTLPROJ_PLANE_KERNEL", "tiles")
    elif reads != "bricks":
        monkeypatch.setenv("CTLPROJ_PLANE_READS", reads)
    normals = np.vstack([np.eye(3), random_normals(rng, 9)])
    for data, voxel, offset in ((sparse, (0.9, 1.1, 0.8), (1.5, -2.0, 0.7)), (phantom, (2.0, 2.0, 2.0), (0, 0, 0))):
        vol = VoxelVolume(np.ascontiguousarray(data.transpose(2, 1, 0)), voxel, offset)
        ext = max(n * v for n, v in zip(vol.dims, voxel))
        dist = np.concatenate([rng.uniform(-0.6 * ext, 0.6 * ext, 22), [0.0, 2 * ext]]).astype(np.float32)
        monkeypatch.setenv("CTLPROJ_EMPTY_SKIP", "1")
        on = RadonTransform3D(vol, storage)
        monkeypatch.setenv("CTLPROJ_EMPTY_SKIP", "0")
        off = RadonTransform3D(vol, storage)
        monkeypatch.delenv("CTLPROJ_EMPTY_SKIP")
        got, want = on.planeIntegrals(normals, dist), off.planeIntegrals(normals, dist)
        assert np.array_equal(got, want)
        side, reso = on.sliceDimensions()[0], on.sliceResolution()
        r32 = oracle.plane_integrals(rounded(vol.data, storage), vol.voxel_size, vol.offset, normals, dist, side,
                                     side, reso)
        assert np.abs(got - r32).max() <= 2e-5 * np.abs(r32).max()
        assert np.count_nonzero(got) > 0


def test_plane_integrals_axis_layer_sums_and_single_integral():
    rng = np.random.default_rng(11)
    X, Y, Z = 10, 12, 6
    data = rng.random((Z, Y, X), dtype=np.float32)
    offset = (2.0, -3.0, 1.0)
    rt = RadonTransform3D(VoxelVolume(data, (1, 1, 1), offset))
    rt.setSliceResolution(1.0)
    assert rt.sliceDimensions() == (32, 32)  # 17 -> 32 (radontransform3d.cpp:283)
    sums = [data.sum(axis=(0, 1)), data.sum(axis=(0, 2)), data.sum(axis=(1, 2))]
    for a in range(3):
        dist = (offset[a] + np.arange((X, Y, Z)[a]) - ((X, Y, Z)[a] - 1) / 2.0).astype(np.float32)
        got = rt.planeIntegrals(np.eye(3)[a:a + 1], dist)[:, 0]
        np.testing.assert_allclose(got, sums[a], rtol=2e-6)
    # planeIntegral(normal, d) and planeIntegral(azimuth, polar, d) (e_z: polar 0)
    assert rt.planeIntegral(np.array([0.0, 0.0, 1.0]), 1.5) == pytest.approx(float(sums[2][3]), rel=2e-6)
    assert rt.planeIntegral(0.0, 0.0, 1.5) == pytest.approx(float(sums[2][3]), rel=2e-6)


def test_sample_transform_grid_order():
    vol = random_volume(12, shape=(16, 18, 20))
    rt = RadonTransform3D(vol)
    azi = SamplingRange(0.0, np.pi)
    pol = SamplingRange(0.1, 3.0)
    dst = SamplingRange(-10.0, 10.0)
    R = rt.sampleTransform(azi, 7, pol, 5, dst, 9)
    assert R.dims == (7, 5, 9)
    assert R.voxel_size == pytest.approx((np.pi / 6, 2.9 / 4, 2.5))
    assert R.offset == pytest.approx((np.pi / 2, 1.55, 0.0))
    a, p, d = azi.linspace(7), pol.linspace(5), dst.linspace(9)
    normals = np.array([[np.sin(float(pp)) * np.cos(float(aa)), np.sin(float(pp)) * np.sin(float(aa)),
                         np.cos(float(pp))] for pp in p for aa in a])
    want = rt.planeIntegrals(normals, d).reshape(9, 5, 7)
    assert np.array_equal(R.data, want)
    side, reso = rt.sliceDimensions()[0], rt.sliceResolution()
    r32 = oracle.plane_integrals(vol.data, vol.voxel_size, vol.offset, normals, d, side, side, reso)
    assert np.abs(R.data.reshape(9, 35) - r32).max() <= 2e-5 * np.abs(r32).max()


def test_plane_integrals_device_output():
    import torch
    vol = random_volume(13)
    rt = RadonTransform3D(vol)
    normals = random_normals(np.random.default_rng(14), 4)
    dist = np.linspace(-5, 5, 11).astype(np.float32)
    out = torch.empty(11 * 4, device="cuda")
    side, reso = rt.sliceDimensions()[0], rt.sliceResolution()
    n = np.ascontiguousarray(normals)
    _lib.check(_lib.lib().ctlproj_radon3d_plane_integrals_device(
        rt._vol.h, _lib.dptr(n), 4, _lib.fptr(dist), 11, side, reso, C.c_void_p(out.data_ptr()), 44))
    assert np.array_equal(out.cpu().numpy().reshape(11, 4), rt.planeIntegrals(normals, dist))


@pytest.mark.parametrize("kernel", ["bricks", "tiles"])
def test_many_orientations_split_over_launches(kernel, monkeypatch):
    # enough orientations x distances that the tile kernel's host splits them over several launches
    if kernel == "tiles":
        monkeypatch.setenv("CTLPROJ_PLANE_KERNEL", "tiles")
    vol = random_volume(15, shape=(12, 12, 12), voxel=(1, 1, 1), offset=(0, 0, 0))
    rt = RadonTransform3D(vol)
    rng = np.random.default_rng(16)
    normals = random_normals(rng, 70000)  # more than 65536 orientations per launch
    dist = rng.uniform(-8, 8, 3).astype(np.float32)
    got = rt.planeIntegrals(normals, dist)
    pick = np.concatenate([rng.choice(70000, 10, replace=False), [65535, 65536, 69999]])
    side, reso = rt.sliceDimensions()[0], rt.sliceResolution()
    r32 = oracle.plane_integrals(vol.data, vol.voxel_size, vol.offset, normals[pick], dist, side, side, reso)
    assert np.abs(got[:, pick] - r32).max() <= 2e-5 * np.abs(r32).max()


def test_processing_error_paths():
    vol = ResidentVolume(random_volume(17, shape=(8, 8, 8)))
    L = _lib.lib()
    n = np.array([0.0, 0.0, 2.0])
    d = np.array([0.0])
    out = np.zeros(16, np.float32)
    assert L.ctlproj_slice(vol.h, _lib.dptr(n), _lib.dptr(d), 1, 4, 4, 1.0, _lib.fptr(out), 16) == _lib.ERR_INVALID_ARGUMENT
    assert "unit" in _lib.last_error()
    n = np.array([0.0, 0.0, 1.0])
    assert L.ctlproj_slice(vol.h, _lib.dptr(n), _lib.dptr(d), 1, 4, 4, 0.0, _lib.fptr(out), 16) == _lib.ERR_INVALID_ARGUMENT
    assert L.ctlproj_slice(vol.h, _lib.dptr(n), _lib.dptr(d), 1, 4, 4, 1.0, _lib.fptr(out), 15) == _lib.ERR_INVALID_ARGUMENT
    dist = np.zeros(2, np.float32)
    assert L.ctlproj_radon3d_plane_integrals(vol.h, _lib.dptr(n), 1, _lib.fptr(dist), 2, 16, 1.0, _lib.fptr(out),
                                             3) == _lib.ERR_INVALID_ARGUMENT
    bad = np.array([np.nan, 0.0, 1.0, 0.0, 0.0, 0.0], np.float32)
    pts = np.zeros(3, np.float32)
    assert L.ctlproj_sample(vol.h, _lib.fptr(bad), _lib.fptr(pts), 1, _lib.fptr(out)) == _lib.ERR_INVALID_ARGUMENT
    with pytest.raises(CtlProjError):
        VolumeSlicer(vol).slice(np.array([1.0, 1.0, 0.0]), 0.0)
    # empty requests succeed without touching the outputs
    assert L.ctlproj_slice(vol.h, _lib.dptr(n), _lib.dptr(d), 0, 4, 4, 1.0, None, 0) == _lib.OK
    assert L.ctlproj_radon3d_plane_integrals(vol.h, _lib.dptr(n), 1, _lib.fptr(dist), 0, 16, 1.0, None, 0) == _lib.OK


def test_cpp_mirror_processing_program(tmp_path):
    """The C++ mirrors (include/ctlproj.hpp: VolumeSlicer, VolumeResampler, RadonTransform3D),
    driven like CTL code, against the oracle."""
    import subprocess
    from pathlib import Path
    exe = Path(__file__).resolve().parent.parent / "ctl_amd" / "lib" / "processing_demo"
    out = tmp_path / "proc.f32"
    res = subprocess.run([str(exe), str(out)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    assert "slice=57x57 radon=64" in res.stdout  # ceil(1.4142 * 40) = 57; 57 -> 64
    raw = np.fromfile(out, np.float32)
    dims, voxel, offset = (40, 36, 32), (0.9, 1.1, 0.8), (1.5, -2.0, 0.7)
    n_vol = 40 * 36 * 32
    vol = raw[:n_vol].reshape(32, 36, 40)
    sl = raw[n_vol:n_vol + 57 * 57].reshape(57, 57)
    smp = raw[n_vol + 57 * 57:n_vol + 57 * 57 + 7]
    tr = raw[n_vol + 57 * 57 + 7:].reshape(4, 2, 3)
    n = np.array([np.sin(1.1) * np.cos(0.4), np.sin(1.1) * np.sin(0.4), np.cos(1.1)])
    want = oracle.slice_plane(vol, voxel, offset, n, 2.5, 57, 57, 0.8)
    assert np.abs(sl - want).max() <= 1e-4 * max(1.0, float(np.abs(vol).max()))
    ranges = []
    for k in range(3):
        half = np.float32(0.5) * np.float32(voxel[k]) * np.float32(dims[k] - 1)
        ranges += [np.float32(offset[k]) - half, np.float32(offset[k]) + half]
    pts = np.array([0, 0, 0, 5.2, -3.1, 4, -17, 16, -11, 1.5, -2, 0.7, 20, 0, 0, -9.3, 7.7, 2.2, 3.3, -18, 11.9],
                   np.float32).reshape(7, 3)
    assert np.abs(smp - oracle.sample_points(vol, ranges, pts)).max() <= 1e-4 * max(1.0, float(np.abs(vol).max()))
    normals = np.array([[np.sin(float(p)) * np.cos(float(a)), np.sin(float(p)) * np.sin(float(a)), np.cos(float(p))]
                        for p in np.float32([0.5, 1.5]) for a in np.float32([0.0, 1.0, 2.0])])
    r32 = oracle.plane_integrals(vol, voxel, offset, normals, np.float32([-6, -1, 0.5, 7]), 64, 64, 0.8)
    assert np.abs(tr.reshape(4, 6) - r32).max() <= 2e-5 * np.abs(r32).max()
