"""Pins the oracle's restatement of VolumeSlicer, VolumeResampler and RadonTransform3D.

The reference's tests hold no vectors for these classes, so the oracle is pinned by what the
mathematics fixes:
  * axis-aligned planes through voxel centres read the voxels themselves: the slice is a
    sub-array of the volume, turned the way the plane rotation turns the template plane. The
    rotation for each axis normal is worked out by hand from volumeslicer.cpp:207-217
    (e_z: slice x -> +y, slice y -> -x; e_x: x -> +z, y -> -y; e_y: x -> +z, y -> +x);
  * OpenCL's linear sampler reproduces a linear function exactly between voxel centres, so
    slices of the coordinate functions x, y, z return each pixel's world position: every pixel
    lies on the plane, neighbours are `resolution` apart, the rows and columns are orthogonal and
    right-handed about the normal;
  * resampling at the voxel centres is the identity; half a spacing beyond the last centre the
    sampler reads half the last voxel (border colour 0), a full spacing beyond it reads zero;
  * plane integrals of axis-aligned planes through voxel centres are the layer sums times the
    pixel area; over all distances they add up to the volume's mass, and their first moment
    puts the normal's projection of the centre of mass where it belongs.
"""
import numpy as np
import pytest

import oracle

AXIS_MAPS = {
    # normal: (world axis, sign) of slice x and of slice y, derived by hand
    (0.0, 0.0, 1.0): ((1, +1), (0, -1)),
    (1.0, 0.0, 0.0): ((2, +1), (1, -1)),
    (0.0, 1.0, 0.0): ((2, +1), (0, +1)),
}


def coordinate_volumes(dims, voxel, offset):
    """Three (Z, Y, X) volumes holding the world x, y and z of every voxel centre."""
    X, Y, Z = dims
    axes = [offset[k] + (np.arange(n) - (n - 1) / 2.0) * voxel[k] for k, n in enumerate((X, Y, Z))]
    zz, yy, xx = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    return [a.astype(np.float32) for a in (xx, yy, zz)]


@pytest.mark.parametrize("normal", list(AXIS_MAPS))
@pytest.mark.parametrize("fp64", [False, True])
def test_slice_axis_planes_read_voxels(normal, fp64):
    rng = np.random.default_rng(11)
    X, Y, Z = 6, 8, 5
    vol = rng.random((Z, Y, X), dtype=np.float32) + 1.0
    (ax_x, sx), (ax_y, sy) = AXIS_MAPS[normal]
    dims = (X, Y, Z)
    n_axis = int(np.argmax(normal))
    # slice sizes with the same parity as the volume extents they run along: pixels on centres
    W, H = dims[ax_x] + 4, dims[ax_y] + 2
    for layer in range(dims[n_axis]):
        d = layer - (dims[n_axis] - 1) / 2.0
        got = oracle.slice_plane(vol, (1, 1, 1), (0, 0, 0), normal, d, W, H, 1.0, fp64=fp64)
        want = np.zeros((H, W), np.float32)
        for y in range(H):
            for x in range(W):
                idx = [0, 0, 0]
                idx[n_axis] = layer
                idx[ax_x] = sx * (x - (W - 1) / 2.0) + (dims[ax_x] - 1) / 2.0
                idx[ax_y] = sy * (y - (H - 1) / 2.0) + (dims[ax_y] - 1) / 2.0
                assert all(float(v).is_integer() for v in idx)
                i, j, k = (int(v) for v in idx)
                if 0 <= i < X and 0 <= j < Y and 0 <= k < Z:
                    want[y, x] = vol[k, j, i]
        np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("fp64", [False, True])
def test_slice_oblique_geometry_from_linear_volumes(fp64):
    dims, voxel, offset = (20, 18, 16), (1.0, 1.5, 0.75), (2.0, -1.0, 0.5)
    vols = coordinate_volumes(dims, voxel, offset)
    n = np.array([0.3, -0.5, 0.8])
    n /= np.linalg.norm(n)
    d, W, H, reso = 1.25, 41, 37, 0.6
    P = np.stack([oracle.slice_plane(v, voxel, offset, n, d, W, H, reso, fp64=fp64) for v in vols], -1)
    # pixels whose eight taps are all voxels (a slice of ones reads 1 there): the sampler is exact
    ones = oracle.slice_plane(np.ones_like(vols[0]), voxel, offset, n, d, W, H, reso, fp64=fp64)
    idx = (P - np.asarray(offset)) / np.asarray(voxel) + (np.asarray(dims) - 1) / 2.0
    inner = (np.abs(ones - 1.0) < 1e-6) & np.all((idx > 0.01) & (idx < np.asarray(dims) - 1.01), axis=-1)
    assert inner.sum() > 300
    tol = 5e-6 if fp64 else 2e-4  # outputs are float32
    assert np.abs(P[inner] @ n - d).max() < tol
    both = inner[:-1, :-1] & inner[1:, :-1] & inner[:-1, 1:]
    ex = (P[:-1, 1:] - P[:-1, :-1])[both]  # one pixel along x
    ey = (P[1:, :-1] - P[:-1, :-1])[both]  # one pixel along y
    assert np.abs(np.linalg.norm(ex, axis=1) - reso).max() < tol * 10
    assert np.abs(np.linalg.norm(ey, axis=1) - reso).max() < tol * 10
    assert np.abs(np.sum(ex * ey, axis=1)).max() < tol * 10
    assert np.abs(np.cross(ex, ey) - reso * reso * n).max() < tol * 10
    assert inner[(H - 1) // 2, (W - 1) // 2]
    assert np.abs(P[(H - 1) // 2, (W - 1) // 2] - d * n).max() < tol  # centre pixel at d * n


@pytest.mark.parametrize("fp64", [False, True])
def test_resample_identity_order_and_border(fp64):
    rng = np.random.default_rng(12)
    X, Y, Z = 5, 4, 3
    vol = rng.random((Z, Y, X), dtype=np.float32) + 1.0
    voxel, offset = (0.5, 2.0, 1.0), (1.0, 0.0, -3.0)
    ranges = []
    for k, n in enumerate((X, Y, Z)):  # the default ranges, volumeresampler.cpp:74-83
        ranges += [np.float32(offset[k]) - np.float32(0.5) * np.float32(voxel[k]) * np.float32(n - 1),
                   np.float32(offset[k]) + np.float32(0.5) * np.float32(voxel[k]) * np.float32(n - 1)]
    ranges = np.asarray(ranges, np.float32)
    centres = [np.linspace(ranges[2 * k], ranges[2 * k + 1], n).astype(np.float32)
               for k, n in enumerate((X, Y, Z))]
    got = oracle.resample(vol, ranges, *centres, fp64=fp64)
    np.testing.assert_allclose(got, vol, rtol=0, atol=2e-6)
    # a point half a spacing beyond the last x centre reads half that voxel; a spacing beyond, 0
    sx = (ranges[1] - ranges[0]) / (X - 1)
    pts = np.array([[ranges[1] + 0.5 * sx, centres[1][2], centres[2][1]],
                    [ranges[1] + 1.0 * sx, centres[1][2], centres[2][1]],
                    [ranges[0] - 0.5 * sx, centres[1][0], centres[2][2]]], np.float32)
    got = oracle.sample_points(vol, ranges, pts, fp64=fp64)
    np.testing.assert_allclose(got, [0.5 * vol[1, 2, X - 1], 0.0, 0.5 * vol[2, 0, 0]], atol=2e-6)


@pytest.mark.parametrize("fp64", [False, True])
def test_sample_linear_function_exact(fp64):
    rng = np.random.default_rng(13)
    X, Y, Z = 9, 7, 6
    ranges = np.array([-4.0, 4.0, 0.0, 3.0, 10.0, 20.0], np.float32)
    ii, jj, kk = np.meshgrid(np.arange(X), np.arange(Y), np.arange(Z), indexing="ij")
    f = lambda i, j, k: 2.0 * i - 1.5 * j + 0.25 * k + 3.0  # linear in the index coordinates
    vol = np.transpose(f(ii, jj, kk), (2, 1, 0)).astype(np.float32)
    t = rng.random((500, 3))
    pts = (ranges[0::2] + t * (ranges[1::2] - ranges[0::2])).astype(np.float32)
    got = oracle.sample_points(vol, ranges, pts, fp64=fp64)
    lo, hi = ranges[0::2].astype(np.float64), ranges[1::2].astype(np.float64)
    cont = (pts - lo) / (hi - lo) * (np.array([X, Y, Z]) - 1)
    np.testing.assert_allclose(got, f(*cont.T), rtol=0, atol=4e-6 if fp64 else 2e-5)  # float32 out


@pytest.mark.parametrize("fp64", [False, True])
def test_plane_integrals_of_axis_planes_are_layer_sums(fp64):
    rng = np.random.default_rng(14)
    X, Y, Z = 10, 12, 6
    vol = rng.random((Z, Y, X), dtype=np.float32)
    offset = (2.0, -3.0, 1.0)  # whole voxels: pixels stay on voxel centres
    normals = np.eye(3)
    dims = (X, Y, Z)
    layer_sum = [vol.sum(axis=(0, 1)), vol.sum(axis=(0, 2)), vol.sum(axis=(1, 2))]  # per x, y, z
    for a in range(3):
        dist = (offset[a] + np.arange(dims[a]) - (dims[a] - 1) / 2.0).astype(np.float32)
        got = oracle.plane_integrals(vol, (1, 1, 1), offset, normals[a:a + 1], dist, 32, 32, 1.0, fp64=fp64)
        np.testing.assert_allclose(got[:, 0], layer_sum[a], rtol=2e-6)
    # (the pixel-area factor for other resolutions is pinned by the mass test below)


@pytest.mark.parametrize("fp64", [False, True])
def test_plane_integrals_mass_and_centre(fp64):
    dims, voxel, offset = (24, 20, 22), (1.0, 1.25, 0.8), (1.5, -2.0, 0.7)
    X, Y, Z = dims
    axes = [offset[k] + (np.arange(n) - (n - 1) / 2.0) * voxel[k] for k, n in enumerate(dims)]
    zz, yy, xx = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    c0 = np.array([2.5, -3.0, 1.2])
    vol = np.exp(-((xx - c0[0]) ** 2 + (yy - c0[1]) ** 2 + (zz - c0[2]) ** 2) / (2 * 3.0 ** 2))
    vol = vol.astype(np.float32)
    mass = vol.astype(np.float64).sum() * np.prod(voxel)
    centre = np.array([(vol * a).sum() for a in (xx, yy, zz)]) / vol.sum()
    normals = []
    for pol, azi in [(0.4, 0.3), (1.1, 2.0), (2.5, -1.0), (1.5708, 0.7)]:
        normals.append([np.sin(pol) * np.cos(azi), np.sin(pol) * np.sin(azi), np.cos(pol)])
    normals = np.array(normals)
    step = 0.5
    dist = np.arange(-25.0, 25.0 + step / 2, step).astype(np.float32)
    side = 48  # ceil(1.4142 * 24) = 34 -> next multiple of 16
    reso = min(voxel)
    R = oracle.plane_integrals(vol, voxel, offset, normals, dist, side, side, reso, fp64=fp64)
    for o, n in enumerate(normals):
        m = R[:, o].astype(np.float64).sum() * step
        assert m == pytest.approx(mass, rel=3e-3)
        c = (R[:, o] * dist).astype(np.float64).sum() * step / m
        assert c == pytest.approx(float(n @ centre), abs=2e-2)


def test_ref32_matches_ref64():
    rng = np.random.default_rng(15)
    vol = rng.random((9, 11, 13), dtype=np.float32)
    voxel, offset = (0.9, 1.1, 1.3), (0.4, -0.2, 0.3)
    n = np.array([[0.2, 0.4, 0.894427191], [-0.6, 0.0, 0.8]])
    dist = np.array([-3.0, -0.5, 0.0, 2.25], np.float32)
    a = oracle.plane_integrals(vol, voxel, offset, n, dist, 32, 32, 0.9)
    b = oracle.plane_integrals(vol, voxel, offset, n, dist, 32, 32, 0.9, fp64=True)
    np.testing.assert_allclose(a, b, rtol=2e-5, atol=1e-5 * np.abs(b).max())
    s = oracle.slice_plane(vol, voxel, offset, n[0], 0.7, 19, 23, 0.8)
    t = oracle.slice_plane(vol, voxel, offset, n[0], 0.7, 19, 23, 0.8, fp64=True)
    np.testing.assert_allclose(s, t, atol=1e-5)
