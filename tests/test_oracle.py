"""Pins the CPU oracle (oracle/ctl_oracle.c) against what the mathematics fixes.

The reference holds no vectors for the projector's output (its projector tests check Poisson
statistics and loose spectral fixtures only), so every oracle function is checked here against
closed forms, invariants, an independent interpolation routine (scipy.ndimage) and brute force.
Each check is chosen so that a plausible slip (a swapped tap, a wrong sign, a dropped half-voxel
margin, a transposed QR) fails it.
"""
import math

import numpy as np
import pytest
from scipy import ndimage

import oracle

DEG = math.pi / 180.0


def rx(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[1, 0, 0], [0, c, -s], [0, s, c]], float)


def compose(K, R, c):
    M = K @ R
    return np.hstack([M, (-M @ c)[:, None]])


# projectionmatrixtest.cpp:8-14, 31: P = -compose(K, Rx(-90 deg), (0, 300, 0)) / 300
K0 = np.array([[1000.0, 0, 300.0], [0, 1000.0, 250.0], [0, 0, 1.0]])
P_TEST = -compose(K0, rx(-90 * DEG), np.array([0.0, 300.0, 0.0])) / 300.0


# ---------------------------------------------------------------- FP64 geometry helpers

def test_source_position_closed_form():
    # projectionmatrixtest.cpp:49-52
    assert np.allclose(oracle.source_position(P_TEST), [0.0, 300.0, 0.0], atol=1e-12)


def test_source_position_is_null_space_of_random_P():
    rng = np.random.default_rng(1)
    for _ in range(20):
        P = rng.normal(size=(3, 4))
        c = oracle.source_position(P)
        assert np.allclose(P @ np.append(c, 1.0), 0.0, atol=1e-10)


def test_qr_factorises_and_is_orthogonal():
    rng = np.random.default_rng(2)
    for _ in range(50):
        A = rng.normal(size=(3, 3)) * rng.uniform(0.1, 1000)
        Q, R = oracle.qr(A)
        assert np.allclose(Q @ R, A, rtol=1e-12, atol=1e-12 * np.abs(A).max())
        assert np.allclose(Q.T @ Q, np.eye(3), atol=1e-13)
        assert np.allclose(np.tril(R, -1), 0.0)
        assert np.linalg.det(Q) == pytest.approx(1.0, abs=1e-12)  # two Householder reflections
        # same factorisation as LAPACK up to the sign of each column/row
        Qn, Rn = np.linalg.qr(A)
        s = np.sign(np.diag(R)) * np.sign(np.diag(Rn))
        assert np.allclose(Q, Qn * s, atol=1e-12)


def test_direction_sign_rule_and_principal_ray():
    """d = sign(det M) M^-1 [x,y,1] points from the source towards the detector; at the
    principal point it is the principal ray (projectionmatrixtest.cpp:34-47)."""
    for P in (P_TEST, -P_TEST, 7.0 * P_TEST):
        q = oracle.decompose_m(P)
        d = oracle.direction64(q, 300.0, 250.0)
        M = P[:, :3]
        principal = np.sign(np.linalg.det(M)) * M[2] / np.linalg.norm(M[2])
        assert np.allclose(d / np.linalg.norm(d), principal, atol=1e-12)
        # a point on the ray in front of the source projects onto the pixel with positive depth
        c = oracle.source_position(P)
        for x, y in [(0.0, 0.0), (511.0, 13.0), (300.0, 250.0)]:
            dd = oracle.direction64(q, x, y)
            h = P @ np.append(c + 10.0 * dd / np.linalg.norm(dd), 1.0)
            assert np.allclose(h[:2] / h[2], [x, y], atol=1e-9)
            assert h[2] * np.sign(np.linalg.det(M)) > 0


def test_decompose_m_packs_q_and_r():
    rng = np.random.default_rng(3)
    P = rng.normal(size=(3, 4))
    q = oracle.decompose_m(P)
    Q = q[:9].reshape(3, 3)
    R = np.array([[q[9], q[10], q[11]], [0, q[12], q[13]], [0, 0, q[14]]])
    M = P[:, :3]
    assert q[15] == 0.0
    assert np.prod(np.diag(R)) > 0  # R negated when its determinant is negative
    assert np.allclose(Q @ R, np.sign(np.linalg.det(M)) * M, atol=1e-12)


# ---------------------------------------------------------------- interpolation

def scipy_sample(vol, p):
    # texel centres of the OpenCL image at k + 1/2  <->  scipy's centres at integer k
    return ndimage.map_coordinates(vol.astype(np.float64), [[p[2] - 0.5], [p[1] - 0.5], [p[0] - 0.5]],
                                   order=1, mode="constant", cval=0.0, prefilter=False)[0]


def test_linear_filter_matches_independent_trilinear():
    rng = np.random.default_rng(4)
    vol = rng.uniform(-1, 1, size=(5, 6, 7)).astype(np.float32)  # (Z, Y, X)
    for _ in range(2000):
        p = rng.uniform(-1.2, [8.2, 7.2, 6.2])
        ref = scipy_sample(vol, p) if np.all(p >= 0.5) and np.all(p <= np.array([7, 6, 5]) - 0.5) else None
        v32 = oracle.sample_linear32(vol, *map(float, p))
        v64 = oracle.sample_linear64(vol, *map(float, p))
        assert v32 == pytest.approx(v64, abs=1e-6)
        if ref is not None:
            assert v64 == pytest.approx(ref, abs=1e-12)


def test_linear_filter_border_is_zero_and_half_voxel_support():
    vol = np.ones((4, 4, 4), np.float32)
    # inside the texel-centre hull the value is exactly 1
    assert oracle.sample_linear64(vol, 0.5, 0.5, 0.5) == 1.0
    assert oracle.sample_linear64(vol, 3.5, 2.0, 1.7) == 1.0
    # half a voxel outside the grid the border texel (0) takes full weight
    assert oracle.sample_linear64(vol, -0.5, 2.0, 2.0) == 0.0
    assert oracle.sample_linear64(vol, 4.5, 2.0, 2.0) == 0.0
    # the tent: a quarter voxel outside the outer centre keeps 3/4 of the edge value
    assert oracle.sample_linear64(vol, 0.25, 2.0, 2.0) == pytest.approx(0.75)
    assert oracle.sample_linear64(vol, 3.75, 2.0, 2.0) == pytest.approx(0.75)
    # far away: 0
    assert oracle.sample_linear32(vol, 1e6, 2.0, 2.0) == 0.0


def test_linear_filter_reproduces_affine_field_inside():
    z, y, x = np.meshgrid(np.arange(6), np.arange(5), np.arange(4), indexing="ij")
    vol = (0.5 + 2.0 * (x + 0.5) - 1.0 * (y + 0.5) + 0.25 * (z + 0.5)).astype(np.float32)
    rng = np.random.default_rng(5)
    for _ in range(300):
        p = rng.uniform([0.5, 0.5, 0.5], [3.5, 4.5, 5.5])
        want = 0.5 + 2.0 * p[0] - 1.0 * p[1] + 0.25 * p[2]
        assert oracle.sample_linear64(vol, *p) == pytest.approx(want, abs=1e-12)


# ---------------------------------------------------------------- slab clip

def brute_force_interval(s, d, lo, hi):
    lam = np.linspace(0, 400, 400001)
    pts = s[None] + lam[:, None] * d[None]
    inside = np.all((pts >= lo) & (pts <= hi), axis=1)
    if not inside.any():
        return None
    return lam[inside][0], lam[inside][-1]


def test_clip_matches_brute_force():
    rng = np.random.default_rng(6)
    size = np.array([10.0, 12.0, 8.0], np.float32)
    corner = np.array([-5.0, -6.0, -4.0], np.float32)
    hits = 0
    for _ in range(200):
        s = rng.uniform(-60, 60, 3).astype(np.float32)
        tgt = rng.uniform(-8, 8, 3)
        d = (tgt - s)
        d = (d / np.linalg.norm(d) * 0.25).astype(np.float32)
        mm = oracle.clip32(s, d, size, corner, True)
        bf = brute_force_interval(s.astype(float), d.astype(float), corner - 0.5, corner + size + 0.5)
        if bf is None:
            assert mm[0] == np.float32(np.finfo(np.float32).max)
        else:
            hits += 1
            assert mm[0] == pytest.approx(max(bf[0], 0.0), abs=2e-3)
            assert mm[1] == pytest.approx(bf[1], abs=2e-3)
    assert hits > 50


def test_clip_axis_parallel_and_nearest_box():
    size = np.array([4.0, 4.0, 4.0], np.float32)
    corner = np.array([0.0, 0.0, 0.0], np.float32)
    s = np.array([-10.0, 2.0, 2.0], np.float32)
    d = np.array([0.5, 0.0, 0.0], np.float32)
    assert np.allclose(oracle.clip32(s, d, size, corner, True), [(10 - 0.5) / 0.5, (14 + 0.5) / 0.5])
    assert np.allclose(oracle.clip32(s, d, size, corner, False), [10 / 0.5, 14 / 0.5])
    # parallel to a face, outside it: miss
    s2 = np.array([-10.0, 5.0, 2.0], np.float32)
    assert oracle.clip32(s2, d, size, corner, True)[0] == np.finfo(np.float32).max


# ---------------------------------------------------------------- projections

def axis_geometry(n_pix=9, pitch=1.0, sdd=1000.0, sid=500.0):
    """A detector looking down +x from a source at (-sid, 0, 0): pixel (x, y) maps to the
    direction (1, (x - cx)*p/sdd, (y - cy)*p/sdd) for a perfect pinhole."""
    cx = cy = 0.5 * (n_pix - 1)
    f = sdd / pitch
    K = np.array([[f, 0, cx], [0, f, cy], [0, 0, 1.0]])
    R = np.array([[0, 1.0, 0], [0, 0, 1.0], [1.0, 0, 0]])  # world x -> depth
    return compose(K, R, np.array([-sid, 0.0, 0.0]))[None, None]


def test_uniform_cube_central_ray_closed_form():
    """Central ray through a uniform cube: integral = mu * (N + 1) voxels... the half-voxel tent
    at each face integrates to mu/2 * 1/2 on both sides, so the line integral is mu * N."""
    n, mu, vs = 16, 0.02, 2.0
    vol = np.full((n, n, n), mu, np.float32)
    P = axis_geometry()
    for fp64 in (False, True):
        out, ns = oracle.project(vol, (vs,) * 3, (0, 0, 0), P, 9, 9, ray_sampling=0.05, fp64=fp64)
        centre = out[0, 0, 4, 4]
        assert centre == pytest.approx(mu * n * vs, rel=2e-3)


def test_ball_chord_lengths():
    """Rays through a ball: the integral is mu * chord length (sampled, tol. set by voxels)."""
    n, vs, r_mm, mu = 48, 1.0, 18.0, 0.03
    g = (np.arange(n) + 0.5) - n / 2
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    vol = np.where(x * x + y * y + z * z <= r_mm * r_mm, mu, 0.0).astype(np.float32)
    P = axis_geometry(n_pix=21, pitch=2.0, sdd=1000.0, sid=500.0)
    out, _ = oracle.project(vol, (vs,) * 3, (0, 0, 0), P, 21, 21, ray_sampling=0.1, fp64=True)
    for px in (10, 13, 16):
        # impact parameter at the ball centre: (px - 10) * 2 mm * 500/1000
        b = (px - 10) * 2.0 * 0.5
        chord = 2.0 * math.sqrt(r_mm * r_mm - b * b)
        assert out[0, 0, 10, px] == pytest.approx(mu * chord, rel=0.04)


def test_linearity_zero_and_ref32_close_to_ref64():
    rng = np.random.default_rng(7)
    v1 = rng.uniform(0, 0.05, (12, 10, 14)).astype(np.float32)
    v2 = rng.uniform(0, 0.05, (12, 10, 14)).astype(np.float32)
    P = axis_geometry(n_pix=15, pitch=3.0)
    kw = dict(ray_sampling=0.3)
    a, _ = oracle.project(v1, (1.0, 1.2, 0.8), (0.5, -1, 2), P, 15, 15, fp64=True, **kw)
    b, _ = oracle.project(v2, (1.0, 1.2, 0.8), (0.5, -1, 2), P, 15, 15, fp64=True, **kw)
    ab, _ = oracle.project(2 * v1 + 3 * v2, (1.0, 1.2, 0.8), (0.5, -1, 2), P, 15, 15, fp64=True, **kw)
    assert np.allclose(ab, 2 * a + 3 * b, rtol=1e-5, atol=1e-7)
    z, _ = oracle.project(np.zeros_like(v1), (1.0, 1.2, 0.8), (0, 0, 0), P, 15, 15, **kw)
    assert np.all(z == 0.0)
    a32, _ = oracle.project(v1, (1.0, 1.2, 0.8), (0.5, -1, 2), P, 15, 15, **kw)
    assert oracle.rel_l2(a, a32) < 1e-5


def test_sample_count_closed_form():
    """Central ray along x through N voxels with step s: lambda in/out at (sid - N/2 - 1/2)/s and
    (sid + N/2 + 1/2)/s, count = floor(out) + 2 - floor(in)."""
    n = 10
    vol = np.ones((n, n, n), np.float32)
    P = axis_geometry(n_pix=1, sdd=1000.0, sid=500.0)
    _, ns = oracle.project(vol, (1.0,) * 3, (0, 0, 0), P, 1, 1, ray_sampling=0.25, fp64=True)
    lin, lout = (500 - 5 - 0.5) / 0.25, (500 + 5 + 0.5) / 0.25
    assert ns == math.floor(lout) + 2 - math.floor(lin)


def test_translation_of_volume_moves_projection():
    """Shifting the volume offset along the detector's row axis by k detector-pixel footprints
    shifts a parallel-like projection by k pixels (far source, small object)."""
    rng = np.random.default_rng(8)
    vol = np.zeros((8, 8, 8), np.float32)
    vol[2:6, 3:5, 2:6] = rng.uniform(0.01, 0.02, (4, 2, 4))
    P = axis_geometry(n_pix=31, pitch=1.0, sdd=1e6, sid=5e5)  # magnification 2
    a, _ = oracle.project(vol, (1.0,) * 3, (0, 0, 0), P, 31, 31, ray_sampling=0.1, fp64=True)
    b, _ = oracle.project(vol, (1.0,) * 3, (0, 2.0, 0), P, 31, 31, ray_sampling=0.1, fp64=True)
    # world +y maps to detector +x; 2 mm at magnification 2 = 4 pixels
    assert np.allclose(b[0, 0, :, 4:], a[0, 0, :, :-4], atol=1e-5 * a.max())


def test_upsampling_factor_semantics():
    rng = np.random.default_rng(9)
    vol = rng.uniform(0, 0.05, (6, 5, 4)).astype(np.float32)
    P = axis_geometry(n_pix=7, pitch=3.0)
    # factor 2 equals a manually replicated volume with half the voxel size
    rep = np.repeat(np.repeat(np.repeat(vol, 2, 0), 2, 1), 2, 2)
    a, _ = oracle.project(vol, (2.0,) * 3, (0, 0, 0), P, 7, 7, ray_sampling=0.3, upsampling=2)
    b, _ = oracle.project(rep, (1.0,) * 3, (0, 0, 0), P, 7, 7, ray_sampling=0.3)
    assert np.array_equal(a, b)
    # factor 0 equals a single voxel holding the mean
    c, _ = oracle.project(vol, (2.0,) * 3, (0, 0, 0), P, 7, 7, ray_sampling=0.3, upsampling=0)
    one = np.full((1, 1, 1), vol.astype(np.float32).sum(dtype=np.float32) / np.float32(vol.size), np.float32)
    d, _ = oracle.project(one, (8.0, 10.0, 12.0), (0, 0, 0), P, 7, 7, ray_sampling=0.3)
    assert np.allclose(c, d, rtol=1e-6)


def test_nearest_mode_cube_and_rays_per_pixel_average():
    n, mu = 12, 0.02
    vol = np.full((n, n, n), mu, np.float32)
    P = axis_geometry(n_pix=5, pitch=1.0, sdd=1000.0, sid=500.0)
    out, _ = oracle.project(vol, (1.0,) * 3, (0, 0, 0), P, 5, 5, ray_sampling=0.05,
                            interpolate=False, fp64=True)
    assert out[0, 0, 2, 2] == pytest.approx(mu * n, rel=1e-2)
    # rays per pixel: the pixel value is the mean of its sub-rays
    rng = np.random.default_rng(10)
    v = rng.uniform(0, 0.05, (8, 8, 8)).astype(np.float32)
    P2 = axis_geometry(n_pix=3, pitch=4.0, sdd=1000.0, sid=500.0)
    sub, _ = oracle.project(v, (1.0,) * 3, (0, 0, 0), axis_geometry(n_pix=6, pitch=2.0), 6, 6,
                            ray_sampling=0.2, fp64=True)
    avg, _ = oracle.project(v, (1.0,) * 3, (0, 0, 0), P2, 3, 3, ray_sampling=0.2,
                            rays_per_pixel=(2, 2), fp64=True)
    want = sub[0, 0].reshape(3, 2, 3, 2).mean(axis=(1, 3))
    assert np.allclose(avg[0, 0], want, rtol=1e-6)
