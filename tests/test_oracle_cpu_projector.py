"""Pins refcpu, the oracle's restatement of the reference's CPU projector RayCasterProjectorCPU
(modules/src/projectors/raycasterprojectorcpu.cpp:132-407), on the CPU.

refcpu differs from ref64 only where the reference's CPU projector does: FP32 interpolation
weights and values, and its reads at the high border. So it is checked against ref64 (itself
pinned by closed forms in test_oracle.py) at a tolerance that an FP32 weight allows but a wrong
step, axis scale, corner or loop bound does not; its single-sample reads against an independent
FP64 interpolation, the affine fields trilinear interpolation reproduces, and hand-worked values
at the border, including the unmasked read at a position exactly n + 1/2 (raycasterprojectorcpu.cpp
:247-272) and the nearest read at exactly n (:287-299).
"""
import numpy as np
import pytest

import oracle
from ctl_amd import workloads as W


def test_matches_ref64_on_the_small_workload():
    w = W.small(views=36)
    pm = w.pmats[::4]
    r64, n64 = oracle.project(w.volume, w.voxel_mm, w.offset_mm, pm, w.channels, w.rows,
                              w.ray_sampling, fp64=True)
    rc, nc = oracle.project(w.volume, w.voxel_mm, w.offset_mm, pm, w.channels, w.rows,
                            w.ray_sampling, cpu_projector=True)
    assert nc == n64  # the same loop bounds
    e, m = oracle.rel_l2(r64, rc), oracle.max_abs_rel(r64, rc)
    assert e <= 2e-7 and m <= 2e-6, (e, m)
    # and it is not ref32, whose FP32 positions move grazing pixels by far more
    r32, _ = oracle.project(w.volume, w.voxel_mm, w.offset_mm, pm, w.channels, w.rows, w.ray_sampling)
    assert oracle.max_abs_rel(r32, rc) > 10 * m


def test_matches_ref64_anisotropic_offset_volume_and_rays_per_pixel():
    """Anisotropic voxels make the per-axis increment (increment_mm / voxel size) and corner
    scaling visible; an offset volume moves the corner; 2 x 3 rays per pixel the sub-ray grid.
    The CPU projector divides the FP32 increment by the FP32 voxel size in FP32 (:164-166), so
    with voxel sizes that are no powers of two its step differs from ref64's by an FP32 rounding
    (relative L2 ~1e-6 here); in nearest mode such a step moves a few samples across a voxel
    boundary, where the value jumps by the phantom's contrast (~1e-4). A wrong axis scale or corner
    moves either by 1e-2 or more."""
    dims, vs, off = (23, 17, 29), (1.3, 0.9, 1.7), (6.0, -4.5, 9.0)
    vol = np.ascontiguousarray(W.ellipsoid_phantom(dims, vs, 7, 77, semi_axes=(3.0, 12.0)))
    pm = W.carm_circle(5, (61, 47), (1.1, 0.9))
    for rpp, interp in (((1, 1), True), ((2, 3), True), ((1, 1), False)):
        r64, _ = oracle.project(vol, vs, off, pm, 61, 47, 0.15, rays_per_pixel=rpp, fp64=True,
                                interpolate=interp)
        rc, _ = oracle.project(vol, vs, off, pm, 61, 47, 0.15, rays_per_pixel=rpp, interpolate=interp,
                               cpu_projector=True)
        e = oracle.rel_l2(r64, rc)
        assert e <= (5e-6 if interp else 5e-4), (rpp, interp, e)


def test_interpolated_read_against_fp64_filter_and_affine_fields():
    rng = np.random.default_rng(3)
    vol = rng.uniform(0.0, 1.0, size=(6, 5, 7)).astype(np.float32)  # (Z, Y, X)
    for _ in range(300):
        p = rng.uniform(-0.4, [7.4, 5.4, 6.4])
        got = oracle.cpu_read(vol, *p)
        ref = oracle.sample_linear64(vol, *p)
        assert abs(got - ref) <= 2e-6, (p, got, ref)
    # an affine field is reproduced wherever all eight taps are inside the grid
    z, y, x = np.meshgrid(np.arange(6), np.arange(5), np.arange(7), indexing="ij")
    aff = (0.25 + 0.5 * x - 0.125 * y + 0.375 * z).astype(np.float32)
    for _ in range(100):
        p = rng.uniform(0.5, [6.5, 4.5, 5.5])
        want = 0.25 + 0.5 * (p[0] - 0.5) - 0.125 * (p[1] - 0.5) + 0.375 * (p[2] - 0.5)
        assert oracle.cpu_read(aff, *p) == pytest.approx(want, abs=4e-6)


def test_border_reads_and_the_unmasked_read_at_n_plus_half():
    """A 3 x 2 x 2 volume v[k] = k + 1 (x fastest). Just inside the high x border only the x = 2
    voxels contribute, with weight ~0; at exactly x = 3 + 1/2 the reference's floor(p - 1/2) = 3
    is not masked for the x taps of that corner, whose reads land on linear indices 3, 6, 9 and 12
    (the last one past the array, read as 0 here), each with weight 1/4 at y = z = 1."""
    v = (np.arange(12, dtype=np.float32) + 1.0).reshape(2, 2, 3)  # (Z, Y, X)
    flat = v.reshape(-1)
    inside = oracle.cpu_read(v, 3.5 - 1e-9, 1.0, 1.0)
    assert abs(inside) < 1e-7
    at = oracle.cpu_read(v, 3.5, 1.0, 1.0)
    assert at == pytest.approx((flat[3] + flat[6] + flat[9] + 0.0) / 4.0, rel=1e-6)
    assert oracle.cpu_read(v, 3.5 + 1e-9, 1.0, 1.0) == 0.0  # beyond n + 1/2: rejected
    # low border: half a voxel out the tent reaches 0, the centre of voxel 0 reads it
    assert oracle.cpu_read(v, -0.5, 1.0, 1.0) == 0.0
    assert oracle.cpu_read(v, 0.5, 0.5, 0.5) == pytest.approx(flat[0])
    assert oracle.cpu_read(v, 0.0, 0.5, 0.5) == pytest.approx(0.5 * flat[0])


def test_nearest_read_edges():
    v = (np.arange(12, dtype=np.float32) + 1.0).reshape(2, 2, 3)
    flat = v.reshape(-1)
    assert oracle.cpu_read(v, 2.999, 0.5, 0.5, interpolate=False) == flat[2]
    # p = n passes the range test (p > n is rejected) and floor(3) = 3 reads linear index 3
    assert oracle.cpu_read(v, 3.0, 0.5, 0.5, interpolate=False) == flat[3]
    assert oracle.cpu_read(v, 3.0001, 0.5, 0.5, interpolate=False) == 0.0
    assert oracle.cpu_read(v, -1e-9, 0.5, 0.5, interpolate=False) == 0.0
    assert oracle.cpu_read(v, 1.5, 1.5, 1.5, interpolate=False) == v[1, 1, 1]
