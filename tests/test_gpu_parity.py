"""Parity of the HIP projector (through the C ABI) with the CPU oracle.

Bars (fixed here, not tuned per run):
  FP32 storage vs ref32:  relative L2 <= 1e-4 and max over pixels of |d| / max(|ref|, 1e-3*max|ref|)
                          <= 1e-3 (errormetrics.cpp:69-85 for the L2 definition, reference first);
  FP16 storage vs ref64:  relative L2 <= 5e-4;
  BF16 storage vs ref64:  relative L2 <= 4e-3 (8-bit mantissa; reported, not a drop-in claim).
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from ctl_amd import (RayCasterProjector, ResidentVolume, VoxelVolume, CtlProjError)
from ctl_amd import _lib
from ctl_amd import workloads as W

pytestmark = pytest.mark.gpu

L2_F32, MAR_F32, L2_F16, L2_BF16 = 1e-4, 1e-3, 5e-4, 4e-3


def run_gpu(w, storage="f32", pm=None, **settings):
    pm = w.pmats if pm is None else pm
    p = RayCasterProjector()
    s = p.settings()
    s.raySampling = w.ray_sampling
    for k, v in settings.items():
        setattr(s, k, v)
    p.configure((pm, w.channels, w.rows))
    vol = ResidentVolume(VoxelVolume(w.volume, w.voxel_mm, w.offset_mm), storage)
    out = p.project(vol).data
    return out, p, vol


def run_ref(w, pm=None, fp64=False, **kw):
    pm = w.pmats if pm is None else pm
    return oracle.project(w.volume, w.voxel_mm, w.offset_mm, pm, w.channels, w.rows,
                          w.ray_sampling, fp64=fp64, **kw)


@pytest.fixture(scope="module")
def small():
    return W.small()


@pytest.fixture(scope="module")
def small_refs(small):
    r32, n32 = run_ref(small)
    r64, n64 = run_ref(small, fp64=True)
    return r32, n32, r64


def test_small_fp32_matches_ref32(small, small_refs):
    r32, _, r64 = small_refs
    got, _, _ = run_gpu(small)
    assert got.shape == (36, 1, 128, 256)
    e, m = oracle.rel_l2(r32, got), oracle.max_abs_rel(r32, got)
    print(f"small f32: rel L2 vs ref32 {e:.3e}, max abs rel {m:.3e}, vs ref64 {oracle.rel_l2(r64, got):.3e}")
    assert e <= L2_F32 and m <= MAR_F32


@pytest.mark.parametrize("storage,bar", [("f16", L2_F16), ("bf16", L2_BF16)])
def test_small_half_storage_vs_ref64(small, small_refs, storage, bar):
    _, _, r64 = small_refs
    got, _, _ = run_gpu(small, storage)
    e = oracle.rel_l2(r64, got)
    print(f"small {storage}: rel L2 vs ref64 {e:.3e}")
    assert e <= bar


def test_sample_count_matches_oracle(small, small_refs):
    _, n32, _ = small_refs
    p = RayCasterProjector()
    p.settings().raySampling = small.ray_sampling
    p.configure((small.pmats, small.channels, small.rows))
    vol = ResidentVolume(VoxelVolume(small.volume, small.voxel_mm, small.offset_mm))
    n = p.count_samples(vol)
    print(f"samples: gpu {n} oracle {n32}")
    assert abs(n - n32) <= 1e-5 * n32


TILES = [(8, 8), (16, 4), (4, 16), (32, 2), (2, 32), (64, 1), (1, 64)]


def test_unroll_and_tile_variants_agree(small, small_refs):
    """Every wavefront tile shape (each maps its lanes to pixels its own way: 2 x 2 quads, or rows
    for tiles one pixel wide or tall), staged and unstaged, against ref32."""
    r32 = small_refs[0]
    for tile in TILES:
        for unroll in (1, 2, 3, 4):
            got, _, _ = run_gpu(small, pm=small.pmats[::6], tile=tile, unroll=unroll)
            e, m = oracle.rel_l2(r32[::6], got), oracle.max_abs_rel(r32[::6], got)
            assert e <= L2_F32 and m <= MAR_F32, (tile, unroll, e, m)
        got, _, _ = run_gpu(small, pm=small.pmats[::6], tile=tile, unroll=4, lds_chunk=-1)
        e, m = oracle.rel_l2(r32[::6], got), oracle.max_abs_rel(r32[::6], got)
        assert e <= L2_F32 and m <= MAR_F32, (tile, "unstaged", e, m)


@pytest.mark.parametrize("storage,bar", [("f16", L2_F16), ("bf16", L2_BF16)])
def test_tile_variants_half_storage(small, small_refs, storage, bar):
    r64 = small_refs[2][::6]
    for tile in TILES:
        got, _, _ = run_gpu(small, storage, pm=small.pmats[::6], tile=tile, unroll=4)
        e = oracle.rel_l2(r64, got)
        assert e <= bar, (storage, tile, e)


# Staged vs unstaged kernel: the same samples with the same per-sample arithmetic, summed in a
# different order (per-chunk sums vs unrolled partial sums). Summation order moves a pixel by far
# less than SAME_SAMPLES; one extra, missing or wrong tap moves it by about 1/(samples per ray),
# at least 5e-4 at these sizes.
SAME_SAMPLES = 2e-4


def check_same_samples(plain, got, what):
    d, dm = oracle.rel_l2(plain, got), oracle.max_abs_rel(plain, got)
    print(f"{what}: vs unstaged rel L2 {d:.3e}, max abs rel {dm:.3e}")
    assert dm <= SAME_SAMPLES, (what, dm)


@pytest.mark.parametrize("chunk,texels", [(16, 0), (64, 0), (256, 0), (64, 1024), (1024, 20352)])
def test_lds_staged_matches_unstaged_and_ref32(small, small_refs, chunk, texels):
    """FP32 trilinear projections stage the taps of each chunk of samples in LDS; 1024 texels
    forces the HBM fallback for the larger boxes, 20352 is the largest box."""
    r32 = small_refs[0]
    pm = small.pmats[::5]
    plain, _, _ = run_gpu(small, pm=pm, lds_chunk=-1, tile=(8, 8), unroll=4)
    got, _, _ = run_gpu(small, pm=pm, lds_chunk=chunk, lds_texels=texels, tile=(8, 8), unroll=4)
    check_same_samples(plain, got, f"lds chunk {chunk} texels {texels}")
    e, m = oracle.rel_l2(r32[::5], got), oracle.max_abs_rel(r32[::5], got)
    print(f"lds chunk {chunk} texels {texels}: vs ref32 {e:.3e} / {m:.3e}")
    assert e <= L2_F32 and m <= MAR_F32


@pytest.mark.parametrize("storage,bar", [("f16", L2_F16), ("bf16", L2_BF16)])
def test_lds_staged_half_storage(small, small_refs, storage, bar):
    """FP16/BF16 (y,z)-quad texels staged in LDS, with unroll 3 (a third of the shared body through
    L1) and 4 (a quarter): against the unstaged kernel with the same tile and unroll, and against
    ref64."""
    r64 = small_refs[2][::5]
    pm = small.pmats[::5]
    for unroll in (3, 4):
        plain, _, _ = run_gpu(small, storage, pm=pm, lds_chunk=-1, tile=(8, 8), unroll=unroll)
        for chunk, texels in ((64, 0), (16, 1024)):
            got, _, _ = run_gpu(small, storage, pm=pm, lds_chunk=chunk, lds_texels=texels, tile=(8, 8),
                                unroll=unroll)
            check_same_samples(plain, got, f"lds {storage} unroll {unroll} chunk {chunk} texels {texels}")
            e = oracle.rel_l2(r64, got)
            print(f"lds {storage} unroll {unroll} chunk {chunk} texels {texels}: vs ref64 {e:.3e}")
            assert e <= bar


def test_lds_staged_real_size_oblique_views():
    """Full-size geometry: views at 0, 45, 137 and 225 degrees (boxes of axis-aligned and
    diagonal ray bundles), staged against unstaged and ref32."""
    w = W.real(views=360)
    pick = [0, 45, 137, 225]
    r32, _ = run_ref(w, views=pick)
    plain, _, _ = run_gpu(w, pm=w.pmats[pick], lds_chunk=-1, tile=(16, 4), unroll=4)
    got, _, _ = run_gpu(w, pm=w.pmats[pick])  # default: staged, 16x4 tiles, unroll 3
    check_same_samples(plain, got, f"lds real views {pick}")
    e, m = oracle.rel_l2(r32, got), oracle.max_abs_rel(r32, got)
    e0 = oracle.rel_l2(r32, plain)
    print(f"lds real views {pick}: vs ref32 {e:.3e} / {m:.3e} (unstaged {e0:.3e})")
    assert e <= L2_F32 and m <= MAR_F32


def test_nearest_mode(small):
    pm = small.pmats[::4]
    r32, _ = run_ref(small, pm=pm, interpolate=False)
    got, _, _ = run_gpu(small, pm=pm, interpolate=False)
    e = oracle.rel_l2(r32, got)
    print(f"nearest: rel L2 {e:.3e}")
    assert e <= L2_F32


def test_rays_per_pixel(small):
    pm = small.pmats[::9]
    r32, _ = run_ref(small, pm=pm, rays_per_pixel=(2, 3))
    got, _, _ = run_gpu(small, pm=pm, raysPerPixel=[2, 3])
    e, m = oracle.rel_l2(r32, got), oracle.max_abs_rel(r32, got)
    assert e <= L2_F32 and m <= MAR_F32


@pytest.mark.parametrize("factor", [0, 2])
def test_volume_upsampling(small, factor):
    pm = small.pmats[::9]
    r32, _ = run_ref(small, pm=pm, upsampling=factor)
    got, _, _ = run_gpu(small, pm=pm, volumeUpSampling=factor)
    e = oracle.rel_l2(r32, got)
    print(f"upsampling {factor}: rel L2 {e:.3e}")
    assert e <= L2_F32


def test_helical_multimodule():
    pm = W.helical_geometry(n_views=12, turns=0.05)
    vol = np.ascontiguousarray(W.ellipsoid_phantom((128, 128, 64), (2.8, 2.8, 4.0), 12, 3001,
                                                   mu=(0.0, 0.04)))
    w = W.Workload("helical small", vol, (2.8, 2.8, 4.0), (0.0, 0.0, 0.0), pm, 16, 64, 0.25)
    r32, _ = run_ref(w)
    got, _, _ = run_gpu(w)
    assert got.shape == (12, 40, 64, 16)
    e, m = oracle.rel_l2(r32, got), oracle.max_abs_rel(r32, got)
    print(f"helical 40 modules: rel L2 {e:.3e}, max abs rel {m:.3e}")
    assert e <= L2_F32 and m <= MAR_F32


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
def test_odd_sizes_anisotropic_voxels_offset_volume(storage):
    """Odd volume sizes (odd padded pitches), anisotropic voxels (per-axis step scaling), a volume
    offset from the isocentre and a detector whose sides are no multiple of any wavefront tile
    (partial tiles on both edges): staged and unstaged FP32 against ref32, half storage against
    ref64."""
    dims, vs, off = (37, 29, 41), (1.3, 0.9, 1.7), (6.0, -4.5, 9.0)
    vol = np.ascontiguousarray(W.ellipsoid_phantom(dims, vs, 9, 3101, semi_axes=(3.0, 14.0)))
    pm = W.carm_circle(9, (77, 53), (1.1, 0.9))
    w = W.Workload("odd", vol, vs, off, pm, 77, 53, 0.15)
    if storage == "f32":
        r32, _ = run_ref(w)
        for kw in ({}, {"lds_chunk": -1}, {"tile": (16, 4), "lds_chunk": 16, "lds_texels": 1024}):
            got, _, _ = run_gpu(w, **kw)
            e, m = oracle.rel_l2(r32, got), oracle.max_abs_rel(r32, got)
            print(f"odd sizes f32 {kw}: rel L2 {e:.3e}, max abs rel {m:.3e}")
            assert e <= L2_F32 and m <= MAR_F32, (kw, e, m)
    else:
        r64, _ = run_ref(w, fp64=True)
        got, _, _ = run_gpu(w, storage)
        e = oracle.rel_l2(r64, got)
        print(f"odd sizes {storage}: rel L2 vs ref64 {e:.3e}")
        assert e <= (L2_F16 if storage == "f16" else L2_BF16)


def test_helical_autotuned_shape():
    """16-channel modules: the autotuner also times the candidates with their wavefronts stacked
    (one column of 1 x 4 wavefronts per module); whichever it keeps, the projection stays within
    the FP32 bar and the recorded shape is a valid wavefront layout."""
    import json
    from ctl_amd import autotune_table
    pm = W.helical_geometry(n_views=12, turns=0.05)
    vol = np.ascontiguousarray(W.ellipsoid_phantom((128, 128, 64), (2.8, 2.8, 4.0), 12, 3002,
                                                   mu=(0.0, 0.04)))
    w = W.Workload("helical small", vol, (2.8, 2.8, 4.0), (0.0, 0.0, 0.0), pm, 16, 64, 0.25)
    r32, _ = run_ref(w)
    p = RayCasterProjector()
    p.settings().raySampling = w.ray_sampling
    p.configure((pm, w.channels, w.rows))
    rv = ResidentVolume(VoxelVolume(w.volume, w.voxel_mm, w.offset_mm))
    assert p.autotune(rv) > 0.0
    table = json.loads(autotune_table())
    assert table and all(v["waves_x"] * v["waves_y"] == 4 for v in table.values())
    got = p.project(rv).data
    e, m = oracle.rel_l2(r32, got), oracle.max_abs_rel(r32, got)
    print(f"helical autotuned {table}: rel L2 {e:.3e}, max abs rel {m:.3e}")
    assert e <= L2_F32 and m <= MAR_F32


def test_real_size_views_against_oracle():
    w = W.real(views=360)
    pick = [0, 137]
    r32, _ = run_ref(w, views=pick)
    got, _, _ = run_gpu(w, pm=w.pmats[pick])
    e, m = oracle.rel_l2(r32, got), oracle.max_abs_rel(r32, got)
    print(f"256^3 1024x768 views {pick}: rel L2 {e:.3e}, max abs rel {m:.3e}")
    assert e <= L2_F32 and m <= MAR_F32


def test_real_size_linearity_and_device_path():
    """At the full 360-view size: project(a*V1 + b*V2) = a*P1 + b*P2 and the device-output path
    equals the host-output path bit for bit."""
    import ctypes
    w = W.real(views=360)
    v2 = np.ascontiguousarray(w.volume[::-1, :, ::-1])
    p = RayCasterProjector()
    p.settings().raySampling = w.ray_sampling
    p.configure((w.pmats, w.channels, w.rows))
    P1 = p.project(ResidentVolume(VoxelVolume(w.volume, w.voxel_mm))).data
    P2 = p.project(ResidentVolume(VoxelVolume(v2, w.voxel_mm))).data
    P12 = p.project(ResidentVolume(VoxelVolume(2.0 * w.volume + 0.5 * v2, w.voxel_mm))).data
    e = oracle.rel_l2(2.0 * P1.astype(np.float64) + 0.5 * P2, P12)
    print(f"linearity rel L2 {e:.3e}")
    assert e <= 1e-5
    import torch
    vol = ResidentVolume(VoxelVolume(w.volume, w.voxel_mm))
    dev = torch.empty(P1.size, dtype=torch.float32, device="cuda")
    p.project_device(vol, dev.data_ptr(), dev.numel())
    p.synchronize()
    assert np.array_equal(dev.cpu().numpy().reshape(P1.shape), P1)


def test_zero_volume_and_miss():
    w = W.small(views=4)
    zero = W.Workload("zero", np.zeros_like(w.volume), w.voxel_mm, w.offset_mm, w.pmats, 256, 128, 0.1)
    got, _, _ = run_gpu(zero)
    assert np.all(got == 0.0)
    # volume shifted far outside every ray: all rays miss
    far = W.Workload("far", w.volume, w.voxel_mm, (0.0, 0.0, 5000.0), w.pmats, 256, 128, 0.1)
    got, _, _ = run_gpu(far)
    assert np.all(got == 0.0)


def test_error_paths():
    w = W.small(views=4)
    p = RayCasterProjector()
    with pytest.raises(CtlProjError) as ei:  # project before configure
        p.project(ResidentVolume(VoxelVolume(w.volume, w.voxel_mm)))
    assert ei.value.status == _lib.ERR_NOT_CONFIGURED
    p.configure((w.pmats, w.channels, w.rows))
    # a volume with no data gives zero views, as the reference does (raycasterprojector.cpp:156-160)
    empty = p.project(VoxelVolume(np.zeros((0, 0, 0), np.float32)))
    assert empty.nbViews() == 0
    with pytest.raises(CtlProjError) as ei:
        _lib.check(_lib.lib().ctlproj_context_create(99, ctypes_ref()))
    assert ei.value.status == _lib.ERR_NO_DEVICE
    # forced device failure: an error status, no partial result reported as success
    _lib.check(_lib.lib().ctlproj_debug_inject_failure(p.ctx.h, 1))
    with pytest.raises(CtlProjError) as ei:
        p.project(ResidentVolume(VoxelVolume(w.volume, w.voxel_mm)))
    assert ei.value.status == _lib.ERR_INJECTED
    out = p.project(ResidentVolume(VoxelVolume(w.volume, w.voxel_mm))).data
    assert np.isfinite(out).all() and out.max() > 0
    # an LDS box beyond one workgroup's LDS (with the kernels' static reserve) is refused on the
    # host, before any launch
    for bad in (dict(lds_texels=20480), dict(lds_chunk=8), dict(blocks_per_cu=9)):
        q = RayCasterProjector()
        for k, v in bad.items():
            setattr(q.settings(), k, v)
        with pytest.raises(CtlProjError) as ei:
            q.configure((w.pmats, w.channels, w.rows))
            q.project(ResidentVolume(VoxelVolume(w.volume, w.voxel_mm)))
        assert ei.value.status == _lib.ERR_INVALID_ARGUMENT, bad


def ctypes_ref():
    import ctypes
    return ctypes.byref(ctypes.c_void_p())


def test_progress_callback_order():
    w = W.small(views=36)
    p = RayCasterProjector()
    p.settings().raySampling = w.ray_sampling
    p.configure((w.pmats, w.channels, w.rows))
    seen = []
    p.progress = seen.append
    p.project(ResidentVolume(VoxelVolume(w.volume, w.voxel_mm)))
    assert seen == list(range(36))


def test_cpp_mirror_program(tmp_path):
    """The header-only C++ mirror (include/ctlproj.hpp), driven like CTL code, matches ref32."""
    import subprocess
    from pathlib import Path
    exe = Path(__file__).resolve().parent.parent / "ctl_amd" / "lib" / "projector_demo"
    out = tmp_path / "proj.f32"
    res = subprocess.run([str(exe), str(out)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    assert "views=36 progress=36 empty_ok=1" in res.stdout
    got = np.fromfile(out, np.float32).reshape(36, 1, 128, 256)
    w = W.small()
    r32, _ = run_ref(w)
    e, m = oracle.rel_l2(r32, got), oracle.max_abs_rel(r32, got)
    print(f"C++ mirror: rel L2 {e:.3e}, max abs rel {m:.3e}")
    assert e <= L2_F32 and m <= MAR_F32


def test_parameter_serialization_round_trip():
    """parameter()/setParameter() and toVariant()/fromVariant() keep the reference's keys and type
    id (raycasterprojector.cpp:60-91, abstractprojector.h:114-123)."""
    from ctl_amd import FlatPanelDetector, Settings
    p = RayCasterProjector()
    s = p.settings()
    s.raysPerPixel, s.raySampling, s.volumeUpSampling, s.interpolate = [2, 3], 0.2, 2, False
    v = p.toVariant()
    assert v["type-id"] == 1 and v["#"] == "RayCasterProjector"
    assert v["parameters"] == {"Rays per pixel X": 2, "Rays per pixel Y": 3,
                               "Ray sampling step length": pytest.approx(0.2),
                               "Volume upsampling factor": 2, "Interpolate": False}
    q = RayCasterProjector()
    q.fromVariant(v)
    assert q.parameter() == p.parameter()
    q.setParameter({})  # missing keys take the reference defaults
    assert q.parameter() == {"Rays per pixel X": 1, "Rays per pixel Y": 1,
                             "Ray sampling step length": pytest.approx(0.3),
                             "Volume upsampling factor": 1, "Interpolate": True}
    o = Settings.optimizedFor(VoxelVolume(np.zeros((4, 4, 4), np.float32), (0.25, 0.25, 0.25)),
                              FlatPanelDetector((16, 16), (1.0, 2.0)))
    assert o.raysPerPixel == [4, 8] and o.volumeUpSampling == 1
