"""Volumes of 4 GiB or more, against the oracle.

A 1024^3 phantom of 0.125 mm voxels (20 ellipsoids, 128 mm across) is 8.6 GB of resident 8-byte
texels, beyond the 32-bit byte offsets of one buffer: the ray casters read it with 64-bit
addresses. Two views 45 degrees apart on a 256 x 192 panel (1.6 mm pixels), ray step 0.2 voxel
(about 5000 samples per ray):
  FP32 trilinear, LDS-staged and unstaged, against ref32;
  FP32 nearest (unstaged) against ref32;
  FP16 trilinear (staged) against ref64.
Bars as in test_gpu_parity.py: FP32 vs ref32 relative L2 <= 1e-4 and max-abs-rel <= 1e-3;
FP16 vs ref64 relative L2 <= 5e-4.

CTLPROJ_VOLUME_WIDE=1 reads a small volume the way large ones are read (64-bit gather addresses,
LDS staging through per-box buffer windows): on the 64^3 workload that equals the ordinary reads
bit for bit, and the paths that take volumes below 4 GiB only (batches, the server, volume
processing, the gather probe) refuse it.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from ctl_amd import Batch, CtlProjError, RayCasterProjector, ResidentVolume, Server, VoxelVolume
from ctl_amd.processing import VolumeSlicer
from ctl_amd import workloads as W

pytestmark = pytest.mark.gpu

L2_F32, MAR_F32, L2_F16 = 1e-4, 1e-3, 5e-4
THREADS = 16  # the oracle's worker threads
VIEWS = [0, 1]


@pytest.fixture(scope="module")
def big():
    return W.real(views=8, n=1024, panel=(256, 192))


@pytest.fixture(scope="module")
def resident(big):
    made = {}

    def get(storage):
        if storage not in made:
            for v in made.values():  # one 8.6 GB volume at a time
                v.close()
            made.clear()
            made[storage] = ResidentVolume(VoxelVolume(big.volume, big.voxel_mm, big.offset_mm), storage)
        return made[storage]

    yield get
    for v in made.values():
        v.close()


def project(w, vol, views=VIEWS, **settings):
    p = RayCasterProjector()
    p.settings().raySampling = w.ray_sampling
    for k, v in settings.items():
        setattr(p.settings(), k, v)
    p.configure((w.pmats[views], w.channels, w.rows))
    return p.project(vol).data


@pytest.mark.parametrize("lds_chunk", [0, -1], ids=["staged", "unstaged"])
def test_large_fp32_vs_ref32(big, resident, lds_chunk):
    vol = resident("f32")
    assert vol.nbytes() >= 4 << 30
    r32, _ = oracle.project(big.volume, big.voxel_mm, big.offset_mm, big.pmats, big.channels, big.rows,
                            big.ray_sampling, views=VIEWS, threads=THREADS)
    got = project(big, vol, lds_chunk=lds_chunk)
    e, m = oracle.rel_l2(r32, got), oracle.max_abs_rel(r32, got)
    print(f"1024^3 f32 ({'staged' if lds_chunk == 0 else 'unstaged'}): rel L2 {e:.3e}, max abs rel {m:.3e}")
    assert np.abs(r32).max() > 0
    assert e <= L2_F32 and m <= MAR_F32


def test_large_fp32_nearest_vs_ref32(big, resident):
    vol = resident("f32")
    r32, _ = oracle.project(big.volume, big.voxel_mm, big.offset_mm, big.pmats, big.channels, big.rows,
                            big.ray_sampling, interpolate=False, views=VIEWS, threads=THREADS)
    got = project(big, vol, interpolate=False)
    e, m = oracle.rel_l2(r32, got), oracle.max_abs_rel(r32, got)
    print(f"1024^3 f32 nearest: rel L2 {e:.3e}, max abs rel {m:.3e}")
    assert e <= L2_F32 and m <= MAR_F32


def test_large_fp16_vs_ref64(big, resident):
    vol = resident("f16")
    assert vol.nbytes() >= 4 << 30
    r64, _ = oracle.project(big.volume, big.voxel_mm, big.offset_mm, big.pmats, big.channels, big.rows,
                            big.ray_sampling, fp64=True, views=VIEWS, threads=THREADS)
    got = project(big, vol)
    e = oracle.rel_l2(r64, got)
    print(f"1024^3 f16: rel L2 vs ref64 {e:.3e}")
    assert e <= L2_F16


MODES = {"staged": {}, "unstaged": {"lds_chunk": -1}, "nearest": {"interpolate": False}}


@pytest.mark.parametrize("mode", list(MODES))
@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
def test_wide_addressing_equals_buffer_reads(storage, mode, monkeypatch):
    w = W.small(views=8)
    views = list(range(8))
    vv = VoxelVolume(w.volume, w.voxel_mm, w.offset_mm)
    ref = project(w, ResidentVolume(vv, storage), views, **MODES[mode])
    monkeypatch.setenv("CTLPROJ_VOLUME_WIDE", "1")
    got = project(w, ResidentVolume(vv, storage), views, **MODES[mode])
    assert np.abs(ref).max() > 0
    assert np.array_equal(ref, got)


def test_wide_volumes_only_project(monkeypatch):
    monkeypatch.setenv("CTLPROJ_VOLUME_WIDE", "1")
    w = W.small(views=2)
    vol = ResidentVolume(VoxelVolume(w.volume, w.voxel_mm, w.offset_mm), "f32")
    with pytest.raises(CtlProjError, match="below 4 GiB"):
        Batch().add(vol, w.pmats, w.channels, w.rows, 0)
    with pytest.raises(CtlProjError, match="below 4 GiB"):
        Server(1 << 30).submit(vol, w.pmats, w.channels, w.rows, np.zeros(2 * w.channels * w.rows, np.float32))
    with pytest.raises(CtlProjError, match="below 4 GiB"):
        VolumeSlicer(vol).slice(0.0, 0.0, 0.0)
    import torch
    out = torch.empty(2 * w.channels * w.rows, device="cuda")
    p = RayCasterProjector()
    p.configure((w.pmats, w.channels, w.rows))
    with pytest.raises(CtlProjError, match="below 4 GiB"):
        p.gather_probe(vol, "gather", out.data_ptr(), out.numel())
