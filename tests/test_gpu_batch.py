"""Batched projection (persistent kernel + hipGraph) against per-instance projections and the
oracle. Bars as in test_gpu_parity.py; batch vs single-projection results with the same launch
shape must be bit-identical (same per-pixel code, same accumulation order)."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from ctl_amd import Batch, CtlProjError, RayCasterProjector, ResidentVolume, Settings, VoxelVolume
from ctl_amd import _lib
from ctl_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    return t


@pytest.fixture(scope="module")
def insts():
    return W.batched(n=6, seed=4100, edge=(32, 96), views=(1, 12), panel=(64, 192))


def single(inst, rs, storage, tile=(8, 8), unroll=4):
    p = RayCasterProjector()
    s = p.settings()
    s.raySampling = rs
    s.tile = tile
    s.unroll = unroll
    p.configure((inst.pmats, inst.channels, inst.rows))
    return p.project(ResidentVolume(VoxelVolume(inst.volume, inst.voxel_mm), storage)).data


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_batch_equals_single_projections(torch, insts, storage):
    items, rs = insts
    vols = [ResidentVolume(VoxelVolume(i.volume, i.voxel_mm), storage) for i in items]
    outs = [torch.zeros(i.pmats.shape[0] * i.rows * i.channels, dtype=torch.float32, device="cuda")
            for i in items]
    b = Batch()
    for i, v, o in zip(items, vols, outs):
        b.add(v, i.pmats, i.channels, i.rows, o.data_ptr(), Settings(raySampling=rs))
    b.finalize(autotune=False)
    info = b.info()
    # default shapes: FP32 trilinear runs the LDS-staged kernel (16x4 tiles, unroll 3), FP16 the
    # gathers from HBM (8x8, unroll 4)
    shape = (info["tile_w"], info["tile_h"], info["unroll"])
    assert shape == ((16, 4, 3) if storage == "f32" else (8, 8, 4))
    b.launch()
    b.synchronize()
    for i, o in zip(items, outs):
        ref = single(i, rs, storage, tile=shape[:2], unroll=shape[2])
        got = o.cpu().numpy().reshape(ref.shape)
        assert np.array_equal(got, ref)


def test_batch_graph_replays_are_stable(torch, insts):
    items, rs = insts
    vols = [ResidentVolume(VoxelVolume(i.volume, i.voxel_mm)) for i in items]
    outs = [torch.zeros(i.pmats.shape[0] * i.rows * i.channels, dtype=torch.float32, device="cuda")
            for i in items]
    b = Batch()
    for i, v, o in zip(items, vols, outs):
        b.add(v, i.pmats, i.channels, i.rows, o.data_ptr(), Settings(raySampling=rs))
    b.finalize(autotune=True)
    print("autotuned batch shape:", b.info())
    b.launch()
    b.synchronize()
    first = [o.cpu().clone() for o in outs]
    for rep in range(3):
        # eager work between replays: overwrite the outputs and run an unrelated projection
        for o in outs:
            o.fill_(-1.0)
        single(items[0], rs, "f32")
        torch.cuda.synchronize()
        b.launch()
        b.synchronize()
        for f, o in zip(first, outs):
            assert torch.equal(f, o.cpu()), f"replay {rep} differs"
    # and the batch agrees with the oracle on one instance
    i = items[2]
    r32, _ = oracle.project(i.volume, i.voxel_mm, (0, 0, 0), i.pmats, i.channels, i.rows, rs)
    got = first[2].numpy().reshape(r32.shape)
    e, m = oracle.rel_l2(r32, got), oracle.max_abs_rel(r32, got)
    print(f"batch instance vs ref32: rel L2 {e:.3e} max abs rel {m:.3e}")
    assert e <= 1e-4 and m <= 1e-3


def test_batch_sample_count_is_sum_of_instances(torch, insts):
    items, rs = insts
    b = Batch()
    total = 0
    keep = []
    for i in items:
        v = ResidentVolume(VoxelVolume(i.volume, i.voxel_mm))
        o = torch.zeros(i.pmats.shape[0] * i.rows * i.channels, dtype=torch.float32, device="cuda")
        keep.append((v, o))
        b.add(v, i.pmats, i.channels, i.rows, o.data_ptr(), Settings(raySampling=rs))
        p = RayCasterProjector()
        p.settings().raySampling = rs
        p.configure((i.pmats, i.channels, i.rows))
        total += p.count_samples(v)
    b.finalize(autotune=False)
    assert b.count_samples() == total


def test_batch_errors(torch, insts):
    items, rs = insts
    b = Batch()
    o = torch.zeros(10, device="cuda")
    with pytest.raises(CtlProjError) as ei:
        b.launch()
    assert ei.value.status == _lib.ERR_NOT_CONFIGURED
    i0, i1 = items[0], items[1]
    v0 = ResidentVolume(VoxelVolume(i0.volume, i0.voxel_mm), "f32")
    v1 = ResidentVolume(VoxelVolume(i1.volume, i1.voxel_mm), "f16")
    o0 = torch.zeros(i0.pmats.shape[0] * i0.rows * i0.channels, device="cuda")
    o1 = torch.zeros(i1.pmats.shape[0] * i1.rows * i1.channels, device="cuda")
    b.add(v0, i0.pmats, i0.channels, i0.rows, o0.data_ptr(), Settings(raySampling=rs))
    with pytest.raises(CtlProjError) as ei:  # one storage type per batch
        b.add(v1, i1.pmats, i1.channels, i1.rows, o1.data_ptr(), Settings(raySampling=rs))
    assert ei.value.status == _lib.ERR_INVALID_ARGUMENT
    with pytest.raises(CtlProjError):  # wrong output size
        b.add(v0, i0.pmats, i0.channels, i0.rows, 0, Settings(raySampling=rs))
    del o
