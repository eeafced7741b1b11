"""Empty-space skipping: chunks (LDS-staged caster) and 64-sample segments (unstaged casters) whose
samples read only zero voxels are skipped, using the volume's occupancy table (a summed-area table
of non-empty 4^3-voxel bricks). Such samples are exactly 0, so every projection must equal, bit for
bit, the one taken with CTLPROJ_EMPTY_SKIP=0 (no table: every sample taken).

  - the 64^3 phantom (60% zero voxels) in FP32 / FP16 / BF16, LDS-staged, unstaged, nearest and
    2 x 2 rays per pixel: skip == no skip, and fewer samples taken;
  - sparse volumes of a few non-zero voxels at random places and at the corners of the grid: a
    brick missing from the table would drop its voxel's rays; skip == no skip and the oracle's
    ref32 on the same voxels;
  - a batch and the server over tabled volumes equal the same launches over untabled ones.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from ctl_amd import Batch, RayCasterProjector, ResidentVolume, Server, Settings, VoxelVolume
from ctl_amd import workloads as W

pytestmark = pytest.mark.gpu

MODES = {
    "staged": {},
    "unstaged": {"lds_chunk": -1},
    "nearest": {"interpolate": False},
    "rpp2x2": {"raysPerPixel": [2, 2]},
}


@pytest.fixture(scope="module")
def torch():
    import torch as t
    return t


def resident(vol, voxel, offset, storage, skip, monkeypatch):
    monkeypatch.setenv("CTLPROJ_EMPTY_SKIP", "1" if skip else "0")
    return ResidentVolume(VoxelVolume(vol, voxel, offset), storage)


def projector(pmats, channels, rows, sampling, **settings):
    p = RayCasterProjector()
    p.settings().raySampling = sampling
    for k, v in settings.items():
        setattr(p.settings(), k, v)
    p.configure((pmats, channels, rows))
    return p


@pytest.mark.parametrize("mode", list(MODES))
@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
def test_skip_equals_all_samples(torch, storage, mode, monkeypatch):
    w = W.small(views=12)
    assert (w.volume == 0).mean() > 0.3  # the phantom has empty space to skip
    on = resident(w.volume, w.voxel_mm, w.offset_mm, storage, True, monkeypatch)
    off = resident(w.volume, w.voxel_mm, w.offset_mm, storage, False, monkeypatch)
    p = projector(w.pmats, w.channels, w.rows, w.ray_sampling, **MODES[mode])
    n = w.views * w.rows * w.channels
    a = torch.empty(n, device="cuda")
    b = torch.empty(n, device="cuda")
    taken_on = p.count_taken(on, a.data_ptr(), n)
    taken_off = p.count_taken(off, b.data_ptr(), n)
    ref_count = p.count_samples(on)
    ga, gb = a.cpu().numpy(), b.cpu().numpy()
    print(f"{storage} {mode}: samples taken {taken_on} of {taken_off} ({taken_on / taken_off:.3f}); "
          f"reference loop count {ref_count}")
    assert np.abs(gb).max() > 0
    assert np.array_equal(ga, gb)
    assert 0 < taken_on < taken_off <= ref_count
    # project() (host output path) takes the same samples
    assert np.array_equal(p.project(on).data.reshape(-1), gb)


@pytest.mark.parametrize("seed", [0, 1, 2])
@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_sparse_voxels(torch, storage, seed, monkeypatch):
    rng = np.random.default_rng(100 + seed)
    n = (45, 38, 52)  # odd sizes: bricks straddle the grid's far faces
    vol = np.zeros(n[::-1], np.float32)
    corners = [(0, 0, 0), (n[0] - 1, n[1] - 1, n[2] - 1), (0, n[1] - 1, 0), (n[0] - 1, 0, n[2] - 1)]
    picks = corners + [tuple(int(rng.integers(0, n[k])) for k in range(3)) for _ in range(6)]
    for x, y, z in picks:
        vol[z, y, x] = float(rng.uniform(0.5, 2.0))
    voxel, offset = (1.0, 1.1, 0.9), (1.5, -2.0, 0.5)
    pm = W.carm_circle(10, (160, 128), (0.9, 0.9))
    for mode in ("staged", "unstaged", "nearest"):
        on = resident(vol, voxel, offset, storage, True, monkeypatch)
        off = resident(vol, voxel, offset, storage, False, monkeypatch)
        p = projector(pm, 160, 128, 0.25, **MODES[mode])
        ga, gb = p.project(on).data, p.project(off).data
        assert np.array_equal(ga, gb), mode
        if storage == "f32":
            r32, _ = oracle.project(vol, voxel, offset, pm, 160, 128, 0.25,
                                    interpolate=(mode != "nearest"), threads=16)
            e = oracle.rel_l2(r32, ga)
            print(f"sparse seed {seed} {mode}: {np.count_nonzero(ga)} non-zero pixels, rel L2 vs ref32 {e:.2e}")
            assert np.count_nonzero(r32) > 0
            # every ray that crosses a non-zero voxel's support is non-zero in both
            assert np.array_equal(r32 != 0, ga != 0)
            assert e <= 1e-4


def test_batch_and_server_skip_equals_all_samples(torch, monkeypatch):
    w = W.small(views=6)
    on = resident(w.volume, w.voxel_mm, w.offset_mm, "f16", True, monkeypatch)
    off = resident(w.volume, w.voxel_mm, w.offset_mm, "f16", False, monkeypatch)
    n = w.views * w.rows * w.channels
    outs = {}
    for name, vol in (("on", on), ("off", off)):
        o = torch.empty(n, device="cuda")
        b = Batch()
        b.add(vol, w.pmats, w.channels, w.rows, o.data_ptr(), Settings(raySampling=w.ray_sampling))
        b.finalize(autotune=False)
        b.launch()
        torch.cuda.synchronize()
        srv = Server(1 << 30)
        host = np.zeros(n, np.float32)
        srv.wait(srv.submit(vol, w.pmats, w.channels, w.rows, host, Settings(raySampling=w.ray_sampling)))
        srv.close()
        outs[name] = (o.cpu().numpy(), host)
    assert np.abs(outs["off"][0]).max() > 0
    assert np.array_equal(outs["on"][0], outs["off"][0])
    assert np.array_equal(outs["on"][1], outs["off"][1])
