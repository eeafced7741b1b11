"""Two projectors driven from two host threads at once (contexts and volumes shared, one stream
per projector): every result equals the one computed alone, bit for bit. The reference's
projector is not re-entrant (raycasterprojector.cpp:278-297 mutates a shared kernel object); here
each projector is single-threaded and the context and volume registry are shared safely."""
from __future__ import annotations

import threading

import numpy as np
import pytest

from ctl_amd import RayCasterProjector, ResidentVolume, VoxelVolume
from ctl_amd import workloads as W

pytestmark = pytest.mark.gpu


def make(w, storage="f32", **settings):
    p = RayCasterProjector()
    p.settings().raySampling = w.ray_sampling
    for k, v in settings.items():
        setattr(p.settings(), k, v)
    p.configure((w.pmats, w.channels, w.rows))
    return p


def test_two_projectors_on_two_threads():
    a = W.small(views=36)
    dims, vs, off = (37, 29, 41), (1.3, 0.9, 1.7), (6.0, -4.5, 9.0)
    vol_b = np.ascontiguousarray(W.ellipsoid_phantom(dims, vs, 9, 3101, semi_axes=(3.0, 14.0)))
    b = W.Workload("odd", vol_b, vs, off, W.carm_circle(9, (77, 53), (1.1, 0.9)), 77, 53, 0.15)
    shared = ResidentVolume(VoxelVolume(a.volume, a.voxel_mm, a.offset_mm))
    own_b = ResidentVolume(VoxelVolume(b.volume, b.voxel_mm, b.offset_mm), "f16")
    pa, pb = make(a), make(b, lds_chunk=-1)
    pc = make(W.Workload("a-nearest", a.volume, a.voxel_mm, a.offset_mm, a.pmats[::3], 256, 128, 0.1),
              interpolate=False)
    jobs = [(pa, shared), (pb, own_b), (pc, shared)]
    alone = [p.project(v).data.copy() for p, v in jobs]
    got = [[] for _ in jobs]
    errors = []

    def worker(k):
        try:
            p, v = jobs[k]
            for _ in range(6):
                got[k].append(p.project(v).data.copy())
        except Exception as e:  # reported below, on the main thread
            errors.append(e)

    threads = [threading.Thread(target=worker, args=(k,)) for k in range(len(jobs))]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not errors, errors
    assert all(not t.is_alive() for t in threads)
    for k, outs in enumerate(got):
        assert len(outs) == 6
        for o in outs:
            assert np.array_equal(o, alone[k]), k
