"""Serving loop: results alone and in a mix are identical, admission by HBM footprint,
rejection of oversized requests, agreement with the oracle."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from ctl_amd import CtlProjError, ResidentVolume, Server, Settings, VoxelVolume
from ctl_amd import _lib
from ctl_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def phantoms():
    vols = []
    for k in range(3):
        data = W.ellipsoid_phantom((64, 64, 64), (2.0, 2.0, 2.0), 12, 5000 + k)
        vols.append((data, ResidentVolume(VoxelVolume(data, (2.0, 2.0, 2.0)), "f16")))
    return vols


def drr_geometry(seed, n=128):
    rng = np.random.default_rng(seed)
    pos, rot = W.carm_arc_poses(rng, 1)
    setup = W.AcquisitionSetup(W.CTSystem(W.CarmGantry(1200.0),
                                          W.FlatPanelDetector((n, n), (2.0, 2.0))))
    setup.setCarmPoses(pos, rot.reshape(1, 9))
    return W.GeometryEncoder.encodeFullGeometry(setup)


def scan_geometry(views=24, n=96):
    return W.carm_circle(views, (n, n), (2.5, 2.5))


def run_alone(server, vol, pm, n):
    out = np.zeros(pm.shape[0] * n * n, np.float32)
    rid = server.submit(vol, pm, n, n, out, Settings(raySampling=0.25))
    server.wait(rid)
    return out


def test_alone_equals_in_mix(phantoms):
    srv = Server(hbm_budget=256 << 20, max_launch_pixels=1 << 18)
    reqs = []
    for k in range(12):
        vol = phantoms[k % 3][1]
        if k % 4 == 3:
            pm, n = scan_geometry(), 96
        else:
            pm, n = drr_geometry(100 + k), 128
        reqs.append((vol, pm, n))
    alone = [run_alone(srv, v, pm, n) for v, pm, n in reqs]
    outs = [np.zeros(pm.shape[0] * n * n, np.float32) for _, pm, n in reqs]
    ids = [srv.submit(v, pm, n, n, o, Settings(raySampling=0.25)) for (v, pm, n), o in zip(reqs, outs)]
    lat = [srv.wait(i) for i in ids]
    for a, o in zip(alone, outs):
        assert np.array_equal(a, o)
    st = srv.stats()
    print("serve stats:", st)
    assert st["completed"] == 24 and st["failed"] == 0 and st["p99_ms"] >= st["p50_ms"] > 0
    assert all(l > 0 for l in lat)
    srv.close()


def test_served_result_matches_oracle(phantoms):
    srv = Server(hbm_budget=64 << 20)
    data, vol = phantoms[1]
    pm = drr_geometry(7)
    got = run_alone(srv, vol, pm, 128).reshape(1, 1, 128, 128)
    r64, _ = oracle.project(data, (2.0, 2.0, 2.0), (0, 0, 0), pm, 128, 128, 0.25, fp64=True)
    e = oracle.rel_l2(r64, got)
    print(f"served DRR (f16 storage) vs ref64: rel L2 {e:.3e}")
    assert e <= 5e-4
    srv.close()


def test_admission_by_hbm_footprint_and_rejection(phantoms):
    # each scan needs 24*96*96*4 B = 0.84 MiB of output (+ geometry): a 2 MiB arena admits two
    srv = Server(hbm_budget=2 << 20)
    vol = phantoms[0][1]
    pm = scan_geometry()
    outs = [np.zeros(24 * 96 * 96, np.float32) for _ in range(6)]
    ids = [srv.submit(vol, pm, 96, 96, o, Settings(raySampling=0.25)) for o in outs]
    for i in ids:
        srv.wait(i)
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])
    big = np.zeros(200 * 128 * 128, np.float32)  # 12.5 MiB > 2 MiB budget
    with pytest.raises(CtlProjError) as ei:
        srv.submit(vol, scan_geometry(200, 128), 128, 128, big, Settings(raySampling=0.25))
    assert ei.value.status == _lib.ERR_OVER_BUDGET
    assert srv.stats()["rejected"] == 1
    srv.close()


def test_wait_timeout_and_drain(phantoms):
    srv = Server(hbm_budget=64 << 20)
    vol = phantoms[2][1]
    out = np.zeros(24 * 96 * 96, np.float32)
    rid = srv.submit(vol, scan_geometry(), 96, 96, out, Settings(raySampling=0.25))
    try:
        srv.wait(rid, timeout=0.0)
    except CtlProjError as e:
        assert e.status == _lib.ERR_TIMEOUT
    srv.drain()
    assert srv.wait(rid) > 0
    srv.close()


def test_single_view_requests_pass_a_running_scan(phantoms):
    # a long scan cut into many launches, then DRRs: the DRRs run on the free launch slot beside
    # the scan's launches and finish first, with the same result as alone; the scan is unaffected.
    # Launches of 16 views, each several times a DRR's duration, so that the completion thread sees
    # a DRR finish while a scan launch still runs.
    srv = Server(hbm_budget=256 << 20, max_launch_pixels=1 << 20)
    vol = phantoms[0][1]
    scan_pm = scan_geometry(360, 256)
    drr_pms = [drr_geometry(300 + k) for k in range(4)]
    scan_alone = run_alone(srv, vol, scan_pm, 256)
    drr_alone = [run_alone(srv, vol, pm, 128) for pm in drr_pms]
    scan_out = np.zeros_like(scan_alone)
    drr_outs = [np.zeros_like(a) for a in drr_alone]
    st = Settings(raySampling=0.25)
    sid = srv.submit(vol, scan_pm, 256, 256, scan_out, st)
    dids = [srv.submit(vol, pm, 128, 128, o, st) for pm, o in zip(drr_pms, drr_outs)]
    drr_lat = [srv.wait(i) for i in dids]
    scan_lat = srv.wait(sid)
    st = srv.stats()
    print(f"scan {scan_lat:.2f} ms, DRRs {', '.join(f'{t:.2f}' for t in drr_lat)} ms; {st['launches']} launches, "
          f"{st['cu_masked_slots']}/{st['slots']} CU-masked slot streams, at most "
          f"{st['max_scan_launches_in_flight']} scan launches in flight, {st['drr_launches_beside_scan']} DRR "
          f"launches done beside a running scan launch")
    assert max(drr_lat) < scan_lat
    # the launch slots: every slot stream has a hardware queue of its own (a CU-masked stream; a
    # silent fall back to plain streams would queue DRRs behind scans), scans never take the last
    # slot, and DRR launches complete while a scan launch of the same run is still running
    assert st["slots"] == 3 and st["cu_masked_slots"] == 3
    assert 1 <= st["max_scan_launches_in_flight"] <= st["slots"] - 1
    assert st["drr_launches_beside_scan"] >= 1
    assert np.array_equal(scan_out, scan_alone)
    for a, o in zip(drr_alone, drr_outs):
        assert np.array_equal(a, o)
    srv.close()
