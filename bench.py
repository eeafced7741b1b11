"""Benchmark of the forward projector on one MI355X.

Headline (`value`): the FP32 drop-in on a 256^3 ellipsoid phantom, 360 views, 1024x768 flat
panel, ray step 0.2 voxel. One step = one full 360-view projection with the volume resident in
HBM and the 1.13 GB output written to HBM. In the same run: FP16/BF16 storage, the helical
workload (40-module curved detector), 64 uneven instances in one persistent graph-captured
launch, and a served request stream (p50/p99 latency); errors against the CPU oracle (ref32 and
ref64) and the oracle timed on the host cores as the CPU baseline.

python bench.py [--steps K] [--warmup W] [--quick]   (one GPU; --gpus N > 1 exits)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBPS = 8000.0
# On-chip data paths the staged ray caster gathers through, per CU per clock on gfx950, times
# 256 CUs x 2.4 GHz:
#   vector L1 -> registers, 16-byte loads: 64 B/clk (39.3 TB/s); the stream probe (coalesced
#     16-byte loads of an L1-resident window) reaches 98% of it;
#   LDS -> registers, ds_read_b128: 256 B/clk (157.3 TB/s); the LDS probe (conflict-free 16-byte
#     reads) reaches 85-90% of it. Two adjacent 8-byte reads per lane (ds_read2_b64, the way the
#     ray caster reads its texel pairs) reach only a quarter of that (the lds_pairs probe).
L1_PEAK_GBPS = 256 * 2.4 * 64.0
LDS_PEAK_GBPS = 256 * 2.4 * 256.0
ONCHIP_PEAK_GBPS = L1_PEAK_GBPS + LDS_PEAK_GBPS

# HBM traffic per launch of the 360-view FP32 projection (LDS-staged kernel), from the rocprofv3
# --pmc FETCH_SIZE and WRITE_SIZE passes in profiles/staged_kernel_f32_256cube_360views_1024x768.md, in
# KiB per dispatch, each scaled by what the counter reads for a plain 1 GiB device copy
# (bench.py --copy-calibration: FETCH_SIZE 524315 KiB -> x2.0, WRITE_SIZE 1048576 KiB -> x1.0).
# Chunks that read only zero voxels are skipped, so the kernel fetches about half the bytes it
# fetches with every sample taken (5.11e7 KiB).
FETCH_KIB, FETCH_CAL = 2.6607e7, 1048576 / 524315
WRITE_KIB, WRITE_CAL = 1.1499e6, 1.0
TRAFFIC_F32_BYTES = (FETCH_KIB * FETCH_CAL + WRITE_KIB * WRITE_CAL) * 1024


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def timed_projection(p, vol, out, steps, warmup):
    n_out = out.numel()
    for _ in range(warmup):
        p.project_device(vol, out.data_ptr(), n_out)
    p.synchronize()
    kern = []
    t0 = time.perf_counter()
    for _ in range(steps):
        p.project_device(vol, out.data_ptr(), n_out)
        kern.append(p.last_kernel_ms())  # waits for this step's kernel
    p.synchronize()
    wall = (time.perf_counter() - t0) / steps
    return wall, float(np.mean(kern))


def bench_real(args, torch, W, ctl, oracle):
    w = W.real(views=360)
    log(f"workload: {w.name}")
    p = ctl.RayCasterProjector()
    p.settings().raySampling = w.ray_sampling
    if args.no_lds:
        p.settings().lds_chunk = -1
    p.configure((w.pmats, w.channels, w.rows))
    out = torch.empty(w.views * w.modules * w.rows * w.channels, dtype=torch.float32, device="cuda")
    res = {}
    for storage in args.variants.split(","):
        vol = ctl.ResidentVolume(ctl.VoxelVolume(w.volume, w.voxel_mm, w.offset_mm), storage)
        n_samples = p.count_samples(vol)
        wall, kms = timed_projection(p, vol, out, args.steps, args.warmup)
        if args.kernels_only:  # profiling: the timed launches and nothing else of this kernel
            res[storage] = dict(ms_per_step=wall * 1e3, kernel_ms=kms, n_samples=n_samples)
            log(f"real {storage}: {json.dumps(res[storage])}")
            del vol
            continue
        # samples the kernels take: chunks that read only zero voxels (60% of this phantom's
        # voxels are 0) are skipped with an identical result
        n_taken = p.count_taken(vol, out.data_ptr(), out.numel())
        eb = 4 if storage == "f32" else 2
        res[storage] = dict(ms_per_step=wall * 1e3, kernel_ms=kms, n_samples=n_samples, n_taken=n_taken,
                            gsamples_s=n_samples / wall / 1e9, views_s=w.views / wall,
                            kernel_gsamples_s=n_samples / (kms * 1e-3) / 1e9,
                            taken_gsamples_s=n_taken / (kms * 1e-3) / 1e9,
                            gather_GBps=n_taken * 8 * eb / (kms * 1e-3) / 1e9)
        # the same projection with every sample taken (a volume without occupancy table)
        os.environ["CTLPROJ_EMPTY_SKIP"] = "0"
        try:
            dense = ctl.ResidentVolume(ctl.VoxelVolume(w.volume, w.voxel_mm, w.offset_mm), storage)
        finally:
            del os.environ["CTLPROJ_EMPTY_SKIP"]
        dwall, dkms = timed_projection(p, dense, out, args.steps, args.warmup)
        res[storage]["all_samples"] = dict(ms_per_step=dwall * 1e3, kernel_ms=dkms,
                                           gsamples_s=n_samples / dwall / 1e9,
                                           gather_GBps=n_samples * 8 * eb / (dkms * 1e-3) / 1e9)
        del dense
        # gather ceiling of this access pattern: the same loads without the interpolation
        pg = [p.gather_probe(vol, "gather", out.data_ptr(), out.numel()) for _ in range(4)][1:]
        ceil = float(np.median([b / (ms * 1e-3) / 1e9 for ms, b in pg]))
        res[storage]["probe_gather_GBps"] = ceil
        res[storage]["frac_of_probe_gather"] = res[storage]["gather_GBps"] / ceil
        if storage == "f32":
            ps = [p.gather_probe(vol, "stream", out.data_ptr(), out.numel()) for _ in range(4)][1:]
            res[storage]["probe_l1_stream_GBps"] = float(np.median([b / (ms * 1e-3) / 1e9 for ms, b in ps]))
            for mode in ("lds", "lds_pairs"):
                pl = [p.gather_probe(vol, mode, out.data_ptr(), out.numel()) for _ in range(4)][1:]
                res[storage][f"probe_{mode}_GBps"] = float(np.median([b / (ms * 1e-3) / 1e9 for ms, b in pl]))
        log(f"real {storage}: {json.dumps(res[storage])}")
        del vol
    if args.kernels_only:
        return w, res
    if not args.no_errors:
        pick = [0, 137]
        r32, _ = oracle.project(w.volume, w.voxel_mm, w.offset_mm, w.pmats, w.channels, w.rows,
                                w.ray_sampling, views=pick, threads=16)
        r64, _ = oracle.project(w.volume, w.voxel_mm, w.offset_mm, w.pmats, w.channels, w.rows,
                                w.ray_sampling, views=pick, threads=16, fp64=True)
        p2 = ctl.RayCasterProjector()
        p2.settings().raySampling = w.ray_sampling
        p2.configure((w.pmats[pick], w.channels, w.rows))
        for storage in res:
            got = p2.project(ctl.ResidentVolume(ctl.VoxelVolume(w.volume, w.voxel_mm), storage)).data
            res[storage]["rel_l2_vs_ref64"] = oracle.rel_l2(r64, got)
            res[storage]["error_views"] = pick
            if storage == "f32":
                res[storage]["rel_l2_vs_ref32"] = oracle.rel_l2(r32, got)
                res[storage]["max_abs_rel_vs_ref32"] = oracle.max_abs_rel(r32, got)
    # end-to-end reference-style project(): host volume upload, projection, host output
    pr = ctl.RayCasterProjector()
    pr.settings().raySampling = w.ray_sampling
    pr.configure((w.pmats, w.channels, w.rows))
    t0 = time.perf_counter()
    pr.project(ctl.VoxelVolume(w.volume, w.voxel_mm, w.offset_mm))
    res["f32"]["end_to_end_s"] = time.perf_counter() - t0
    log(f"real f32 end to end (upload + project + 1.13 GB to host): {res['f32']['end_to_end_s']:.3f} s")
    return w, res


def bench_helical(args, torch, W, ctl, oracle):
    w = W.helical(views=2000)
    log(f"workload: {w.name}")
    p = ctl.RayCasterProjector()
    p.settings().raySampling = w.ray_sampling
    p.configure((w.pmats, w.channels, w.rows))
    out = torch.empty(w.views * w.modules * w.rows * w.channels, dtype=torch.float32, device="cuda")
    res = {}
    for storage in ("f32", "f16"):
        vol = ctl.ResidentVolume(ctl.VoxelVolume(w.volume, w.voxel_mm, w.offset_mm), storage)
        n_samples = p.count_samples(vol)
        wall, kms = timed_projection(p, vol, out, max(1, args.steps // 2), 1)
        res[storage] = dict(ms_per_step=wall * 1e3, kernel_ms=kms, n_samples=n_samples,
                            gsamples_s=n_samples / wall / 1e9, views_s=w.views / wall)
        if not args.kernels_only:
            res[storage]["n_taken"] = p.count_taken(vol, out.data_ptr(), out.numel())
        log(f"helical {storage}: {json.dumps(res[storage])}")
        if not args.no_errors:
            pick = [0, 1234]
            r64, _ = oracle.project(w.volume, w.voxel_mm, w.offset_mm, w.pmats, w.channels, w.rows,
                                    w.ray_sampling, views=pick, threads=16, fp64=True)
            p2 = ctl.RayCasterProjector()
            p2.settings().raySampling = w.ray_sampling
            p2.configure((w.pmats[pick], w.channels, w.rows))
            got = p2.project(vol).data
            res[storage]["rel_l2_vs_ref64"] = oracle.rel_l2(r64, got)
            if storage == "f32":
                r32, _ = oracle.project(w.volume, w.voxel_mm, w.offset_mm, w.pmats, w.channels,
                                        w.rows, w.ray_sampling, views=pick, threads=16)
                res[storage]["rel_l2_vs_ref32"] = oracle.rel_l2(r32, got)
        del vol
    return res


def bench_batched(args, torch, W, ctl):
    items, rs = W.batched(n=64)
    views = sum(i.pmats.shape[0] for i in items)
    log(f"batched: 64 instances, {views} views")
    outs = [torch.empty(i.pmats.shape[0] * i.rows * i.channels, dtype=torch.float32, device="cuda")
            for i in items]
    res = {"instances": 64, "views": views}
    for storage in ("f32", "f16"):
        vols = [ctl.ResidentVolume(ctl.VoxelVolume(i.volume, i.voxel_mm), storage) for i in items]
        b = ctl.Batch()
        for i, v, o in zip(items, vols, outs):
            b.add(v, i.pmats, i.channels, i.rows, o.data_ptr(), ctl.Settings(raySampling=rs))
        t0 = time.perf_counter()
        b.finalize(autotune=True)
        fin = time.perf_counter() - t0
        ns = b.count_samples()
        b.launch()
        b.synchronize()
        t0 = time.perf_counter()
        ms = []
        for _ in range(max(1, args.steps // 2)):
            b.launch()
            ms.append(b.last_ms())
        wall = (time.perf_counter() - t0) / len(ms)
        res[storage] = dict(ms_per_launch=wall * 1e3, kernel_ms=float(np.mean(ms)), n_samples=ns,
                            gsamples_s=ns / wall / 1e9, views_s=views / wall, finalize_s=fin,
                            shape=b.info())
        log(f"batched {storage}: {json.dumps(res[storage])}")
        b.close()
        del vols
    # baseline: 64 sequential plain FP32 project() calls (upload, project, output to host)
    t0 = time.perf_counter()
    for i in items:
        p = ctl.RayCasterProjector()
        p.settings().raySampling = rs
        p.configure((i.pmats, i.channels, i.rows))
        p.project(ctl.VoxelVolume(i.volume, i.voxel_mm))
    seq = time.perf_counter() - t0
    res["sequential_plain_f32"] = dict(s=seq, views_s=views / seq)
    res["speedup_f16_batched_vs_sequential_plain_f32"] = res["f16"]["views_s"] / (views / seq)
    res["speedup_f32_batched_vs_sequential_plain_f32"] = res["f32"]["views_s"] / (views / seq)
    log(f"batched sequential plain f32: {seq:.3f} s")
    return res


def bench_served(args, torch, W, ctl):
    """8 resident 256^3 FP16 phantoms; Poisson arrivals; 80% one-view 512x512 DRRs near a
    reference C-arm pose (+-5 deg, +-10 mm), 20% 180-view 768x768 scans."""
    rng = np.random.default_rng(5)
    phantoms = W.served_phantoms()
    vols = [ctl.ResidentVolume(ctl.VoxelVolume(d, (0.5, 0.5, 0.5)), "f16") for d in phantoms]

    def drr_pose():
        return W.served_drr_geometry(rng)

    scan_pm = W.served_scan_geometry()
    drr_out = [torch.empty(512 * 512, dtype=torch.float32).pin_memory() for _ in range(64)]
    scan_out = [torch.empty(180 * 768 * 768, dtype=torch.float32).pin_memory() for _ in range(4)]
    st = ctl.Settings(raySampling=0.2)
    srv = ctl.Server(hbm_budget=16 << 30)
    rid = srv.submit_ptr(vols[0], scan_pm, 768, 768, scan_out[0].data_ptr(), scan_out[0].numel(), st)
    srv.wait(rid)
    res = {"phantoms": "8 x 256^3 FP16 resident", "mix": "80% DRR 512x512, 20% scan 180x768x768"}
    n = 120 if args.quick else 1000  # requests per arrival rate
    for rate in (200, 500, 1000, 2000):  # the arrival rates of the served workload
        gaps = rng.exponential(1.0 / rate, n)
        kinds = rng.uniform(size=n) < 0.8
        geoms = [drr_pose() if k else None for k in kinds]
        srv.reset_stats()
        t = time.perf_counter()
        for i in range(n):
            t += gaps[i]
            while time.perf_counter() < t:
                pass
            v = vols[int(rng.integers(8))]
            if kinds[i]:
                o = drr_out[i % 64]
                srv.submit_ptr(v, geoms[i], 512, 512, o.data_ptr(), o.numel(), st)
            else:
                o = scan_out[i % 4]
                srv.submit_ptr(v, scan_pm, 768, 768, o.data_ptr(), o.numel(), st)
        srv.drain()
        s = srv.stats()
        # offered load: the mix's mean of 0.8 x 1 + 0.2 x 180 views per request at this rate
        offered = rate * (kinds.sum() + 180 * (n - kinds.sum())) / n
        res[f"rate_{rate}"] = dict(requests=n, offered_views_s=offered, p50_ms=s["p50_ms"], p99_ms=s["p99_ms"],
                                   p50_drr_ms=s["p50_single_ms"], p99_drr_ms=s["p99_single_ms"],
                                   p50_scan_ms=s["p50_multi_ms"], p99_scan_ms=s["p99_multi_ms"],
                                   views_s=s["views"] / s["span_s"], launches=s["launches"])
        log(f"served {rate}/s: {json.dumps(res[f'rate_{rate}'])}")
    srv.close()
    # plain per-request FP32 path: upload the FP32 volume, project, copy to host, per request
    p = ctl.RayCasterProjector()
    p.settings().raySampling = 0.2
    t0 = time.perf_counter()
    nv = 0
    for k in range(10):
        pm = drr_pose() if k < 8 else scan_pm
        p.configure((pm, 512 if k < 8 else 768, 512 if k < 8 else 768))
        p.project(ctl.VoxelVolume(phantoms[k % 8], (0.5, 0.5, 0.5)))
        nv += pm.shape[0]
    dt = time.perf_counter() - t0
    res["plain_per_request_f32"] = dict(requests=10, views=nv, s=dt, views_s=nv / dt)
    res["speedup_views_s"] = res["rate_1000"]["views_s"] / (nv / dt)
    log(f"served plain per-request f32: {nv / dt:.1f} views/s")
    return res


def bench_processing(args, torch, w, ctl, oracle):
    """Volume processing on the resident 256^3 phantom of the headline workload (FP32):
    RadonTransform3D plane integrals (32 x 32 orientations x 128 distances over +-90 mm on the
    default 368^2 template plane of 0.5 mm pixels), 256 VolumeSlicer planes of 363^2 pixels and
    16 Mi VolumeResampler points, every output in device memory."""
    import ctypes as C
    from ctl_amd import _lib
    L = _lib.lib()
    vol = ctl.VoxelVolume(w.volume, w.voxel_mm, w.offset_mm)
    rt = ctl.RadonTransform3D(vol)
    side, reso = rt.sliceDimensions()[0], rt.sliceResolution()
    azi = np.linspace(0.0, np.pi, 32, endpoint=False)
    pol = np.linspace(0.02, np.pi - 0.02, 32)
    normals = np.ascontiguousarray([[np.sin(b) * np.cos(a), np.sin(b) * np.sin(a), np.cos(b)]
                                    for b in pol for a in azi])
    dist = np.linspace(-90.0, 90.0, 128).astype(np.float32)
    out = torch.empty(len(normals) * len(dist), dtype=torch.float32, device="cuda")

    def planes(h):
        _lib.check(L.ctlproj_radon3d_plane_integrals_device(h, _lib.dptr(normals), len(normals),
                                                            _lib.fptr(dist), len(dist), side, reso,
                                                            C.c_void_p(out.data_ptr()), out.numel()))

    planes(rt._vol.h)
    walls = []
    for _ in range(3):
        t0 = time.perf_counter()
        planes(rt._vol.h)
        walls.append(time.perf_counter() - t0)
    wall = float(np.median(walls))
    samples = len(normals) * len(dist) * side * side
    # samples that read the volume: the plane integrals of a volume of ones / pixel area
    ones = ctl.RadonTransform3D(ctl.VoxelVolume(np.ones_like(w.volume), w.voxel_mm, w.offset_mm))
    planes(ones._vol.h)
    inside = float(out.double().sum().item()) / (reso * reso)
    gather = inside * 32 / wall / 1e9
    res = {"radon3d": dict(
        orientations=len(normals), distances=len(dist), side=side, resolution_mm=reso, ms=wall * 1e3,
        template_samples=samples, template_gsamples_s=samples / wall / 1e9,
        inside_samples=inside, inside_gsamples_s=inside / wall / 1e9,
        roofline={"bound": "lds", "achieved": gather, "peak": LDS_PEAK_GBPS, "unit": "GB/s",
                  "frac": gather / LDS_PEAK_GBPS, "frac_of_l1_peak": gather / L1_PEAK_GBPS,
                  "note": "achieved = samples inside the support x 8 taps x 4 B per wall second "
                          "(samples of empty bricks are skipped with the same sums); the brick "
                          "kernel reads its taps from bricks of 14^3 cells staged in LDS (two "
                          "ds_read2_b64 per sample) and is bound by instruction issue and LDS bank "
                          "conflicts (counters in profiles/plane_brick_kernel_f32_256cube_1024orient_128dist.md)"})}
    log(f"radon3d: {json.dumps(res['radon3d'])}")
    del ones
    if oracle is not None and not args.no_cpu:
        pick = normals[::128]  # 8 orientations spread over the sphere
        t0 = time.perf_counter()
        oracle.plane_integrals(w.volume, w.voxel_mm, w.offset_mm, pick, dist[32:96:4], side, side, reso)
        dt = time.perf_counter() - t0
        n_cpu = len(pick) * 16 * side * side
        res["radon3d"]["cpu_baseline"] = {
            "value": n_cpu / dt / 1e9, "unit": "Gsamples/s (template)", "cores": 1, "kind": "port",
            "sample": f"ref32 oracle plane integrals (single thread), 8 orientations x 16 distances "
                      f"of the same workload, {n_cpu} template samples in {dt:.2f} s"}
    # slicer: 256 planes through the volume
    rng = np.random.default_rng(7)
    sl = ctl.VolumeSlicer(rt._vol)
    sw, sh = sl.sliceDimensions()
    nrm = rng.normal(size=(256, 3))
    nrm = np.ascontiguousarray(nrm / np.linalg.norm(nrm, axis=1, keepdims=True))
    dd = rng.uniform(-40, 40, 256)
    sout = torch.empty(256 * sw * sh, dtype=torch.float32, device="cuda")

    def slices():
        _lib.check(L.ctlproj_slice_device(rt._vol.h, _lib.dptr(nrm), _lib.dptr(dd), 256, sw, sh,
                                          sl.sliceResolution(), C.c_void_p(sout.data_ptr()), sout.numel()))

    slices()
    t0 = time.perf_counter()
    for _ in range(3):
        slices()
    swall = (time.perf_counter() - t0) / 3
    res["slicer"] = dict(planes=256, width=sw, height=sh, ms=swall * 1e3,
                         gsamples_s=256 * sw * sh / swall / 1e9,
                         write_GBps=256 * sw * sh * 4 / swall / 1e9)
    # resampler: 16 Mi points inside the ranges
    rsm = ctl.VolumeResampler(rt._vol)
    lo, hi = rsm._ranges[0::2], rsm._ranges[1::2]
    n = 1 << 24
    pts = torch.rand(n, 3, device="cuda") * torch.tensor(hi - lo, device="cuda") + torch.tensor(lo, device="cuda")
    rout = torch.empty(n, dtype=torch.float32, device="cuda")
    rsm.sample_device(pts.data_ptr(), n, rout.data_ptr())
    t0 = time.perf_counter()
    for _ in range(3):
        rsm.sample_device(pts.data_ptr(), n, rout.data_ptr())
    rwall = (time.perf_counter() - t0) / 3
    res["resampler"] = dict(points=n, ms=rwall * 1e3, gsamples_s=n / rwall / 1e9,
                            stream_GBps=n * 16 / rwall / 1e9)
    log(f"processing: slicer {json.dumps(res['slicer'])} resampler {json.dumps(res['resampler'])}")
    return res


def cpu_baseline(w, oracle):
    threads = min(16, os.cpu_count() or 1)
    sub = list(range(0, 360, 45))  # 8 evenly strided views
    t0 = time.perf_counter()
    _, n = oracle.project(w.volume, w.voxel_mm, w.offset_mm, w.pmats, w.channels, w.rows,
                          w.ray_sampling, views=sub, threads=threads)
    dt = time.perf_counter() - t0
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), model)
    except OSError:
        pass
    return {"value": n / dt / 1e9, "unit": "Gsamples/s", "cores": threads, "kind": "port",
            "host_cpus": os.cpu_count(), "cpu_model": model,
            "sample": f"ref32 oracle (C restatement of the reference kernel, pthreads) on views "
                      f"{sub[0]}..{sub[-1]} step 45 (8 of 360 views) of the same workload, "
                      f"{n} samples in {dt:.1f} s",
            "views_per_s": len(sub) / dt}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--variants", default="f32,f16,bf16")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-errors", action="store_true", help="skip the oracle error statements")
    ap.add_argument("--only-real", action="store_true", help="only the 256^3 / 360-view workload")
    ap.add_argument("--quick", action="store_true", help="fewer served requests")
    ap.add_argument("--no-lds", action="store_true",
                    help="real workload with the gathers from HBM through L1 (no LDS staging)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="with --only-real: one timed projection per variant and nothing else (profiling)")
    ap.add_argument("--only-helical", action="store_true",
                    help="only the helical workload (2000 views of 40 modules; profiling)")
    ap.add_argument("--only-processing", action="store_true",
                    help="only the volume processing leg (3D Radon, slicer, resampler; profiling)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB device to device three times (calibrates the FETCH_SIZE / "
                         "WRITE_SIZE counters behind roofline.traffic) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")  # 1 GiB
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": src.numel() * 4, "copies": 3}), flush=True)
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported here. The "
              "projector would shard over views with no collective; this build runs on one GPU.",
              file=sys.stderr)
        return 2

    import torch

    import ctl_amd as ctl
    from ctl_amd import workloads as W
    oracle = None
    if not (args.no_cpu and args.no_errors):
        import oracle

    if args.only_processing:
        res = bench_processing(args, torch, W.real(views=1), ctl, oracle)
        print(json.dumps({"processing": res}), flush=True)
        return 0
    if args.only_helical:
        print(json.dumps({"helical": bench_helical(args, torch, W, ctl, oracle)}), flush=True)
        return 0
    w, real = bench_real(args, torch, W, ctl, oracle)
    if args.kernels_only:
        print(json.dumps({"real": real}), flush=True)
        return 0
    extra = {}
    if not args.only_real:
        extra["helical"] = bench_helical(args, torch, W, ctl, oracle)
        extra["batched"] = bench_batched(args, torch, W, ctl)
        extra["served"] = bench_served(args, torch, W, ctl)
        extra["processing"] = bench_processing(args, torch, w, ctl, oracle)
    cpu = None if args.no_cpu else cpu_baseline(w, oracle)
    if cpu:
        log(f"cpu baseline: {cpu}")

    f32 = real["f32"] if "f32" in real else next(iter(real.values()))  # headline: FP32 unless not run
    achieved = f32["gather_GBps"]
    line = {
        "metric": "Gsamples/s & views/s at 256³/360 views vs plain FP32; served p50/p99 latency",
        "value": f32["gsamples_s"], "unit": "Gsamples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": f32["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": w.name, "volume": "256^3 x 0.5 mm, 20 seeded ellipsoids, FP32",
                   "views": 360, "detector": "1024x768 @ 0.4 mm, C-arm 1200 mm / SID 750 mm",
                   "ray_step_vox": 0.2, "parallelism": "single GPU"},
        "views_per_s": f32["views_s"],
        "roofline": {"bound": "lds+l1", "achieved": achieved, "peak": ONCHIP_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / ONCHIP_PEAK_GBPS, "traffic": TRAFFIC_F32_BYTES,
                     "algorithmic_bytes_per_sample": 32, "samples_per_launch": f32["n_taken"],
                     "reference_samples_per_launch": f32["n_samples"],
                     "kernel_ms": f32["kernel_ms"],
                     "note": "achieved = algorithmic gather bytes (8 taps x 4 B per sample taken; the "
                             "kernel skips chunks whose samples read only zero voxels, "
                             "samples_per_launch of reference_samples_per_launch are taken) / kernel "
                             "time; peak = the on-chip data paths the staged kernel gathers through "
                             "together: LDS (ds_read_b128, 256 B/clk/CU) + vector L1 (64 B/clk/CU), "
                             "256 CUs x 2.4 GHz, each checked by a probe (probe_lds_GBps, "
                             "probe_l1_stream_GBps). traffic = calibrated HBM bytes per launch "
                             "(hbm_GBps, hbm_frac). The kernel reads its texel pairs with "
                             "ds_read2_b64, which the lds_pairs probe measures at a quarter of the "
                             "b128 rate; it is bound by latency and issue together (VALU, the LDS "
                             "pair reads and the texture path with the third of the samples it "
                             "gathers through L1), not by any one data path",
                     "lds_peak": LDS_PEAK_GBPS, "l1_peak": L1_PEAK_GBPS,
                     "probe_lds_GBps": f32.get("probe_lds_GBps"),
                     "probe_lds_pairs_GBps": f32.get("probe_lds_pairs_GBps"),
                     "probe_l1_stream_GBps": f32.get("probe_l1_stream_GBps"),
                     "probe_gather_GBps": f32.get("probe_gather_GBps"),
                     "frac_of_probe_gather": f32.get("frac_of_probe_gather"),
                     "hbm_GBps": TRAFFIC_F32_BYTES / (f32["kernel_ms"] * 1e-3) / 1e9,
                     "hbm_frac": TRAFFIC_F32_BYTES / (f32["kernel_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBPS},
        "cpu_baseline": cpu,
        "variants": real,
    }
    line.update(extra)
    if "f16" in real:
        line["f16_vs_f32_views_s"] = real["f16"]["views_s"] / f32["views_s"]
    if "served" in extra:
        sv = extra["served"]
        line["served_p50_ms"] = sv["rate_200"]["p50_ms"]
        line["served_p99_ms"] = sv["rate_200"]["p99_ms"]
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
