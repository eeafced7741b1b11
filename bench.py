"""Benchmark: forward projection of a 256^3 ellipsoid phantom, 360 views, 1024x768 flat panel.

One step = one full 360-view projection with the volume resident in HBM and the 1.13 GB output
written to HBM. Prints one JSON line (see the keys below). Run: python bench.py [--steps K]
[--warmup W]. Only one GPU is used; --gpus N > 1 exits with a message.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBPS = 8000.0
# gfx950 vector L1: 256 CUs x 2.4 GHz x 128 B/clk
L1_PEAK_GBPS = 256 * 2.4 * 128.0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-errors", action="store_true", help="skip the oracle error statements")
    ap.add_argument("--variants", default="f32,f16,bf16")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the projector shards over views with no collective; see DESIGN.md)", file=sys.stderr)
        return 2

    import torch

    from ctl_amd import RayCasterProjector, ResidentVolume, VoxelVolume
    from ctl_amd import workloads as W

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    w = W.real(views=360)
    log(f"workload: {w.name}")
    p = RayCasterProjector()
    p.settings().raySampling = w.ray_sampling
    p.configure((w.pmats, w.channels, w.rows))
    n_out = w.views * w.modules * w.rows * w.channels
    out = torch.empty(n_out, dtype=torch.float32, device="cuda")

    results = {}
    for storage in args.variants.split(","):
        vol = ResidentVolume(VoxelVolume(w.volume, w.voxel_mm, w.offset_mm), storage)
        n_samples = p.count_samples(vol)
        for _ in range(args.warmup):
            p.project_device(vol, out.data_ptr(), n_out)
        p.synchronize()
        torch.cuda.synchronize()
        kern = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            p.project_device(vol, out.data_ptr(), n_out)
            kern.append(p.last_kernel_ms())  # waits for this step's kernel
        p.synchronize()
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / args.steps
        kms = float(np.mean(kern))
        eb = 4 if storage == "f32" else 2
        results[storage] = dict(
            ms_per_step=wall * 1e3, kernel_ms=kms, n_samples=n_samples,
            gsamples_s=n_samples / wall / 1e9, views_s=w.views / wall,
            kernel_gsamples_s=n_samples / (kms * 1e-3) / 1e9,
            gather_GBps=n_samples * 8 * eb / (kms * 1e-3) / 1e9)
        log(f"{storage}: {json.dumps(results[storage])}")
        del vol

    # error statements against the oracle on two views (ref32 for FP32, ref64 for all)
    if not args.no_errors:
        import oracle
        pick = [0, 137]
        r32, _ = oracle.project(w.volume, w.voxel_mm, w.offset_mm, w.pmats, w.channels, w.rows,
                                w.ray_sampling, views=pick, threads=16)
        r64, _ = oracle.project(w.volume, w.voxel_mm, w.offset_mm, w.pmats, w.channels, w.rows,
                                w.ray_sampling, views=pick, threads=16, fp64=True)
        p2 = RayCasterProjector()
        p2.settings().raySampling = w.ray_sampling
        p2.configure((w.pmats[pick], w.channels, w.rows))
        for storage in results:
            vol = ResidentVolume(VoxelVolume(w.volume, w.voxel_mm, w.offset_mm), storage)
            got = p2.project(vol).data
            results[storage]["rel_l2_vs_ref64"] = oracle.rel_l2(r64, got)
            if storage == "f32":
                results[storage]["rel_l2_vs_ref32"] = oracle.rel_l2(r32, got)
                results[storage]["max_abs_rel_vs_ref32"] = oracle.max_abs_rel(r32, got)

    cpu = None
    if not args.no_cpu:
        import oracle
        threads = min(16, os.cpu_count() or 1)
        sub = list(range(0, 360, 45))  # 8 evenly strided views
        t0 = time.perf_counter()
        _, n = oracle.project(w.volume, w.voxel_mm, w.offset_mm, w.pmats, w.channels, w.rows,
                              w.ray_sampling, views=sub, threads=threads)
        dt = time.perf_counter() - t0
        cpu = {"value": n / dt / 1e9, "unit": "Gsamples/s", "cores": threads, "kind": "port",
               "sample": f"ref32 oracle on views {sub[0]}..{sub[-1]} step 45 (8 of 360 views), "
                         f"{n} samples in {dt:.1f} s; views/s {len(sub) / dt:.3f}",
               "views_per_s": len(sub) / dt}
        log(f"cpu baseline: {cpu}")

    f32 = results.get("f32") or next(iter(results.values()))
    achieved = f32["gather_GBps"]
    line = {
        "metric": "Gsamples/s & views/s at 256³/360 views vs plain FP32; served p50/p99 latency",
        "value": f32["gsamples_s"], "unit": "Gsamples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": f32["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": w.name, "volume": "256^3 x 0.5 mm, 20 seeded ellipsoids",
                   "views": 360, "detector": "1024x768 @ 0.4 mm", "ray_step_vox": 0.2,
                   "parallelism": "single GPU"},
        "views_per_s": f32["views_s"],
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBPS, "traffic": None,
                     "algorithmic_bytes_per_sample": 32, "samples_per_launch": f32["n_samples"],
                     "kernel_ms": f32["kernel_ms"],
                     "l1_gather_peak": L1_PEAK_GBPS, "l1_gather_frac": achieved / L1_PEAK_GBPS},
        "cpu_baseline": cpu,
        "variants": results,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
