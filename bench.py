#!/usr/bin/env python
"""Plain VR throughput of the render path on one MI355X.

Workload: the body-sized volume, 512x512x1559 int16 (spacing 0.7422 x 0.7422
x 1.0 mm, blob chambers + 40 spline vessels), with a 4-label uint8 ellipsoid
mask (each label its own 100-entry TF LUT, window and alpha cap; label 0 alpha
0.4), rendered as 1024x1024 VR frames over 12 views rotating about z. Volume,
mask and LUTs are resident in HBM before timing starts and every frame is
written to HBM. One step = one frame.

The JSON line carries:
  * value / ms_per_step: frames per second over K back-to-back frames;
  * roofline: algorithmic bytes per frame (2 B per distinct voxel the
    reference rule reads, incl. gradient taps, + 1 B per mask voxel read
    + 3 B per output pixel) over the VR kernel's mean duration from HIP events
    on the launch stream, against the 8 TB/s HBM peak;
  * parity: a band of rows of one frame against the FP64 oracle;
  * cpu_baseline: the oracle's float instance (8-bit filter weights) on a
    band of rows of the same frame, converted to frames per second.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_scene():
    from monkeygl_amd import params as P
    from monkeygl_amd import workloads as W

    spec = W.BODY
    t = time.time()
    vol = W.make_volume(spec)
    mask = W.ellipsoid_mask(spec.dims, 4, seed=1)
    log(f"generated {spec.name} + mask in {time.time() - t:.1f}s")
    luts, labels = {}, {}
    for l in range(5):
        luts[l] = P.transfer_function(W.tf_points(100 + l))
        labels[l] = P.LabelState(alpha=0.4 if l == 0 else 1.0, ww=1200.0 + 200.0 * l, wl=-200.0 + 60.0 * l)
    return spec, vol, mask, luts, labels


def view_params(spec, labels, size: int, n_views: int):
    from monkeygl_amd import params as P

    ps = []
    for k in range(n_views):
        info = P.RenderInfo()
        info.set_vr_size(size, size)
        info.rotate(360.0 * k / n_views, 0.0)  # the x angle drives Rz
        ps.append(P.render_params(spec.dims, spec.spacing, info, labels, width=size, height=size))
    return ps


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=48)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--views", type=int, default=12)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-footprint", action="store_true")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="after the timed frames, copy this many MiB device-to-device once (a known-size "
                         "kernel against which FETCH_SIZE/WRITE_SIZE counters are scaled)")
    ap.add_argument("--traffic", type=str, default=str(ROOT / "profiles" / "traffic_vr_body1024.json"),
                    help="measured HBM bytes per frame from rocprof counters (see the file), if present")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build measures one GPU; run with --gpus 1 (no multi-GPU path)"}))
        return 2

    import torch

    from monkeygl_amd._capi import Context
    from oracle import oracle as O

    spec, vol, mask, luts, labels = build_scene()
    ctx = Context(0)
    t = time.time()
    ctx.set_volume(vol)
    ctx.set_mask(mask)
    for l, a in luts.items():
        ctx.set_lut(l, a)
    ctx.synchronize()
    log(f"uploaded and bricked in {time.time() - t:.1f}s")

    W_, H_ = args.size, args.size
    views = view_params(spec, labels, args.size, args.views)
    out = torch.empty(W_ * H_ * 3, dtype=torch.uint8, device="cuda")
    stream = torch.cuda.Stream()
    sh = stream.cuda_stream

    # warmup (builds the skip tables and step lattice once)
    with torch.cuda.stream(stream):
        for i in range(args.warmup):
            ctx.render_async(views[i % len(views)], out, stream=sh)
    torch.cuda.synchronize()

    K = args.steps
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(K + 1)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    with torch.cuda.stream(stream):
        evs[0].record(stream)
        for i in range(K):
            ctx.render_async(views[i % len(views)], out, stream=sh)
            evs[i + 1].record(stream)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    per = [evs[i].elapsed_time(evs[i + 1]) for i in range(K)]
    kernel_ms = float(np.mean(per))
    fps = K / elapsed
    log(f"{K} frames in {elapsed * 1e3:.1f} ms -> {fps:.1f} fps; kernel mean {kernel_ms:.3f} ms "
        f"(min {min(per):.3f}, max {max(per):.3f})")

    if args.calibrate_copy_mib:
        n = args.calibrate_copy_mib << 20
        a = torch.empty(n, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        log(f"calibration copy of {n} bytes done")
        del a, b

    # algorithmic bytes per frame, on the exact traversal of every view
    bytes_per_view = []
    if not args.no_footprint:
        tv = torch.zeros(vol.size, dtype=torch.uint8, device="cuda")
        tm = torch.zeros(vol.size, dtype=torch.uint8, device="cuda")
        scratch = torch.empty_like(out)
        for p in views:
            tv.zero_()
            tm.zero_()
            ctx.render_footprint(p, scratch, tv, tm, stream=torch.cuda.current_stream().cuda_stream)
            torch.cuda.synchronize()
            V = int(tv.sum(dtype=torch.int64).item())
            M = int(tm.sum(dtype=torch.int64).item())
            bytes_per_view.append(2 * V + M + 3 * W_ * H_)
        del tv, tm
        torch.cuda.synchronize()
    b_frame = float(np.mean([bytes_per_view[i % len(views)] for i in range(K)])) if bytes_per_view else None
    achieved = b_frame / (kernel_ms * 1e-3) / 1e9 if b_frame else None
    traffic = None
    tp = Path(args.traffic)
    if tp.exists() and args.size == 1024:
        traffic = json.loads(tp.read_text()).get("hbm_bytes_per_frame")

    # parity band and CPU baseline on view 1 (an oblique one)
    p = views[1 % len(views)]
    ctx.render_async(p, out, stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    gpu_img = out.cpu().numpy().reshape(H_, W_, 3)
    ncores = min(len(os.sched_getaffinity(0)), 16)
    O.set_threads(ncores)
    parity = None
    cpu = None
    if not args.no_cpu:
        # rows through the middle of the frame: the volume covers them fully
        band = 16
        y0 = H_ // 2 - band // 2
        t = time.perf_counter()
        ref = O.render_rows(p, vol, luts, y0, y0 + band, mask_zyx=mask, precision="f64")
        t_ref = time.perf_counter() - t
        d = np.abs(gpu_img[y0:y0 + band].astype(int) - ref.astype(int)).max(-1)
        parity = {"oracle": "fp64", "rows": band, "frac_within_1_255": float((d <= 1).mean()),
                  "max_diff": int(d.max()), "oracle_s": round(t_ref, 2)}
        log(f"parity on {band} rows: {parity}")
        # CPU baseline: float instance with 8-bit weights over whole frames of the same views
        # (bands of 64 rows) until the time budget is spent
        rows_done, t_cpu, nviews = 0, 0.0, 0
        band_rows = 64
        for q in views:
            for y in range(0, H_, band_rows):
                t = time.perf_counter()
                O.render_rows(q, vol, luts, y, min(H_, y + band_rows), mask_zyx=mask, precision="f32", weights8=True)
                t_cpu += time.perf_counter() - t
                rows_done += min(H_, y + band_rows) - y
                if t_cpu >= args.cpu_seconds:
                    break
            nviews += 1
            if t_cpu >= args.cpu_seconds:
                break
        cpu_fps = (rows_done / H_) / t_cpu
        cpu = {"value": round(cpu_fps, 4), "unit": "frames/s", "cores": ncores, "kind": "port",
               "sample": f"{rows_done / H_:.2f} frames ({rows_done} rows of {W_}x{H_} over {nviews} views), float "
                         f"oracle with 8-bit filter weights, {t_cpu:.1f}s on {ncores} threads"}
        log(f"cpu baseline: {cpu}")

    line = {
        "metric": "VR fps 512^2x1559 i16 -> 1024^2 (plain), 4-label mask",
        "value": round(fps, 3),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed * 1e3 / K, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int16->fp32",
        "data": "synthetic",
        "config": {"workload": f"body {spec.dims[0]}x{spec.dims[1]}x{spec.dims[2]} int16 + 4-label mask, "
                               f"VR {W_}x{H_}, {args.views} views about z",
                   "volume": list(spec.dims), "spacing": list(spec.spacing), "image": [W_, H_],
                   "views": args.views, "parallelism": "none (1 GPU)"},
        "kernel_ms": round(kernel_ms, 4),
        "roofline": None if achieved is None else {
            "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
            "bytes_per_frame": int(b_frame), "kernel": "k_vr"},
        "parity": parity,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
