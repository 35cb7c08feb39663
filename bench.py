#!/usr/bin/env python
"""Plain VR throughput and served latency of the render path on one MI355X.

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
    + 3 B per output pixel) over the VR frame's mean duration from HIP events
    on the launch stream (both passes: k_vr, then k_vr_tail for the rays it
    queued), against the 8 TB/s HBM peak;
  * parity: one whole frame against the FP64 oracle (pixels within 1/255,
    pixels beyond, their maximum difference and cause);
  * cpu_baseline: the oracle compiled for the host it runs on (-march=native),
    float instance with 8-bit filter weights, on whole frames of the same
    views for --cpu-seconds, in frames per second; the FP64 oracle timed
    beside it; threads from OMP_NUM_THREADS when set, with nproc recorded;
  * fit: transfer-function fitting at its configuration (256^3, 8 views at
    512^2, 256-entry LUT + opacity scale): Adam steps/s, forward and adjoint
    ms per view, the loss reduction and the adjoint against FP64 finite
    differences;
  * served: the render-serving mix over three resident volumes (the
    128^3 blob set, the 512x512x361 cardiac-sized set and the body set above
    with its mask): Poisson arrivals at --serve-qps, 60% VR / 20% MIP /
    15% MPR / 5% CPR, viewports uniform over 256^2..1024^2, random cameras;
    p50/p99 latency (arrival to result in host memory) and frames/s of the
    continuous-batching server next to one-at-a-time plain rendering of the
    same request stream, the same two at saturation (every request queued at
    once), and the count of served images bit-identical to the plain ones,
    then the same stream served as base64 PNG text (served.encoded);
  * encode: the headline frame to PNG + base64 on the device, next to the
    reference's fpng on one host core;
  * series: a GetBatchData series of parallel MIP slabs through the batcher
    against one plane launch per image;
  * mask_ingest: an object-mask edit merged on the device against the
    reference's host merge + full mask upload;
  * orientation_figure: the VR overlay's figure merged on the device against
    the reference's host rasteriser.
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


def fit_section(args) -> dict:
    """Transfer-function fitting at its configuration (256^3 16-blob volume,
    8 views at 512^2, 256-entry RGBA LUT + opacity scale, MSE to a seeded
    ground truth): Adam steps per second through the hand-written adjoint,
    the forward and adjoint times per 512^2 view from HIP events, the loss
    reduction, and the adjoint against FP64 central differences (h = 1e-4) on
    64 live LUT entries and the scale, with the loss over rows 252..259 of
    every 512^2 view."""
    import torch

    from monkeygl_amd import workloads as W
    from monkeygl_amd._capi import Context
    from monkeygl_amd.tf_fit import fit, render_tf
    from oracle import oracle as O
    from oracle.fit_check import fd_check

    spec = W.TF_FIT
    vol = W.make_volume(spec)
    ctx = Context(0)
    try:
        ctx.set_volume(vol)
        ctx.synchronize()
        n_lut, size = W.FIT_LUT_ENTRIES, 512
        vs = W.fit_views(size)
        lut_gt, lut0 = W.fit_lut(1), W.fit_lut(2)
        with torch.no_grad():
            target = render_tf(ctx, vs, torch.tensor(lut_gt, device="cuda"), torch.tensor(W.FIT_SCALE_TRUE))
        fit(ctx, vs, target, torch.tensor(lut0, device="cuda"), W.FIT_SCALE_START, steps=2)   # warm
        torch.cuda.synchronize()
        t = time.perf_counter()
        _, s_f, losses = fit(ctx, vs, target, torch.tensor(lut0, device="cuda"), W.FIT_SCALE_START,
                             steps=args.fit_steps, lr=0.01)
        torch.cuda.synchronize()
        el = time.perf_counter() - t
        log(f"fit: {args.fit_steps} Adam steps in {el:.2f}s; loss {losses[0]:.4g} -> {losses[-1]:.4g}")

        # forward and adjoint per view, HIP events on the launch stream
        stream = torch.cuda.Stream()
        sh = stream.cuda_stream
        dl = torch.tensor(lut0, device="cuda")
        rgb = torch.empty((len(vs), size, size, 3), dtype=torch.float32, device="cuda")
        rgb64 = torch.empty((len(vs), size, size, 3), dtype=torch.float64, device="cuda")
        g = torch.randn((len(vs), size, size, 3), generator=torch.Generator(device="cuda").manual_seed(3),
                        device="cuda")
        g_lut = torch.zeros(n_lut * 4, dtype=torch.float64, device="cuda")
        g_s = torch.zeros(1, dtype=torch.float64, device="cuda")
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        reps = 3
        torch.cuda.synchronize()
        ev[0].record(stream)
        for _ in range(reps):
            for i, p in enumerate(vs):
                ctx.tf_forward(p, dl, n_lut, W.FIT_SCALE_START, rgb[i], stream=sh, dev_rgb64=rgb64[i])
        ev[1].record(stream)
        for _ in range(reps):
            for i, p in enumerate(vs):
                ctx.tf_backward(p, dl, n_lut, W.FIT_SCALE_START, g[i], g_lut, g_s, stream=sh, dev_rgb64=rgb64[i])
        ev[2].record(stream)
        torch.cuda.synchronize()
        fwd_ms = ev[0].elapsed_time(ev[1]) / (reps * len(vs))
        bwd_ms = ev[1].elapsed_time(ev[2]) / (reps * len(vs))
        del rgb, rgb64, g
        log(f"fit: forward {fwd_ms:.3f} ms, adjoint {bwd_ms:.3f} ms per {size}^2 view")

        fd = None
        if not args.no_cpu:
            O.set_threads(cpu_threads()[0])
            t = time.perf_counter()
            r = fd_check(ctx, vol, vs, lut_gt, lut0, W.FIT_SCALE_TRUE, W.FIT_SCALE_START, rows=(252, 260))
            fd = {"entries": r["checked"], "compared": r["big"], "live_entries": r["live"], "h": 1e-4,
                  "image": [size, size], "views": len(vs), "loss_rows": [252, 260],
                  "max_rel_err": r["max_rel_err"], "scale_rel_err": r["scale_rel_err"], "bound": 1e-3,
                  "oracle_s": round(time.perf_counter() - t, 1)}
            log(f"fit: FD check {fd}")
        return {"steps_per_s": round(args.fit_steps / el, 3), "steps": args.fit_steps,
                "ms_per_step": round(el * 1e3 / args.fit_steps, 3), "forward_ms_per_view": round(fwd_ms, 4),
                "adjoint_ms_per_view": round(bwd_ms, 4), "views": len(vs), "image": [size, size],
                "volume": list(spec.dims), "lut_entries": n_lut, "loss_first": losses[0], "loss_last": losses[-1],
                "loss_ratio": losses[-1] / losses[0], "scale_final": s_f, "fd_check": fd}
    finally:
        ctx.close()


def fit_main(args) -> int:
    """--fit: the fitting block alone, as its own JSON line."""
    f = fit_section(args)
    line = {
        "metric": "TF fitting Adam steps/s (adjoint), 256^3 i16, 8 views 512^2, 256-entry LUT + opacity scale",
        "value": f["steps_per_s"], "unit": "steps/s", "n_gpus": 1, "steps": f["steps"], "warmup": 2,
        "ms_per_step": f["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 render, fp64 gradient accumulation", "data": "synthetic",
        "config": {"workload": "TF fitting: 256^3 16-blob volume, 8 views at 512^2, MSE to a seeded ground-truth TF",
                   "volume": f["volume"], "views": f["views"], "image": f["image"], "lut_entries": f["lut_entries"]},
        "fit": f,
    }
    print(json.dumps(line), flush=True)
    return 0


def cpu_threads():
    """(threads, note): OpenMP threads for the CPU legs. OMP_NUM_THREADS when
    set (the GPU box sets it to its CPU share for one GPU), else every core
    this process may run on."""
    aff = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env), f"OMP_NUM_THREADS={env}"
    return aff, "all cores in the affinity mask"


def series_section(ctx, spec, n_images: int = 256, size: int = 512):
    """GetBatchData: a series of n_images parallel 3 mm MIP slabs of size^2 on
    the body volume (an oblique frame through the z axis, 1 mm apart), through
    the batcher against one mk_render_plane_async launch per image."""
    import torch

    from monkeygl_amd import mpr as M
    from monkeygl_amd._capi import MK_MPR_MIP
    from monkeygl_amd.series import SeriesRenderer

    h, v = M.batch_direction_3d((1.0, 0.0, 0.0), (0.0, 1.0, 0.0), 30.0)
    ps = min(spec.spacing)
    c = tuple(0.5 * d * s for d, s in zip(spec.dims, spec.spacing))
    info = M.BatchInfo(h, v, c, size * ps, size * ps, ps, 3.0, 1.0, n_images, MK_MPR_MIP)
    params = M.batch_plane_params(info, spec.spacing)
    w, hgt = params[0].width, params[0].height
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    a = torch.empty(n_images * w * hgt, dtype=torch.int16, device="cuda")
    b = torch.empty_like(a)
    sr = SeriesRenderer(ctx)
    try:
        sr.render_async(params, a, stream=sh)   # warm
        for k, p in enumerate(params):
            ctx.render_plane_async(p, b[k * w * hgt:(k + 1) * w * hgt], stream=sh)
        torch.cuda.synchronize()
        e = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        e[0].record(stream)
        sr.render_async(params, a, stream=sh)
        e[1].record(stream)
        e[2].record(stream)
        for k, p in enumerate(params):
            ctx.render_plane_async(p, b[k * w * hgt:(k + 1) * w * hgt], stream=sh)
        e[3].record(stream)
        torch.cuda.synchronize()
    finally:
        sr.close()
    same = int((a.view(n_images, -1) == b.view(n_images, -1)).all(dim=1).sum().item())
    tb, ts = e[0].elapsed_time(e[1]), e[2].elapsed_time(e[3])
    out = {"images": n_images, "image": [w, hgt], "slab_mm": 3.0, "slice_distance_mm": 1.0, "mpr_type": "MIP",
           "batched_ms": round(tb, 3), "batched_images_per_s": round(n_images / tb * 1e3, 1),
           "per_image_launches_ms": round(ts, 3), "per_image_launches_images_per_s": round(n_images / ts * 1e3, 1),
           "identical": f"{same}/{n_images}"}
    log(f"batch series: {out}")
    return out


def mask_ingest_section(spec, vol, mask):
    """An object-mask edit on the body volume with its 4-label mask: the device
    merge (mk_merge_object_mask from host memory, and from a device copy)
    against the reference's path, a host merge of the whole mask followed by a
    full mask upload (VolumeInfo::AddNewObjectMask + CopyMaskData)."""
    import torch

    from monkeygl_amd import workloads as W
    from monkeygl_amd._capi import MK_MASK_ADD, Context

    obj = (W.ellipsoid_mask(spec.dims, 1, seed=9) > 0).astype(np.uint8)
    c = Context(0)
    try:
        c.set_volume(vol)
        c.set_mask(mask)
        t = time.perf_counter()
        c.merge_mask(obj, 5, MK_MASK_ADD)
        t_host = time.perf_counter() - t
        merged = c.get_mask()
        dev = torch.from_numpy(obj).cuda()
        torch.cuda.synchronize()
        c.set_mask(mask)
        t = time.perf_counter()
        c.merge_mask_device(dev, obj.shape, 5, MK_MASK_ADD)
        t_dev = time.perf_counter() - t
        # the reference's path: merge on the host, upload the whole mask
        t = time.perf_counter()
        m = mask.copy()
        m[obj > 0] = 5
        c.set_mask(m)
        t_ref = time.perf_counter() - t
        same = bool(np.array_equal(merged, m))
    finally:
        c.close()
    out = {"voxels": int(obj.size), "object_voxels": int(obj.sum()), "device_merge_from_host_ms": round(t_host * 1e3, 2),
           "device_merge_from_device_ms": round(t_dev * 1e3, 2), "host_merge_and_upload_ms": round(t_ref * 1e3, 2),
           "same_mask": same}
    log(f"mask ingest: {out}")
    return out


def overlay_section(ctx, size: int, with_cpu: bool):
    """The orientation figure (a 35,124-facet synthetic figure, the size of the
    reference's resources/human.stl) merged into a size^2 VR frame on the
    device, against the reference's host rasteriser (MergeOrientationBox over
    Methods::FillHoleInImage_Ch1, oracle/_ref) on the same frame."""
    import ctypes as C
    import struct

    import torch

    from monkeygl_amd import params as P
    from monkeygl_amd import workloads as W

    stl = W.stl_bytes(W.figure(0, fine=True))
    n = ctx.set_orientation_mesh(stl)
    info = P.RenderInfo()
    info.rotate(30.0, 20.0)
    rot = np.asarray(info.rot, np.float32)
    ws = torch.empty(ctx.orientation_workspace_bytes(), dtype=torch.uint8, device="cuda")
    frame = np.random.default_rng(5).integers(0, 256, (size, size, 3), dtype=np.uint8)
    dev = torch.tensor(frame, device="cuda")
    stream = torch.cuda.current_stream()
    for _ in range(2):
        ctx.merge_orientation_box_async(rot, dev, size, size, ws, stream=stream.cuda_stream)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(20):
        ctx.merge_orientation_box_async(rot, dev, size, size, ws, stream=stream.cuda_stream)
    e1.record(stream)
    torch.cuda.synchronize()
    out = {"facets": n, "image": [size, size], "gpu_ms": round(e0.elapsed_time(e1) / 20, 4)}
    ref_lib = ROOT / "oracle" / "_ref" / "libmkref.so"
    if with_cpu and ref_lib.exists():
        L = C.CDLL(str(ref_lib))
        L.ref_stl_normalize.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.ref_orientation_box.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int]
        body = np.frombuffer(stl, np.uint8, 50 * n, 84).copy()
        fac = np.zeros(n * 9, np.float32)
        L.ref_stl_normalize(body.ctypes.data, n, fac.ctypes.data)
        ref = frame.copy()
        t = time.perf_counter()
        for _ in range(5):
            L.ref_orientation_box(fac.ctypes.data, n, rot.ctypes.data, ref.ctypes.data, size, size)
        out["reference_cpu_ms"] = round((time.perf_counter() - t) / 5 * 1e3, 3)
        out["reference_cpu_cores"] = 1
        out["identical"] = bool(np.array_equal(dev.cpu().numpy(), ref))
    ctx.set_orientation_mesh(None)
    log(f"orientation figure: {out}")
    return out


def configs_section(scene2, body_ctx, spec, vr_trip_budget=None) -> dict:
    """One timing per remaining workload of the configurations, each on its own
    plain path with inputs resident: the 128^3 blob set (256^2 VR with the
    4096-entry LUT, half-voxel step and exit at 0.99; 256^2 MIP), the
    cardiac-sized set (1024^2 VR and MIP over 12 views about z; axial, sagittal
    and coronal MPR at 512^2, nearest and trilinear), and on the body set the
    stretched and straightened CPR at 512 px width along the 200-point
    centreline and marching cubes at iso 200 on a 256^3 sub-block. Mean time
    per item from HIP events over repeated back-to-back launches."""
    import torch

    from monkeygl_amd import cpr as G
    from monkeygl_amd import params as P
    from monkeygl_amd import workloads as W
    from monkeygl_amd._capi import (MK_MPR_AVERAGE, MK_RENDER_MIP, MK_RENDER_VR, MK_SAMPLE_LINEAR,
                                    MK_SAMPLE_NEAREST, Context)

    stream = torch.cuda.Stream()
    sh = stream.cuda_stream

    def timed(launch, n_items, reps):
        for i in range(n_items):
            launch(i)
        stream.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            for i in range(n_items):
                launch(i)
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_time(e1) / (reps * n_items)

    def views(sv, labels, rt, size, n=12, **kw):
        out = []
        for k in range(n):
            info = P.RenderInfo()
            info.set_vr_size(size, size)
            info.rotate(360.0 * k / n, 0.0)
            out.append(P.render_params(sv.spec.dims, sv.spec.spacing, info, labels, rt, width=size, height=size, **kw))
        return out

    res = {}
    small, cardiac = scene2[0], scene2[1]
    frame = torch.empty(1024 * 1024 * 3, dtype=torch.uint8, device="cuda")

    cs = Context(0)
    cc = Context(0)
    if vr_trip_budget is not None:
        cs.set_vr_trip_budget(vr_trip_budget)
        cc.set_vr_trip_budget(vr_trip_budget)
    cs.set_volume(small.vol)
    pts = W.tf_points(0)
    cs.set_lut(0, P.transfer_function({int(round(k * 4095 / 99)): v for k, v in pts.items()}, n=4096))
    lab = {0: P.LabelState(1.0, 2000.0, 0.0)}
    for rt, name, kw in ((MK_RENDER_VR, "blobs128_vr_256", dict(step_scale=0.5, exit_alpha=0.99)),
                         (MK_RENDER_MIP, "blobs128_mip_256", {})):
        vs = views(small, lab, rt, 256, **kw)
        ms = timed(lambda i: cs.render_async(vs[i], frame, stream=sh), len(vs), 8)
        res[name] = {"ms": round(ms, 4), "frames_per_s": round(1e3 / ms, 1)}

    cc.set_volume(cardiac.vol)
    cc.set_lut(0, P.transfer_function(W.tf_points(1)))
    lab = {0: P.LabelState(1.0, 900.0, 250.0)}
    for rt, name in ((MK_RENDER_VR, "cardiac_vr_1024"), (MK_RENDER_MIP, "cardiac_mip_1024")):
        vs = views(cardiac, lab, rt, 1024)
        ms = timed(lambda i: cc.render_async(vs[i], frame, stream=sh), len(vs), 2)
        res[name] = {"ms": round(ms, 4), "frames_per_s": round(1e3 / ms, 1)}
    img = torch.empty(512 * 512, dtype=torch.int16, device="cuda")
    for sampling, name in ((MK_SAMPLE_NEAREST, "cardiac_mpr_512_nearest"), (MK_SAMPLE_LINEAR, "cardiac_mpr_512_linear")):
        qs = [P.plane_params(cardiac.spec.dims, cardiac.spec.spacing, plane, mpr_type=MK_MPR_AVERAGE,
                             sampling=sampling, size=(512, 512))
              for plane in (P.PLANE_AXIAL, P.PLANE_SAGITTAL, P.PLANE_CORONAL)]
        ms = timed(lambda i: cc.render_plane_async(qs[i], img, stream=sh), len(qs), 20)
        res[name] = {"ms": round(ms, 4), "images_per_s": round(1e3 / ms, 1)}

    info = G.CPRInfo()
    info.set_spacing(spec.spacing)
    info.set_cpr_line_voxel(W.centerline(spec.dims, 200, seed=4))
    for which, name in ((G.PLANE_STRETCHED_CPR, "body_cpr_stretched_512"),
                        (G.PLANE_STRAIGHTENED_CPR, "body_cpr_straightened_512")):
        pts_, dirs_, _, h = info.get_cpr_info(which)
        dp = torch.from_numpy(np.ascontiguousarray(pts_, dtype=np.float64).reshape(-1)).cuda()
        dd = torch.from_numpy(np.ascontiguousarray(dirs_, dtype=np.float64).reshape(-1)).cuda()
        out = torch.empty(h * 512, dtype=torch.int16, device="cuda")
        ms = timed(lambda i: body_ctx.render_cpr_async(512, h, dp, dd, False, MK_SAMPLE_LINEAR, out, stream=sh), 1, 20)
        res[name] = {"ms": round(ms, 4), "rows": int(h), "images_per_s": round(1e3 / ms, 1)}

    box = (128, 128, 651, 384, 384, 907)
    n_tri = len(body_ctx.marching_cubes(box, 200.0))
    tri = torch.empty(max(n_tri, 1) * 9, dtype=torch.float32, device="cuda")
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    ms = timed(lambda i: body_ctx.marching_cubes_async(box, 200.0, tri, max(n_tri, 1), cnt, stream=sh), 1, 10)
    res["body_mc_256_iso200"] = {"ms": round(ms, 4), "triangles": n_tri, "box": list(box)}
    cs.close()
    cc.close()
    return res


def serve_section(args, body_ctx, spec, vol, mask, luts, labels, scene2=None):
    """Served vs sequential plain rendering of one seeded request stream."""
    import dataclasses

    from monkeygl_amd import workloads as W
    from monkeygl_amd._capi import Context
    from monkeygl_amd.serve import RenderServer, latency_stats, run_sequential

    t = time.time()
    scene = list(scene2) if scene2 is not None else W.serving_scene([(W.SMALL, 0), (W.CARDIAC, 0)])
    scene.append(W.ServedVolume(spec, vol, mask, luts, labels, W.centerline(spec.dims, 200, seed=4)))
    ctxs = []
    for sv in scene[:2]:
        c = Context(0)
        if args.vr_trip_budget is not None:
            c.set_vr_trip_budget(args.vr_trip_budget)
        c.set_volume(sv.vol)
        for l, lut in sv.luts.items():
            c.set_lut(l, lut)
        c.synchronize()
        ctxs.append(c)
    ctxs.append(body_ctx)
    log(f"serving scene ready in {time.time() - t:.1f}s")

    def fresh(rs, keep_time=True):
        return [dataclasses.replace(r, t_arrival=(r.t_arrival if keep_time else 0.0), result=None, error=None,
                                    t_admit=0.0, t_done=0.0, solo=False) for r in rs]

    reqs = W.serving_requests(scene, args.serve_requests, args.serve_qps, seed=0)
    srv = RenderServer(ctxs, hbm_budget=512 << 20, nslots=args.serve_slots, max_items=64, max_tiles=65536)
    try:
        # warm both paths: tables, lattices, graphs, pinned buffers
        warm = fresh(reqs[:48], keep_time=False)
        srv.run(warm)
        run_sequential(ctxs, fresh(reqs[:48], keep_time=False))
        b0 = srv.batches
        served = fresh(reqs)
        srv.run(served)
        n_batches = srv.batches - b0
        sizes = srv.batch_sizes[b0:]
        seq = fresh(reqs)
        run_sequential(ctxs, seq)
        same, mism = 0, []
        for a, b in zip(served, seq):
            if a.error is None and np.array_equal(a.result, b.result):
                same += 1
            elif len(mism) < 8:
                d = (np.abs(a.result.astype(np.int32) - b.result.astype(np.int32)) if a.error is None
                     else None)
                mism.append({"kind": a.tag[1], "volume": a.volume, "size": list(a.size), "error": a.error,
                             "max_diff": None if d is None else int(d.max()),
                             "values_differing": None if d is None else int((d > 0).sum())})
        st_srv, st_seq = latency_stats(served), latency_stats(seq)
        log(f"served @ {args.serve_qps} qps: {st_srv}; sequential: {st_seq}; identical {same}/{len(reqs)}")
        # saturation: every request queued at t = 0
        sat_srv = fresh(reqs, keep_time=False)
        srv.run(sat_srv)
        sat_seq = fresh(reqs, keep_time=False)
        run_sequential(ctxs, sat_seq)
        sat = {"served_frames_per_s": latency_stats(sat_srv)["frames_per_s"],
               "sequential_frames_per_s": latency_stats(sat_seq)["frames_per_s"]}
        log(f"saturation: {sat}")
    finally:
        srv.close()
    # the same stream with the reference's wire format: base64 PNG per result
    srv_enc = RenderServer(ctxs, hbm_budget=1024 << 20, nslots=args.serve_slots, max_items=64, max_tiles=65536,
                           encode=True)
    try:
        srv_enc.run(fresh(reqs[:48], keep_time=False))
        enc_reqs = fresh(reqs)
        srv_enc.run(enc_reqs)
        st_enc = latency_stats(enc_reqs)
        text_bytes = sum(len(r.result) for r in enc_reqs if r.error is None)
        log(f"served with PNG+base64 @ {args.serve_qps} qps: {st_enc}; {text_bytes / max(1, st_enc['n']) / 1e3:.1f} "
            f"kB text per result")
    finally:
        srv_enc.close()
        for c in ctxs[:2]:
            c.close()
    rnd = lambda d: {k: (round(v, 3) if isinstance(v, float) else v) for k, v in d.items()}
    return {
        "qps": args.serve_qps, "requests": len(reqs), "mix": dict(W.SERVE_MIX),
        "viewports": list(W.SERVE_SIZES), "volumes": [sv.spec.name for sv in scene],
        "slots": args.serve_slots, "hbm_budget_bytes": 512 << 20,
        "p50_ms": round(st_srv["p50_ms"], 3), "p99_ms": round(st_srv["p99_ms"], 3),
        "frames_per_s": round(st_srv["frames_per_s"], 2),
        "sequential_plain": rnd(st_seq), "served_detail": rnd(st_srv),
        "batches": n_batches, "mean_batch": round(float(np.mean(sizes)), 2) if sizes else 0,
        "saturation": rnd(sat), "identical_to_plain": f"{same}/{len(reqs)}", "mismatches": mism,
        "encoded": {"format": "base64 PNG per result (RGB frames, int16 images as RGBA of width W/2)",
                    "p50_ms": round(st_enc["p50_ms"], 3), "p99_ms": round(st_enc["p99_ms"], 3),
                    "frames_per_s": round(st_enc["frames_per_s"], 2),
                    "mean_text_bytes": int(text_bytes / max(1, st_enc["n"]))},
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=48)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--views", type=int, default=12)
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="time budget of the float CPU baseline (whole frames of the views)")
    ap.add_argument("--cpu-seconds-fp64", type=float, default=8.0,
                    help="time budget of the FP64 oracle timed beside it")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-footprint", action="store_true")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="after the timed frames, copy this many MiB device-to-device once (a known-size "
                         "kernel against which FETCH_SIZE/WRITE_SIZE counters are scaled)")
    ap.add_argument("--fit", action="store_true",
                    help="measure only transfer-function fitting (256^3 blob volume, 8 views at 512^2, "
                         "256-entry LUT + opacity scale, Adam steps, FD check of the adjoint at 512^2)")
    ap.add_argument("--no-fit", action="store_true", help="skip the fitting block of the default line")
    ap.add_argument("--fit-steps", type=int, default=200)
    ap.add_argument("--serve-requests", type=int, default=600)
    ap.add_argument("--serve-qps", type=float, default=500.0)
    ap.add_argument("--serve-slots", type=int, default=2)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the timings of the other configurations' workloads")
    ap.add_argument("--vr-trip-budget", type=int, default=None,
                    help="first-pass trip bound of plain VR frames for every context (0: one pass; default: the "
                         "library's); the frames are the same for every value")
    ap.add_argument("--traffic", type=str, default=str(ROOT / "profiles" / "traffic_vr_body1024.json"),
                    help="measured HBM bytes per frame from rocprof counters (see the file), if present")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build measures one GPU; run with --gpus 1 (no multi-GPU path)"}))
        return 2

    if args.fit:
        return fit_main(args)

    import torch

    from monkeygl_amd._capi import Context
    from oracle import oracle as O

    spec, vol, mask, luts, labels = build_scene()
    ctx = Context(0)
    if args.vr_trip_budget is not None:
        ctx.set_vr_trip_budget(args.vr_trip_budget)
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

    # PCIe-inclusive: the synchronous entry point writing caller-owned host
    # memory, as cu_render does (never reported as value)
    host_frames = max(args.views, 12)
    t = time.perf_counter()
    for i in range(host_frames):
        ctx.render(views[i % len(views)])
    host_fps = host_frames / (time.perf_counter() - t)
    log(f"synchronous host-output rate: {host_fps:.1f} fps")

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
    # achievable HBM bandwidth on this box: a 1 GiB device-to-device copy
    # (read + write bytes over the copy's event time), beside the spec peak
    cp_a = torch.empty(1 << 30, dtype=torch.uint8, device="cuda")
    cp_b = torch.empty_like(cp_a)
    cp_b.copy_(cp_a)
    cs = torch.cuda.current_stream()   # the stream the copies run on
    c0, c1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    c0.record(cs)
    for _ in range(5):
        cp_b.copy_(cp_a)
    c1.record(cs)
    torch.cuda.synchronize()
    copy_gbs = 2.0 * cp_a.numel() * 5 / (c0.elapsed_time(c1) * 1e-3) / 1e9
    del cp_a, cp_b
    log(f"device copy: {copy_gbs:.0f} GB/s achievable")

    # device PNG + base64 of the headline frame (the reference encodes on the host)
    from monkeygl_amd.encode import Encoder
    enc = Encoder(W_, H_, 3)
    ctx.render_async(views[1 % len(views)], out, stream=sh)
    for _ in range(3):
        enc.encode_async(out, W_, H_, 3, stream=sh)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(10):
        enc.encode_async(out, W_, H_, 3, stream=sh)
    e1.record(stream)
    torch.cuda.synchronize()
    png_n, b64_n = enc.lengths()
    encode = {"gpu_png_base64_ms": round(e0.elapsed_time(e1) / 10, 4), "png_bytes": png_n, "base64_bytes": b64_n,
              "image": [W_, H_, 3]}
    log(f"device PNG+base64 of a {W_}x{H_} frame: {encode}")
    del enc

    # parity (one whole frame) and CPU baseline on view 1 (an oblique one)
    p = views[1 % len(views)]
    ctx.render_async(p, out, stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    gpu_img = out.cpu().numpy().reshape(H_, W_, 3)
    threads, thread_note = cpu_threads()
    O.set_threads(threads)
    parity = None
    cpu = None
    if not args.no_cpu:
        t = time.perf_counter()
        ref = O.render(p, vol, luts, mask_zyx=mask, precision="f64")
        t_ref = time.perf_counter() - t
        d = np.abs(gpu_img.astype(int) - ref.astype(int)).max(-1)
        parity = {"oracle": "fp64", "image": [W_, H_], "view": 1 % len(views),
                  "frac_within_1_255": float((d <= 1).mean()), "pixels_beyond_1_255": int((d > 1).sum()),
                  "max_diff": int(d.max()), "oracle_s": round(t_ref, 2),
                  "beyond_cause": None if not (d > 1).any() else
                  "rays whose per-label opacity cap or 0.001 refinement threshold is decided differently by "
                  "float and double rounding of the accumulated values, so they composite one sample more or fewer"}
        log(f"parity on the whole frame: {parity}")
        # CPU baseline: the oracle compiled for this host (-O3 -march=native
        # -ffp-contract=off -fopenmp), whole frames of the same views, for
        # the float instance with 8-bit filter weights (the reference's
        # arithmetic) and for the FP64 instance, each until its time budget
        import tempfile
        tmpd = tempfile.mkdtemp(prefix="mk_oracle_")
        try:
            O.use_library(O.build_native(tmpd))
            build = "-O3 -march=native -ffp-contract=off -fopenmp, built on this host"
        except Exception as e:   # no compiler: the package's x86-64-v3 build
            build = f"-O3 -march=x86-64-v3 -ffp-contract=off -fopenmp (native build failed: {e})"
        O.set_threads(threads)

        def frames_per_s(precision, weights8, budget):
            rows, spent, nv = 0, 0.0, 0
            for q in views:
                for y in range(0, H_, 64):
                    t0 = time.perf_counter()
                    O.render_rows(q, vol, luts, y, min(H_, y + 64), mask_zyx=mask, precision=precision,
                                  weights8=weights8)
                    spent += time.perf_counter() - t0
                    rows += min(H_, y + 64) - y
                    if spent >= budget:
                        break
                nv += 1
                if spent >= budget:
                    break
            return rows / H_ / spent, rows, spent, nv

        f32, rows, spent, nv = frames_per_s("f32", True, args.cpu_seconds)
        f64, rows64, spent64, _ = frames_per_s("f64", False, args.cpu_seconds_fp64)
        O.use_library(None)
        O.set_threads(threads)
        cpu = {"value": round(f32, 4), "unit": "frames/s", "cores": threads, "kind": "port",
               "sample": f"{rows / H_:.2f} frames ({rows} rows of {W_}x{H_} over {nv} views), float oracle with "
                         f"8-bit filter weights, {spent:.1f}s on {threads} threads",
               "threads_from": thread_note, "nproc": os.cpu_count(), "affinity_cores": len(os.sched_getaffinity(0)),
               "omp_num_threads": os.environ.get("OMP_NUM_THREADS"), "build": build,
               "fp64_oracle": {"value": round(f64, 4), "unit": "frames/s",
                               "sample": f"{rows64 / H_:.2f} frames, {spent64:.1f}s on {threads} threads"}}
        log(f"cpu baseline: {cpu}")
        # the reference's own encoder (core/fpng/fpng.cpp, built into oracle/_ref) on the same frame
        ref_lib = ROOT / "oracle" / "_ref" / "libmkref.so"
        if ref_lib.exists():
            import ctypes as C
            L = C.CDLL(str(ref_lib))
            L.ref_fpng_encode.restype = C.c_ulonglong
            L.ref_fpng_encode.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_ulonglong]
            frame = np.ascontiguousarray(gpu_img)
            buf = np.zeros(frame.size * 2 + 4096, np.uint8)
            t = time.perf_counter()
            for _ in range(10):
                nb = L.ref_fpng_encode(frame.ctypes.data, W_, H_, 3, buf.ctypes.data, buf.size)
            encode["fpng_cpu_ms"] = round((time.perf_counter() - t) / 10 * 1e3, 4)
            encode["fpng_bytes"] = int(nb)
            encode["fpng_cores"] = 1
            log(f"fpng on the host: {encode}")

    fitb = None if args.no_fit else fit_section(args)
    series = series_section(ctx, spec)
    mask_ingest = mask_ingest_section(spec, vol, mask)
    overlay = overlay_section(ctx, W_, not args.no_cpu)
    scene2 = None
    if not (args.no_serve and args.no_configs):
        from monkeygl_amd import workloads as W
        t = time.time()
        scene2 = W.serving_scene([(W.SMALL, 0), (W.CARDIAC, 0)])
        log(f"blob and cardiac sets generated in {time.time() - t:.1f}s")
    configs = None if args.no_configs else configs_section(scene2, ctx, spec, args.vr_trip_budget)
    if configs:
        log(f"configs: {configs}")
    served = None if args.no_serve else serve_section(args, ctx, spec, vol, mask, luts, labels, scene2)

    line = {
        "metric": "VR fps 512\u00b2\u00d71559 i16\u21921024\u00b2 plain vs served p50/p99 @\u22641/255 err; HBM GB/s",
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
        "plain_host_output_fps": round(host_fps, 2),
        "roofline": None if achieved is None else {
            "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
            "bytes_per_frame": int(b_frame), "kernel": "k_vr + k_vr_tail", "achievable_copy_gbs": round(copy_gbs, 1),
            "frac_of_copy": round(achieved / copy_gbs, 5)},
        "parity": parity,
        "cpu_baseline": cpu,
        "encode": encode,
        "fit": fitb,
        "series": series,
        "mask_ingest": mask_ingest,
        "orientation_figure": overlay,
        "configs": configs,
        "served": served,
    }
    print(json.dumps(line), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
