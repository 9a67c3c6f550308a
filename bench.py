#!/usr/bin/env python3
"""Benchmark of the per-brick render-and-composite path on one MI355X.

One step = one frame (aloha_render_frame): every virtual rank ray-marches
its bricks into its float RGBA + RGBA8 + depth image, and the ranks' images
are 'over'-composited in kd visibility order into the frame, reading each
rank image only inside the pixel rectangle its bricks project to (outside
it the image is transparent black).  Two launches: the per-rank launch
(which also clears the frame outside the rectangles' bounding box) and the
composite of that box.  The same frame as two calls (render_ranks, then the
sparse composite of the whole image), the dense composite and the fused
one-launch form are timed beside it, each checked equal.  Inputs (bricks, TF, camera) are resident in
HBM before timing starts.

Default workload: 512^3 float32 fBm volume, chopSuey-split into 8 bricks
(2x2x2, ghost voxels), one brick per virtual rank, 1024x1024, dt = 0.5.

Prints one JSON line (metric = Gsamples/s; ms_per_step = ms/frame).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("Gsamples/s & ms/frame, 1024² DVR of 512³ 8-brick vol; "
          "fwd+adjoint vs fwd ms")
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="fbm512", choices=["fbm512", "blobs128"])
    ap.add_argument("--size", type=int, default=None, help="volume edge (default per workload)")
    ap.add_argument("--res", type=int, default=None, help="image edge (default per workload)")
    ap.add_argument("--ranks", type=int, default=None, help="virtual ranks (default: bricks)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="target CPU work of the baseline sample")
    ap.add_argument("--profile", action="store_true",
                    help="few frames, no CPU baseline (for rocprofv3 runs)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the differentiable form")
    ap.add_argument("--adam-steps", type=int, default=200)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched form")
    ap.add_argument("--no-surface", action="store_true", help="skip the triangle surface path")
    ap.add_argument("--only", choices=["adjoint", "batched", "surface"], default=None,
                    help="run only this section (for profiling its kernels)")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(json.dumps({"error": "this build measures one GPU: every virtual rank's bricks "
                          "live on a single MI355X; run with --gpus 1"}))
        sys.exit(2)
    if args.profile:
        args.no_cpu = True

    import torch
    import aloha_amd as A
    from aloha_amd import RankImages
    from aloha_amd import workloads as WL

    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU visible"}))
        sys.exit(1)
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    if args.only:
        return only_section(args)

    t0 = time.time()
    if args.workload == "fbm512":
        size, res = args.size or 512, args.res or 1024
        wl = WL.fbm512(num_ranks=args.ranks or 8, W=res, H=res, size=size)
        desc = f"fbm{size} f32 8 bricks (2x2x2, ghost voxels), {res}x{res}, dt 0.5"
    else:
        size, res = args.size or 128, args.res or 512
        wl = WL.blobs128(num_ranks=args.ranks or 2, W=res, H=res, size=size)
        desc = f"blobs{size} u8 2 bricks, {res}x{res}, dt 0.5"
    log(f"[bench] workload {desc} built in {time.time() - t0:.1f}s")
    P, W, H = wl.num_ranks, wl.W, wl.H

    ctx = A.Context(0, stream=stream)
    t_up = time.time()
    ids = [ctx.upload_brick(d, v) for d, v in zip(wl.bricks, wl.voxels)]
    ctx.sync()
    upload_s = time.time() - t_up
    ctx.set_tf(wl.tf)
    ctx.set_camera(wl.cam["org"], wl.cam["dir_00"], wl.cam["dir_du"], wl.cam["dir_dv"])
    ranks = [[ids[i] for i in r] for r in wl.rank_bricks]
    params = wl.params()
    imgs = RankImages.alloc(P, W, H, rgba=True, rgba8=True, depth=True)
    out = torch.empty((H, W, 4), dtype=torch.float32, device=dev)
    out8 = torch.empty((H, W), dtype=torch.int32, device=dev)
    seq = torch.empty_like(out)
    seq8 = torch.empty_like(out8)
    log(f"[bench] {len(ids)} bricks uploaded ({ctx.brick_bytes() / 1e6:.0f} MB padded) "
        f"in {upload_s:.2f}s incl. PCIe")

    # the pixel rectangles outside of which each rank image is transparent:
    # the composite reads the rank images only inside them (IceT's sparse
    # images); the dense composite is timed beside it and gives the same frame
    rects = ctx.rank_rects(params, W, H, ranks)

    def frame():
        ctx.render_frame(params, W, H, ranks, imgs, wl.order, out, out8)

    def two_calls():
        ctx.render_ranks(params, W, H, ranks, imgs)
        ctx.composite_over(imgs.rgba, wl.order, W, H, seq, seq8, rects=rects)

    # lanes-per-ray schedule of the per-rank and fused launches, chosen by
    # timing every choice on these exact launches (untimed; results are
    # identical for every schedule)
    fout = torch.empty_like(out)
    fout8 = torch.empty_like(out8)
    schedule, sched_table = ctx.autotune_march(params, W, H, ranks, imgs, wl.order, fout, fout8,
                                               reps=10)
    log(f"[bench] march schedule (lanes per ray: ranks, fused) = {schedule}")

    def timed(fn, n):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(n):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / n

    if args.profile:
        # calibration for HBM-traffic counters: one copy of exactly 256 MiB
        src = torch.empty(64 << 20, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    for _ in range(args.warmup):
        frame()
    torch.cuda.synchronize()
    ms_frame = timed(frame, args.steps)
    log(f"[bench] {args.steps} frames: {ms_frame:.4f} ms/frame")
    for _ in range(args.warmup):
        two_calls()
    ms_two = timed(two_calls, args.steps)
    torch.cuda.synchronize()
    two_equal = bool(torch.equal(seq, out) and torch.equal(seq8, out8))
    log(f"[bench] as two calls (render_ranks, composite): {ms_two:.4f} ms/frame (equal {two_equal})")

    # per-kernel averages over the same number of launches
    ms_render = timed(lambda: ctx.render_ranks(params, W, H, ranks, imgs), args.steps)
    ms_comp = timed(lambda: ctx.composite_over(imgs.rgba, wl.order, W, H, out, out8, rects=rects),
                    args.steps)
    dense = torch.empty_like(out)
    dense8 = torch.empty_like(out8)
    ms_dense = timed(lambda: ctx.composite_over(imgs.rgba, wl.order, W, H, dense, dense8),
                     args.steps)
    torch.cuda.synchronize()
    sparse_equal = bool(torch.equal(dense, out) and torch.equal(dense8, out8))
    log(f"[bench] render {ms_render:.4f} ms, composite {ms_comp:.4f} ms (sparse; dense "
        f"{ms_dense:.4f} ms, equal {sparse_equal})")
    # the reference's own data path: the same rank images, IceT's ordered blend
    # of their RGBA8 (ubyte colour format) -- 8-bit, so its own result
    out8b = torch.empty_like(out8)
    rgba8_frame = lambda: (ctx.render_ranks(params, W, H, ranks, imgs),
                           ctx.composite_over_rgba8(imgs.rgba8, wl.order, W, H, out8b))
    for _ in range(args.warmup):
        rgba8_frame()
    ms_rgba8 = timed(rgba8_frame, args.steps)
    log(f"[bench] RGBA8 composite form (render_ranks + 8-bit ordered blend): {ms_rgba8:.4f} ms/frame")
    # fused form: the same frame (bit-identical result) in one launch
    fused = lambda: ctx.render_composite(params, W, H, ranks, wl.order, fout, fout8)
    for _ in range(args.warmup):
        fused()
    ms_fused = timed(fused, args.steps)
    torch.cuda.synchronize()
    fused_equal = bool(torch.equal(fout, out) and torch.equal(fout8, out8))
    log(f"[bench] fused render+composite {ms_fused:.4f} ms (equal to plain: {fused_equal})")

    # exact sample count and compulsory voxel footprint of one frame (untimed)
    cnt = RankImages.alloc(P, W, H, rgba=False, rgba8=False, depth=False, nsamples=True)
    ctx.render_ranks(params, W, H, ranks, cnt)
    torch.cuda.synchronize()
    samples = int(cnt.nsamples.to(torch.int64).sum().item())
    touched = ctx.footprint(params, W, H, ranks)
    gsamples = samples / (ms_frame * 1e-3) / 1e9

    # algorithmic bytes per launch
    npix = W * H
    render_bytes = touched * 4 + wl.tf.nbytes + P * npix * (16 + 4 + 4)
    rect_pix = sum(max(0, x1 - x0) * max(0, y1 - y0) for x0, y0, x1, y1 in rects)
    comp_bytes = rect_pix * 16 + npix * (16 + 4)  # sparse: rank pixels inside the rectangles
    dense_bytes = P * npix * 16 + npix * (16 + 4)
    dense_gbs = dense_bytes / (ms_dense * 1e-3) / 1e9
    render_gbs = render_bytes / (ms_render * 1e-3) / 1e9
    comp_gbs = comp_bytes / (ms_comp * 1e-3) / 1e9
    traffic = {}
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath) and args.workload == "fbm512" and (args.size or 512) == 512:
        with open(tpath) as f:
            traffic = {k: v["read_bytes"] + v["write_bytes"] for k, v in json.load(f).items()
                       if not k.startswith("_")}
    dominant = "render" if ms_render >= ms_comp else "composite"
    roof = {
        "kernel": f"march_persist_kernel<{schedule[0]}>" if dominant == "render" else "over_sparse_kernel",
        "bound": "hbm",
        "achieved": round(render_gbs if dominant == "render" else comp_gbs, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round((render_gbs if dominant == "render" else comp_gbs) / HBM_PEAK_GBS, 4),
        "traffic": traffic.get(f"march_persist_kernel<{schedule[0]}>" if dominant == "render"
                               else "over_sparse_kernel"),
        "traffic_source": "profiles/traffic.json (rocprofv3 --pmc, calibrated)" if traffic else None,
        "algorithmic_bytes": int(render_bytes if dominant == "render" else comp_bytes),
        "launch_ms": round(ms_render if dominant == "render" else ms_comp, 5),
    }
    kernels = {
        f"march_persist_kernel<{schedule[0]}>": {"ms": round(ms_render, 5), "bytes": int(render_bytes),
                         "GB/s": round(render_gbs, 1), "frac_hbm": round(render_gbs / HBM_PEAK_GBS, 4),
                         "Gsamples/s": round(samples / (ms_render * 1e-3) / 1e9, 3)},
        "over_sparse_kernel": {"ms": round(ms_comp, 5), "bytes": int(comp_bytes),
                               "GB/s": round(comp_gbs, 1),
                               "frac_hbm": round(comp_gbs / HBM_PEAK_GBS, 4),
                               "equal_to_dense": sparse_equal},
        "over_kernel": {"ms": round(ms_dense, 5), "bytes": int(dense_bytes),
                        "GB/s": round(dense_gbs, 1), "frac_hbm": round(dense_gbs / HBM_PEAK_GBS, 4)},
    }

    cpu = None
    parity = None
    if not args.no_cpu:
        cpu, parity = cpu_baseline(wl, imgs, args.cpu_seconds)
    adjoint = None
    if not args.no_adjoint and not args.profile:
        adjoint = adjoint_bench(args, timed)
    batched = None
    if not args.no_batch and not args.profile:
        batched = batch_bench(args, timed)
    surface = None
    if not args.no_surface:
        surface = surface_bench(args, timed)

    res = {
        "metric": METRIC,
        "value": round(gsamples, 4),
        "unit": "Gsamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_frame, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded)",
        "schedule": {"lanes_per_ray_ranks": schedule[0], "lanes_per_ray_fused": schedule[1],
                     "autotune": [{"lanes": l, "ranks_ms": round(r, 4), "fused_ms": round(f, 4)}
                                  for l, r, f in sched_table]},
        "config": {"workload": desc, "virtual_ranks": P, "bricks": len(ids),
                   "samples_per_frame": samples, "touched_voxels": touched,
                   "brick_bytes_padded": ctx.brick_bytes(), "upload_s_incl_pcie": round(upload_s, 3)},
        "roofline": roof,
        "kernels": kernels,
        "forms": {
            "render_frame": {"ms_per_frame": round(ms_frame, 5), "Gsamples/s": round(gsamples, 4)},
            "render_ranks_then_sparse_composite": {
                "ms_per_frame": round(ms_two, 5),
                "Gsamples/s": round(samples / (ms_two * 1e-3) / 1e9, 4),
                "bit_identical": two_equal},
            "render_ranks_then_rgba8_composite": {
                "ms_per_frame": round(ms_rgba8, 5),
                "Gsamples/s": round(samples / (ms_rgba8 * 1e-3) / 1e9, 4),
                "note": "IceT's ubyte path: the ranks' RGBA8 images blended in 8 bits"},
            "fused_render_composite": {"ms_per_frame": round(ms_fused, 5),
                                       "Gsamples/s": round(samples / (ms_fused * 1e-3) / 1e9, 4),
                                       "bit_identical_to_plain": fused_equal},
        },
        "adjoint": adjoint,
        "batched": batched,
        "surface": surface,
        "cpu_baseline": cpu,
        "parity_sample": parity,
    }
    print(json.dumps(res), flush=True)
    ctx.close()


def _timer():
    import torch

    def timed(fn, n):
        stream = torch.cuda.current_stream()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(n):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / n
    return timed


def only_section(args):
    """One section alone (its JSON object on stdout), for profiling runs."""
    fn = {"adjoint": adjoint_bench, "batched": batch_bench, "surface": surface_bench}[args.only]
    print(json.dumps({args.only: fn(args, _timer())}), flush=True)


def adjoint_bench(args, timed):
    """Differentiable form at 256^3, 8 views of 512^2: forward-only vs
    forward+adjoint ms, and an Adam fit of TF + density to targets rendered
    from a perturbed TF and density."""
    import torch
    from aloha_amd import workloads as WL
    from aloha_amd.autograd import DifferentiableVolume
    t0 = time.time()
    dw = WL.diff256()
    vol = DifferentiableVolume(dw.density.shape, dw.cams, dw.W, dw.H, dt=dw.dt, jitter_value=0.5)
    tf = torch.from_numpy(dw.tf).cuda()
    den = torch.from_numpy(dw.density).cuda()
    with torch.no_grad():
        target, _ = vol.forward_only(torch.from_numpy(dw.target_tf).cuda(),
                                     torch.from_numpy(dw.target_density).cuda())
        target = target.clone()
    log(f"[bench] adjoint workload built in {time.time() - t0:.1f}s")
    tfp = tf.clone().requires_grad_(True)
    denp = den.clone().requires_grad_(True)

    def fwd():
        vol.forward_only(tfp, denp)

    def fwd_bwd():
        tfp.grad = None
        denp.grad = None
        ((vol.render(tfp, denp) - target) ** 2).sum().backward()

    n = max(3, min(args.steps, 20))
    for _ in range(2):
        fwd_bwd()
    ms_f = timed(fwd, n)
    ms_fb = timed(fwd_bwd, n)
    _, ns = vol.forward_only(tfp, denp)
    samples = int(ns.to(torch.int64).sum())
    log(f"[bench] adjoint: forward {ms_f:.3f} ms, forward+adjoint {ms_fb:.3f} ms")
    opt = torch.optim.Adam([tfp, denp], lr=2e-3)
    losses = []
    t0 = time.time()
    for _ in range(args.adam_steps):
        opt.zero_grad()
        loss = ((vol.render(tfp, denp) - target) ** 2).sum()
        loss.backward()
        opt.step()
        with torch.no_grad():
            tfp.clamp_(0, 1)
        losses.append(float(loss.detach()))
    if losses:
        log(f"[bench] Adam {args.adam_steps} steps: loss {losses[0]:.3f} -> {losses[-1]:.3f} "
            f"in {time.time() - t0:.1f}s")
    return {"workload": "256^3 f32 blob density, 256-entry TF, 8 views 512x512, dt 0.5",
            "samples_per_forward": samples, "fwd_ms": round(ms_f, 4),
            "fwd_adjoint_ms": round(ms_fb, 4), "ratio": round(ms_fb / ms_f, 3),
            "adam_steps": args.adam_steps,
            "loss_first": round(losses[0], 4) if losses else None,
            "loss_last": round(losses[-1], 4) if losses else None}


def batch_bench(args, timed):
    """Batched form: 64 uneven u8/f32 bricks (axes 32..256) on a 4x4x4 lattice,
    one virtual rank each, 16 views of 768x768 with per-view sort-last
    compositing, in one autotuned, hipGraph-captured launch -- against the
    plain per-view sequence (render_ranks over 64 slots + composite per view)
    and the per-brick launch sequence (one render launch per brick per view)."""
    import torch
    import aloha_amd as A
    from aloha_amd import RankImages
    from aloha_amd import workloads as WL
    t0 = time.time()
    bw = WL.batched()
    ctx = A.Context(0, stream=torch.cuda.current_stream())
    ids = [ctx.upload_brick(d, v) for d, v in zip(bw.bricks, bw.voxels)]
    ctx.set_tf(bw.tf)
    ranks = [[ids[i] for i in r] for r in bw.rank_bricks]
    V, W, H, P = len(bw.cams), bw.W, bw.H, len(ranks)
    log(f"[bench] batched workload built in {time.time() - t0:.1f}s "
        f"({sum(v.size for v in bw.voxels) / 1e6:.0f} M voxels)")
    out = torch.empty((V, H, W, 4), dtype=torch.float32, device="cuda")
    out8 = torch.empty((V, H, W), dtype=torch.int32, device="cuda")
    ns = torch.empty((V, H, W), dtype=torch.int32, device="cuda")
    b = A.Batch(ctx, bw.params(), W, H, bw.cams, ranks, bw.orders, out, out8, ns)
    b.launch()
    torch.cuda.synchronize()
    samples = int(ns.to(torch.int64).sum())
    best, table = b.autotune(reps=3)
    b.capture()
    sched = b.schedule()  # the timed schedule, after the graph capture
    n = max(3, min(args.steps, 10))
    ms_batch = timed(b.launch, n)
    # plain per-view sequence
    imgs = RankImages.alloc(P, W, H, rgba=True, rgba8=False, depth=False)
    ref = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
    p = bw.params()

    def per_view():
        for cam, order in zip(bw.cams, bw.orders):
            ctx.set_camera(cam["org"], cam["dir_00"], cam["dir_du"], cam["dir_dv"])
            ctx.render_ranks(p, W, H, ranks, imgs)
            ctx.composite_over(imgs.rgba, order, W, H, ref, None)

    per_view()
    ms_views = timed(per_view, max(2, n // 3))
    equal = bool(torch.equal(ref, out[V - 1]))
    # per-brick launch sequence: one render launch per (view, brick)
    one = RankImages.alloc(1, W, H, rgba=True, rgba8=False, depth=False)

    def per_brick():
        for cam, order in zip(bw.cams, bw.orders):
            ctx.set_camera(cam["org"], cam["dir_00"], cam["dir_du"], cam["dir_dv"])
            for s in range(P):
                sub = RankImages(rgba=imgs.rgba[s:s + 1])
                ctx.render_ranks(p, W, H, [ranks[s]], sub)
            ctx.composite_over(imgs.rgba, order, W, H, ref, None)

    per_brick()
    ms_bricks = timed(per_brick, 2)
    del one
    log(f"[bench] batched: graph {ms_batch:.3f} ms ({sched}), per-view sequence {ms_views:.3f} ms, "
        f"per-brick sequence {ms_bricks:.3f} ms, equal {equal}")
    b.close()
    ctx.close()
    return {"workload": "64 bricks (axes U[32,256], u8/f32), 16 views 768x768, per-view kd order",
            "samples": samples, "ms_graph_launch": round(ms_batch, 4),
            "Gsamples/s": round(samples / (ms_batch * 1e-3) / 1e9, 3),
            "ms_per_view_sequence": round(ms_views, 4), "ms_per_brick_sequence": round(ms_bricks, 4),
            "speedup_vs_per_view": round(ms_views / ms_batch, 2),
            "speedup_vs_per_brick": round(ms_bricks / ms_batch, 2),
            "autotune": {"chosen": sched, "best_ms": round(best, 4),
                         "table": [{"lanes_per_ray": t, "persistent_blocks": pb, "ms": round(m, 4)}
                                   for t, pb, m in table]},
            "last_view_equal_to_per_view_sequence": equal}


def surface_bench(args, timed):
    """Triangle surface path: 4096 displaced spheres (9.75 M triangles),
    MeshSplitter into 8 clusters, one per virtual rank, 1920x1080.  One frame
    = per-rank RGBA8 + depth images (one launch, BVH traversal) + depth
    composite (one launch); the fused form does both in one launch.  Rays
    count one primary ray per pixel per rank."""
    import torch
    import aloha_amd as A
    from aloha_amd import RankImages
    from aloha_amd import workloads as WL
    t0 = time.time()
    sw = WL.surface()
    t_host = time.time() - t0
    stream = torch.cuda.current_stream()
    ctx = A.Context(0, stream=stream)
    ctx.set_camera(sw.cam["org"], sw.cam["dir_00"], sw.cam["dir_du"], sw.cam["dir_dv"])
    t0 = time.time()
    geoms = ctx.upload_tris(sw.mesh, sw.ranks, sw.colors)
    t_up = time.time() - t0
    info = [ctx.tris_info(g) for g in geoms]
    log(f"[bench] surface: {sw.ntris} triangles, scene {t_host:.1f}s, BVH build + upload "
        f"{t_up:.1f}s, depth {max(i['depth'] for i in info)}")
    P, W, H = len(geoms), sw.W, sw.H
    p = sw.params()
    imgs = RankImages.alloc(P, W, H, rgba=False, rgba8=True, depth=True)
    out8 = torch.empty((H, W), dtype=torch.int32, device="cuda")
    outz = torch.empty((H, W), dtype=torch.float32, device="cuda")

    def frame():
        ctx.render_tris(p, W, H, geoms, imgs)
        ctx.composite_zmin(imgs.rgba8, imgs.depth, P, W, H, out8, outz)

    for _ in range(args.warmup):
        frame()
    n = max(3, min(args.steps, 20))
    ms_frame = timed(frame, n)
    ms_render = timed(lambda: ctx.render_tris(p, W, H, geoms, imgs), n)
    ms_zmin = timed(lambda: ctx.composite_zmin(imgs.rgba8, imgs.depth, P, W, H, out8, outz), n)
    f8 = torch.empty_like(out8)
    fz = torch.empty_like(outz)
    fused = lambda: ctx.render_tris_composite(p, W, H, geoms, f8, fz)
    fused()
    ms_fused = timed(fused, n)
    torch.cuda.synchronize()
    equal = bool(torch.equal(f8, out8) and torch.equal(fz, outz))
    # traversal work of one frame (untimed): node fetches and triangle tests
    st = torch.zeros((P, H, W, 2), dtype=torch.int32, device="cuda")
    ctx.render_tris(p, W, H, geoms, imgs, stats=st)
    fst = torch.zeros((H, W, 2), dtype=torch.int32, device="cuda")
    ctx.render_tris_composite(p, W, H, geoms, f8, fz, stats=fst)
    torch.cuda.synchronize()
    tot = st.to(torch.int64).sum(dim=(0, 1, 2)).tolist()
    ftot = fst.to(torch.int64).sum(dim=(0, 1)).tolist()
    rays = P * W * H
    hit_frac = float((outz < 1).float().mean())
    # algorithmic bytes of one per-rank launch: nodes x 64 + triangle tests x 48
    # + the RGBA8 and depth images written
    bytes_render = tot[0] * 64 + tot[1] * 48 + rays * 8
    log(f"[bench] surface: frame {ms_frame:.3f} ms (render {ms_render:.3f}, zmin {ms_zmin:.3f}), "
        f"fused {ms_fused:.3f} ms (equal {equal}), {tot[0] / rays:.1f} nodes + {tot[1] / rays:.1f} "
        f"tris per ray")
    res = {"workload": f"{sw.ntris} triangles (4096 displaced spheres), 8 clusters = 8 ranks, "
                       f"{W}x{H}, headlight shading",
           "rays_per_frame": rays, "hit_fraction": round(hit_frac, 4),
           "bvh": {"nodes": sum(i["nnodes"] for i in info), "max_depth": max(i["depth"] for i in info),
                   "bytes": sum(i["bytes"] for i in info), "build_upload_s": round(t_up, 2)},
           "ms_per_frame": round(ms_frame, 4), "Mrays/s": round(rays / (ms_frame * 1e-3) / 1e6, 1),
           "render_ms": round(ms_render, 4), "zmin_ms": round(ms_zmin, 4),
           "render_Mrays/s": round(rays / (ms_render * 1e-3) / 1e6, 1),
           "nodes_per_ray": round(tot[0] / rays, 2), "tris_per_ray": round(tot[1] / rays, 2),
           "render_algorithmic_GB/s": round(bytes_render / (ms_render * 1e-3) / 1e9, 1),
           "fused_ms": round(ms_fused, 4), "fused_Mrays/s": round(rays / (ms_fused * 1e-3) / 1e6, 1),
           "fused_nodes_per_pixel": round(ftot[0] / (W * H), 2),
           "fused_bit_identical": equal}
    if not args.no_cpu:
        res["cpu_baseline"] = surface_cpu_baseline(ctx, geoms, sw, imgs, args.cpu_seconds / 3)
    ctx.close()
    return res


def surface_cpu_baseline(ctx, geoms, sw, imgs, seconds):
    """The oracle's CPU walk of the same BVHs (OpenMP) on a band of rows of
    every rank; also checks those rows against the GPU images."""
    import oracle as O
    threads = min(16, os.cpu_count() or 1)
    exported = [ctx.tris_export(g) for g in geoms]
    d8 = imgs.rgba8.cpu().numpy().view(np.uint32)
    dz = imgs.depth.cpu().numpy()
    band, H = 4, sw.H
    while True:
        rows = (H // 2 - band // 2, H // 2 + band // 2)
        t = time.time()
        outs = [O.tri_render_bvh(sw.cam, nd, tr, rt, sw.W, sw.H, max_depth=sw.max_depth,
                                 rows=rows, threads=threads) for nd, tr, rt in exported]
        el = time.time() - t
        if el * 2 > seconds or band >= H:
            break
        band = min(H, band * max(2, int(seconds / max(el, 1e-3) / 2)))
    rays = len(geoms) * sw.W * (rows[1] - rows[0])
    sl = slice(rows[0], rows[1])
    mism = sum(int((o["rgba8"][sl] != d8[r, sl]).sum() + (o["depth"][sl] != dz[r, sl]).sum())
               for r, o in enumerate(outs))
    return {"value": round(rays / el / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"rows {rows[0]}..{rows[1]} of all {len(geoms)} ranks ({rays} rays, {el:.1f}s), "
                      f"oracle BVH walk, OpenMP",
            "gpu_mismatched_values_in_sample": mism}


def cpu_baseline(wl, imgs, seconds):
    """The oracle (FP32 CPU restatement, OpenMP) on a bounded band of rows of
    every rank, on this host's cores; also the FP64 error of both paths on
    that band."""
    import oracle as O
    threads = min(16, os.cpu_count() or 1)
    H = wl.H
    gpu_rgba = imgs.rgba.cpu().numpy()
    # band through the image centre, grown until it holds ~`seconds` of work
    band = 8
    per_rank = lambda r, rows: O.render_rank(
        wl.cam, wl.tf, [(wl.bricks[i], wl.voxels[i]) for i in wl.rank_bricks[r]], wl.W, wl.H,
        dt=wl.dt, max_depth=wl.max_depth, rows=rows, threads=threads)
    while True:
        rows = (H // 2 - band // 2, H // 2 + band // 2)
        t = time.time()
        ns = 0
        for r in range(wl.num_ranks):
            ns += int(per_rank(r, rows)["nsamples"].sum())
        el = time.time() - t
        if el * 2 > seconds or band >= H:
            break
        band = min(H, band * max(2, int(seconds / max(el, 1e-3) / 2)))
    reps = 1
    if el < seconds / 2:  # whole frames are cheap: repeat them up to the budget
        reps = max(1, int(seconds / max(el, 1e-3)))
        t = time.time()
        ns = 0
        for _ in range(reps):
            for r in range(wl.num_ranks):
                ns += int(per_rank(r, rows)["nsamples"].sum())
        el = time.time() - t
    cpu_gs = ns / el / 1e9
    # FP64 error statement on a small band
    r0 = (H // 2 - 4, H // 2 + 4)
    err_gpu, err_cpu, max_ulp = 0.0, 0.0, 0.0
    for r in range(wl.num_ranks):
        o = O.render_rank(wl.cam, wl.tf, [(wl.bricks[i], wl.voxels[i]) for i in wl.rank_bricks[r]],
                          wl.W, wl.H, dt=wl.dt, max_depth=wl.max_depth, rows=r0, want64=True,
                          threads=threads)
        sl = slice(r0[0], r0[1])
        err_gpu = max(err_gpu, float(np.abs(gpu_rgba[r, sl] - o["rgba64"][sl]).max()))
        err_cpu = max(err_cpu, float(np.abs(o["rgba"][sl] - o["rgba64"][sl]).max()))
        max_ulp = max(max_ulp, float(np.abs(gpu_rgba[r, sl] - o["rgba"][sl]).max()))
    cpu = {"value": round(cpu_gs, 5), "unit": "Gsamples/s", "cores": threads, "kind": "port",
           "sample": f"{reps} x rows {rows[0]}..{rows[1]} of all {wl.num_ranks} ranks ({ns} "
                     f"samples, {el:.1f}s), FP32 oracle, OpenMP"}
    parity = {"rows": f"{r0[0]}..{r0[1]}", "max_abs_gpu_vs_cpu_fp32": max_ulp,
              "max_abs_gpu_vs_fp64": err_gpu, "max_abs_cpu_fp32_vs_fp64": err_cpu}
    return cpu, parity


if __name__ == "__main__":
    main()
