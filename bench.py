#!/usr/bin/env python3
"""Benchmark of the per-brick render-and-composite path on one MI355X.

One step = one frame (aloha_render_frame, the frame only): every virtual
rank ray-marches its bricks into its rank image -- stored compact, only
inside the pixel rectangle its bricks project to (IceT's sparse images; the
rank image is transparent black outside it) -- and the ranks' images are
'over'-composited in kd visibility order into the float RGBA + RGBA8 frame.
Two launches: the per-rank launch (which also clears the frame outside the
rectangles' bounding box) and the composite of that box.  Timed beside it,
each checked bit-identical to it: the same frame with full-size rank images
(float RGBA + RGBA8 + depth per rank, as the reference's ranks write their
framebuffers), as two calls (render_ranks, then the sparse composite), with
the dense composite, from the allocation-free entry point, and the fused
one-launch form.  Inputs (bricks, TF, camera) are resident in HBM before
timing starts.

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

    def frame():  # the frame only: compact rank images in the context's workspace
        ctx.render_frame(params, W, H, ranks, None, wl.order, out, out8)

    dout = torch.empty_like(out)
    dout8 = torch.empty_like(out8)

    def frame_with_images():  # the same frame, every rank's full image written too
        ctx.render_frame(params, W, H, ranks, imgs, wl.order, dout, dout8)

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

    timed = _timer()

    if args.profile:
        calibration_copy()
    for _ in range(args.warmup):
        frame()
    torch.cuda.synchronize()
    ms_frame = timed(frame, args.steps)
    log(f"[bench] {args.steps} frames: {ms_frame:.4f} ms/frame")
    for _ in range(args.warmup):
        frame_with_images()
    ms_frame_img = timed(frame_with_images, args.steps)
    torch.cuda.synchronize()
    img_equal = bool(torch.equal(dout, out) and torch.equal(dout8, out8))
    log(f"[bench] the frame with full-size rank images: {ms_frame_img:.4f} ms/frame "
        f"(equal {img_equal})")
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
    torch.cuda.synchronize()
    # 8-bit levels of the float path's frame (packed from its float 'over')
    # against the 8-bit blend of the ranks' RGBA8 images, per channel
    lv = lambda t: t.view(torch.uint8).reshape(H, W, 4).to(torch.int16)
    level_diff = (lv(out8) - lv(out8b)).abs()
    rgba8_levels = {"max": int(level_diff.max()),
                    "pixels_differing": int((level_diff.amax(dim=2) > 0).sum()),
                    "histogram": [int((level_diff.amax(dim=2) == k).sum()) for k in range(4)]}
    log(f"[bench] RGBA8 composite form (render_ranks + 8-bit ordered blend): {ms_rgba8:.4f} ms/frame, "
        f"levels vs float path {rgba8_levels}")
    # the allocation-free entry point (the same compact rank images in a static
    # per-GPU workspace), same result
    nout = torch.empty_like(out)
    nout8 = torch.empty_like(out8)
    noalloc = lambda: ctx.render_frame_noalloc(params, W, H, ranks, wl.order, nout, nout8)
    for _ in range(args.warmup):
        noalloc()
    ms_noalloc = timed(noalloc, args.steps)
    torch.cuda.synchronize()
    noalloc_equal = bool(torch.equal(nout, out) and torch.equal(nout8, out8))
    log(f"[bench] allocation-free entry point {ms_noalloc:.4f} ms "
        f"(equal {noalloc_equal})")
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
    roof["note"] = ("the per-rank launch of render_ranks (full-size rank images), timed alone; "
                    "the headline frame's two launches are frame_roofline")
    # the headline frame: voxels read, compact rank pixels written and read
    # back by the composite, the frame's float RGBA + RGBA8 written
    frame_bytes = touched * 4 + wl.tf.nbytes + rect_pix * (16 + 16) + npix * (16 + 4)
    frame_gbs = frame_bytes / (ms_frame * 1e-3) / 1e9
    frame_roof = {"kernels": f"march_persist_kernel<{schedule[0]}> (compact) + over_compact",
                  "bound": "hbm", "achieved": round(frame_gbs, 1), "peak": HBM_PEAK_GBS,
                  "unit": "GB/s", "frac": round(frame_gbs / HBM_PEAK_GBS, 4),
                  "algorithmic_bytes": int(frame_bytes), "ms": round(ms_frame, 5)}
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
    fp64_c0 = None
    if not args.no_cpu:
        cpu, parity = cpu_baseline(wl, imgs, args.cpu_seconds)
        fp64_c0 = fp64_blobs128()
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
        "frame_roofline": frame_roof,
        "kernels": kernels,
        "forms": {
            "render_frame": {"ms_per_frame": round(ms_frame, 5), "Gsamples/s": round(gsamples, 4),
                             "note": "the headline step: compact rank images (only inside their "
                                     "rectangles) in the context's workspace"},
            "render_frame_with_full_rank_images": {
                "ms_per_frame": round(ms_frame_img, 5),
                "Gsamples/s": round(samples / (ms_frame_img * 1e-3) / 1e9, 4),
                "bit_identical": img_equal,
                "note": "every rank's 1024x1024 float RGBA + RGBA8 + depth image written too"},
            "render_ranks_then_sparse_composite": {
                "ms_per_frame": round(ms_two, 5),
                "Gsamples/s": round(samples / (ms_two * 1e-3) / 1e9, 4),
                "bit_identical": two_equal},
            "render_ranks_then_rgba8_composite": {
                "ms_per_frame": round(ms_rgba8, 5),
                "Gsamples/s": round(samples / (ms_rgba8 * 1e-3) / 1e9, 4),
                "note": "IceT's ubyte path: the ranks' RGBA8 images blended in 8 bits",
                "levels_vs_float_path": rgba8_levels},
            "render_frame_noalloc": {
                "ms_per_frame": round(ms_noalloc, 5),
                "Gsamples/s": round(samples / (ms_noalloc * 1e-3) / 1e9, 4),
                "bit_identical": noalloc_equal,
                "note": "the allocation-free entry point: the same compact rank images in a "
                        "static per-GPU workspace"},
            "fused_render_composite": {"ms_per_frame": round(ms_fused, 5),
                                       "Gsamples/s": round(samples / (ms_fused * 1e-3) / 1e9, 4),
                                       "bit_identical_to_plain": fused_equal},
        },
        "adjoint": adjoint,
        "batched": batched,
        "surface": surface,
        "cpu_baseline": cpu,
        "parity_sample": parity,
        "fp64_error": {
            "C0 blobs128 2 bricks 512^2 (frame)": fp64_c0,
            "C1 fbm512 8 bricks 1024^2 (rank images, band)": parity,
            "C2 surface 1920x1080 (hit pixels, band)": (surface or {}).get("cpu_baseline", {}).get("fp64"),
            "C3 batched view 0 (frame, band)": (batched or {}).get("fp64"),
            "C4 differentiable forward view 0": (adjoint or {}).get("fp64"),
        },
    }
    print(json.dumps(res), flush=True)
    ctx.close()


def _timer():
    """timed(fn, n): mean ms of n calls of fn, from HIP events recorded on
    the stream current at the call (the stream the timed work runs on)."""
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


def calibration_copy():
    """For HBM-traffic counters: one fill and one copy of exactly 256 MiB."""
    import torch
    src = torch.empty(64 << 20, dtype=torch.float32, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst


def only_section(args):
    """One section alone (its JSON object on stdout), for profiling runs."""
    fn = {"adjoint": adjoint_bench, "batched": batch_bench, "surface": surface_bench}[args.only]
    if args.profile:
        calibration_copy()
    print(json.dumps({args.only: fn(args, _timer())}), flush=True)


def adjoint_bench(args, timed):
    """Differentiable form at 256^3, 8 views of 512^2: forward-only vs
    forward+adjoint ms (and the adjoint alone), and an Adam fit of TF +
    density to targets rendered from a perturbed TF and density; the timings
    and samples per forward are taken again after the Adam steps."""
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
    V, H, W = len(dw.cams), dw.H, dw.W

    def fwd():
        vol.forward_only(tfp, denp)

    def fwd_bwd():
        tfp.grad = None
        denp.grad = None
        ((vol.render(tfp, denp) - target) ** 2).sum().backward()

    n = max(3, min(args.steps, 20))

    def measure():
        for _ in range(2):
            fwd_bwd()
        ms_f = timed(fwd, n)
        ms_fb = timed(fwd_bwd, n)
        # the adjoint launches alone (aloha_diff_backward on a fixed dL/dimage)
        rgba, ns = vol.forward_only(tfp, denp)
        grad = (2 * (rgba - target)).contiguous()
        dtf = torch.zeros_like(tfp)
        dd = torch.zeros_like(denp)
        ms_b = timed(lambda: vol._backward(rgba, ns, grad, dtf, dd), n)
        return ms_f, ms_fb, ms_b, int(ns.to(torch.int64).sum())

    ms_f, ms_fb, ms_b, samples = measure()
    log(f"[bench] adjoint: forward {ms_f:.3f} ms, forward+adjoint {ms_fb:.3f} ms, "
        f"adjoint launches {ms_b:.3f} ms, {samples} samples")
    plain = plain_march_views(vol, dw, tfp, denp, timed, n)
    log(f"[bench] the same grid, TF and views through the plain per-brick march "
        f"(render_ranks per view): {plain['ms']:.3f} ms, {plain['samples']} samples")
    # compulsory voxel reads: the footprint of the same views' march of this
    # grid (the differentiable march samples the same lattice)
    touched = 0
    for cam in dw.cams:
        vol.ctx.set_camera(cam["org"], cam["dir_00"], cam["dir_du"], cam["dir_dv"])
        touched += vol.ctx.footprint(vol.params, W, H, [[vol.brick]])
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
    adam_s = time.time() - t0
    end = None
    if losses:
        log(f"[bench] Adam {args.adam_steps} steps: loss {losses[0]:.3f} -> {losses[-1]:.3f} "
            f"in {adam_s:.1f}s")
        e_f, e_fb, e_b, e_samples = measure()
        log(f"[bench] adjoint after Adam: forward {e_f:.3f} ms, forward+adjoint {e_fb:.3f} ms, "
            f"adjoint launches {e_b:.3f} ms, {e_samples} samples")
        end = {"fwd_ms": round(e_f, 4), "fwd_adjoint_ms": round(e_fb, 4),
               "adjoint_launches_ms": round(e_b, 4), "ratio": round(e_fb / e_f, 3),
               "samples_per_forward": e_samples}
    # algorithmic bytes: forward = compulsory voxel reads + the images and
    # sample counts written; adjoint = the same voxel reads + the forward
    # image and dL/dimage read + dDensity (whole grid) and dTF written
    grid = int(np.prod(dw.density.shape))
    img = V * H * W
    fwd_bytes = touched * 4 + dw.tf.nbytes + img * (16 + 4)
    bwd_bytes = touched * 4 + dw.tf.nbytes + img * (16 + 16 + 4) + grid * 4 + dw.tf.nbytes
    roof = lambda b, ms: {"bound": "hbm", "achieved": round(b / (ms * 1e-3) / 1e9, 1),
                          "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": round(b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                          "algorithmic_bytes": int(b), "ms": round(ms, 4)}
    res = {"workload": "256^3 f32 blob density, 256-entry TF, 8 views 512x512, dt 0.5",
           "samples_per_forward": samples, "touched_voxels": touched,
           "fwd_ms": round(ms_f, 4), "fwd_adjoint_ms": round(ms_fb, 4),
           "adjoint_launches_ms": round(ms_b, 4), "ratio": round(ms_fb / ms_f, 3),
           "plain_per_brick_march": plain,
           "roofline_forward": roof(fwd_bytes, ms_f),
           "roofline_adjoint": roof(bwd_bytes, ms_b),
           "adam_steps": args.adam_steps, "adam_s": round(adam_s, 2),
           "loss_first": round(losses[0], 4) if losses else None,
           "loss_last": round(losses[-1], 4) if losses else None,
           "after_adam": end}
    if not args.no_cpu:
        res["fp64"] = fp64_adjoint_forward(dw, vol, tf, den)
    return res


def plain_march_views(vol, dw, tf, den, timed, n):
    """The plain per-brick form on the differentiable workload: the same
    256^3 grid (one brick, one rank), TF and 8 views rendered by render_ranks
    one view after another -- renderFrame's semantics (repeated t += dt, no
    opacity clamp), so its images differ from the differentiable forward's by
    those two rules, not by accuracy."""
    import torch
    import aloha_amd as A
    from aloha_amd import RankImages
    vol.forward_only(tf, den)  # the context holds these exact inputs
    p = A.render_params(dt=dw.dt, jitter_mode=1, jitter_value=0.5)
    V, H, W = len(dw.cams), dw.H, dw.W
    imgs = [RankImages.alloc(1, W, H, rgba=True, rgba8=False, depth=False, nsamples=True)
            for _ in range(V)]

    def views():
        for cam, im in zip(dw.cams, imgs):
            vol.ctx.set_camera(cam["org"], cam["dir_00"], cam["dir_du"], cam["dir_dv"])
            vol.ctx.render_ranks(p, W, H, [[vol.brick]], im)

    views()
    ms = timed(views, n)
    torch.cuda.synchronize()
    samples = int(sum(int(im.nsamples.to(torch.int64).sum()) for im in imgs))
    return {"ms": round(ms, 4), "samples": samples,
            "note": "render_ranks per view (8 launches) on the same grid, TF and views"}


def fp64_adjoint_forward(dw, vol, tf, den):
    """CPU leg: the float64 re-render of view 0 at the start parameters (on
    the float march's decisions) against the GPU forward and the oracle's
    float32 forward."""
    import torch
    import oracle as O
    threads, _ = host_threads()
    rgba, ns = vol.forward_only(tf, den)
    torch.cuda.synchronize()
    g = rgba[0].cpu().numpy()
    n = dw.density.shape[::-1]
    blist = [({"cell": ((0, 0, 0), n), "voxel": ((0, 0, 0), n),
               "space": ((0.0, 0.0, 0.0), tuple(float(v) for v in n))}, dw.density)]
    kw = dict(dt=dw.dt, jitter_mode=1, jitter_value=0.5, threads=threads)
    r32, _ = O.diff_render(dw.cams[0], dw.tf, blist, dw.W, dw.H, **kw)
    r64 = O.diff_render64(dw.cams[0], dw.tf, blist, dw.W, dw.H, **kw)
    return _fp64_line(g, r32, r64, ns[0].cpu().numpy() > 0, "view 0, all pixels")


def _fp64_line(gpu, cpu32, ref64, mask, what):
    e = np.abs(gpu - ref64).max(-1)
    sel = e[mask] if mask.any() else e.ravel()
    return {"sample": what, "max_abs_gpu_vs_fp64": float(e.max()),
            "p999_abs_gpu_vs_fp64": float(np.quantile(sel, 0.999)),
            "max_abs_cpu_fp32_vs_fp64": float(np.abs(cpu32 - ref64).max()),
            "max_abs_gpu_vs_cpu_fp32": float(np.abs(gpu - cpu32).max())}


def batch_bench(args, timed):
    """Batched form: 64 uneven u8/f32 bricks (axes 32..256) on a 4x4x4 lattice,
    one virtual rank each, 16 views of 768x768 with per-view sort-last
    compositing, in one autotuned, hipGraph-captured launch -- against the
    per-view sequence (render_ranks over the 64 slots + sparse composite, per
    view) and the per-brick sequence (one render launch per brick per view +
    sparse composite per view), each captured as one graph too, so all three
    time the device and not the host's launch loop."""
    import torch
    import aloha_amd as A
    from aloha_amd import RankImages
    from aloha_amd import workloads as WL
    t0 = time.time()
    bw = WL.batched()
    side = torch.cuda.Stream()  # graph capture needs a stream other than the null stream
    ctx = A.Context(0, stream=side)
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
    with torch.cuda.stream(side):
        ms_batch = timed(b.launch, n)
    # the baselines: per-view camera, rectangles and order fixed at capture
    imgs = RankImages.alloc(P, W, H, rgba=True, rgba8=False, depth=False)
    ref = torch.empty((V, H, W, 4), dtype=torch.float32, device="cuda")
    p = bw.params()
    rects = []
    for cam in bw.cams:
        ctx.set_camera(cam["org"], cam["dir_00"], cam["dir_du"], cam["dir_dv"])
        rects.append(ctx.rank_rects(p, W, H, ranks))

    def per_view():
        for v, (cam, order) in enumerate(zip(bw.cams, bw.orders)):
            ctx.set_camera(cam["org"], cam["dir_00"], cam["dir_du"], cam["dir_dv"])
            ctx.render_ranks(p, W, H, ranks, imgs)
            ctx.composite_over(imgs.rgba, order, W, H, ref[v], None, rects=rects[v])

    def per_brick():
        for v, (cam, order) in enumerate(zip(bw.cams, bw.orders)):
            ctx.set_camera(cam["org"], cam["dir_00"], cam["dir_du"], cam["dir_dv"])
            for s in range(P):
                ctx.render_ranks(p, W, H, [ranks[s]], RankImages(rgba=imgs.rgba[s:s + 1]))
            ctx.composite_over(imgs.rgba, order, W, H, ref[v], None, rects=rects[v])

    def graph_of(fn):
        with torch.cuda.stream(side):
            fn()  # warm-up outside the capture (lazy state, occupancy queries)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=side, capture_error_mode="thread_local"):
            fn()
        return g

    res_seq = {}
    for name, fn in (("per_view", per_view), ("per_brick", per_brick)):
        ref.fill_(float("nan"))
        g = graph_of(fn)
        with torch.cuda.stream(side):
            ms_g = timed(g.replay, max(2, n // 2))
            ms_h = timed(fn, 2)  # the same launches issued from Python
        torch.cuda.synchronize()
        res_seq[name] = {"ms_graph": round(ms_g, 4), "ms_host_loop": round(ms_h, 4),
                         "launches": V * (1 + (1 if name == "per_view" else P)),
                         "equal_to_batch": bool(torch.equal(ref, out))}
        del g
    log(f"[bench] batched: graph {ms_batch:.3f} ms ({sched}); captured per-view sequence "
        f"{res_seq['per_view']['ms_graph']:.3f} ms, per-brick sequence "
        f"{res_seq['per_brick']['ms_graph']:.3f} ms; {res_seq}")
    # compulsory voxel reads of every view (the footprint pass)
    touched = 0
    for cam in bw.cams:
        ctx.set_camera(cam["org"], cam["dir_00"], cam["dir_du"], cam["dir_dv"])
        touched += ctx.footprint(p, W, H, ranks)
    # algorithmic bytes: voxels read (4 B: bricks are stored decoded to f32) +
    # the TF + the V frames written (float RGBA, RGBA8, sample counts)
    bbytes = touched * 4 + bw.tf.nbytes + V * H * W * (16 + 4 + 4)
    gbs = bbytes / (ms_batch * 1e-3) / 1e9
    res = {"workload": "64 bricks (axes U[32,256], u8/f32), 16 views 768x768, per-view kd order",
           "samples": samples, "touched_voxels": touched, "ms_graph_launch": round(ms_batch, 4),
           "Gsamples/s": round(samples / (ms_batch * 1e-3) / 1e9, 3),
           "roofline": {"kernel": f"batch_kernel<{sched['lanes_per_ray']}, "
                                  f"{str(sched['persistent_blocks'] > 0).lower()}>",
                        "bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                        "algorithmic_bytes": int(bbytes)},
           "per_view_sequence": res_seq["per_view"], "per_brick_sequence": res_seq["per_brick"],
           "speedup_vs_per_view_graph": round(res_seq["per_view"]["ms_graph"] / ms_batch, 2),
           "speedup_vs_per_brick_graph": round(res_seq["per_brick"]["ms_graph"] / ms_batch, 2),
           "autotune": {"chosen": sched, "best_ms": round(best, 4),
                        "table": [{"tile_w": tw, "lanes_per_ray": t, "persistent_blocks": pb,
                                   "ms": round(m, 4)} for tw, t, pb, m in table]}}
    res["u8_storage"] = batch_u8_storage(bw, out, ms_batch, ctx.brick_bytes(), timed)
    if not args.no_cpu:
        res["fp64"] = fp64_batched(bw, out)
    b.close()
    ctx.close()
    return res


def batch_u8_storage(bw, ref_out, ms_f32, f32_bytes, timed):
    """The same batch with the u8 bricks kept as bytes in HBM (decoded at
    sample time through u/255.f): the launch time, the brick memory and the
    frames' equality with the float32-storage batch."""
    import torch
    import aloha_amd as A
    side = torch.cuda.Stream()
    ctx = A.Context(0, stream=side)
    ctx.set_u8_storage(True)
    ids = [ctx.upload_brick(d, v) for d, v in zip(bw.bricks, bw.voxels)]
    ctx.set_tf(bw.tf)
    ranks = [[ids[i] for i in r] for r in bw.rank_bricks]
    V, W, H = len(bw.cams), bw.W, bw.H
    out = torch.empty((V, H, W, 4), dtype=torch.float32, device="cuda")
    b = A.Batch(ctx, bw.params(), W, H, bw.cams, ranks, bw.orders, out)
    best, _ = b.autotune(reps=3)
    b.capture()
    sched = b.schedule()
    with torch.cuda.stream(side):
        ms = timed(b.launch, 10)
    torch.cuda.synchronize()
    res = {"ms_graph_launch": round(ms, 4), "schedule": sched,
           "brick_bytes": ctx.brick_bytes(), "brick_bytes_f32_storage": f32_bytes,
           "u8_bricks": sum(1 for v in bw.voxels if v.dtype == np.uint8),
           "time_vs_f32_storage": round(ms / ms_f32, 3),
           "bit_identical_to_f32_storage": bool(torch.equal(out, ref_out))}
    log(f"[bench] batched with u8 bricks kept as bytes: {ms:.3f} ms ({sched}), "
        f"{res['brick_bytes'] / 1e6:.0f} MB of bricks against {f32_bytes / 1e6:.0f} MB, "
        f"equal {res['bit_identical_to_f32_storage']}")
    b.close()
    ctx.close()
    return res


def fp64_batched(bw, out):
    """CPU leg: view 0 of the batch on a band of 8 rows through the centre:
    each rank rendered by the oracle in float32 and float64, composited in
    the view's order, against the GPU frame."""
    import torch
    import aloha_amd as A
    import oracle as O
    threads, _ = host_threads()
    torch.cuda.synchronize()
    W, H = bw.W, bw.H
    rows = (H // 2 - 4, H // 2 + 4)
    sl = slice(*rows)
    cam, order = bw.cams[0], bw.orders[0]
    C = np.zeros((H, W, 4), np.float32)
    C64 = np.zeros((H, W, 4), np.float64)
    cnt = np.zeros((H, W), np.uint32)
    for s in order:
        (bi,) = bw.rank_bricks[s]
        d = bw.bricks[bi]
        x0, y0, x1, y1 = A.project_boxes(
            cam, [(tuple(d.space_range.lower), tuple(d.space_range.upper))], W, H)
        if x0 >= x1 or max(y0, rows[0]) >= min(y1, rows[1]):
            continue
        o = O.render_rank(cam, bw.tf, [(d, bw.voxels[bi])], W, H, dt=bw.dt, rows=rows,
                          want64=True, threads=threads)
        C[sl] = C[sl] + (np.float32(1.0) - C[sl, :, 3:4]) * o["rgba"][sl]
        C64[sl] = C64[sl] + (1.0 - C64[sl, :, 3:4]) * o["rgba64"][sl]
        cnt[sl] += o["nsamples"][sl]
    g = out[0, rows[0]:rows[1]].cpu().numpy()
    return _fp64_line(g, C[sl], C64[sl], cnt[sl] > 0, f"view 0, rows {rows[0]}..{rows[1]}")


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
    stn = st.cpu().numpy().reshape(-1, 2)
    tail = {"nodes_per_ray_p99": float(np.percentile(stn[:, 0], 99)),
            "nodes_per_ray_max": int(stn[:, 0].max()),
            "tris_per_ray_p99": float(np.percentile(stn[:, 1], 99)),
            "tris_per_ray_max": int(stn[:, 1].max())}
    del stn
    ftot = fst.to(torch.int64).sum(dim=(0, 1)).tolist()
    rays = P * W * H
    hit_frac = float((outz < 1).float().mean())
    # algorithmic bytes of one per-rank launch: nodes x 128 + triangle tests x 48
    # + the RGBA8 and depth images written
    bytes_render = tot[0] * 128 + tot[1] * 48 + rays * 8
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
           **tail,
           "render_algorithmic_GB/s": round(bytes_render / (ms_render * 1e-3) / 1e9, 1),
           "roofline": {"kernel": "tris_rank_kernel<false>", "bound": "hbm",
                        "achieved": round(bytes_render / (ms_render * 1e-3) / 1e9, 1),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(bytes_render / (ms_render * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                        "algorithmic_bytes": int(bytes_render),
                        "note": "bytes of the traversal this BVH needs (128 B per 4-wide node visit, 48 B "
                                "per triangle test, not deduplicated) + the images written"},
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
    threads, host = host_threads()
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
    fp64 = fp64_surface(sw, exported, outs, d8, dz, rows)
    return {"value": round(rays / el / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"rows {rows[0]}..{rows[1]} of all {len(geoms)} ranks ({rays} rays, {el:.1f}s), "
                      f"oracle BVH walk, OpenMP",
            "gpu_mismatched_values_in_sample": mism, "host": host, "fp64": fp64}


def fp64_surface(sw, exported, outs, d8, dz, rows):
    """CPU leg: for the band's hit pixels, the closest hit's distance and
    headlight colour recomputed in float64 (Moller-Trumbore on the hit
    triangle the oracle's walk found, ray direction and normal in float64)
    against the GPU's depth and 8-bit colour."""
    import oracle as O
    cam = {k: np.asarray(v, np.float64) for k, v in sw.cam.items()}
    sl = slice(*rows)
    ys, xs = np.mgrid[sl, 0:sw.W]
    d = (cam["dir_00"][None, None] + (xs[..., None] + .5) * cam["dir_du"] +
         (ys[..., None] + .5) * cam["dir_dv"])
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    d = np.where(np.abs(d) < 1e-5, 1e-5, d)
    worst_z, worst_rel, worst_lv, hits = 0.0, 0.0, 0, 0
    for r, ((nodes, tris, root), o) in enumerate(zip(exported, outs)):
        prim = o["prim"][sl]
        hit = prim >= 0
        if not hit.any():
            continue
        t12 = np.asarray(tris, np.float32).reshape(-1, 12)
        ids = t12[:, 3].view(np.int32)
        order = np.argsort(ids)
        row = order[np.searchsorted(ids[order], prim[hit])]
        v0 = t12[row, 0:3].astype(np.float64)
        e1 = t12[row, 4:7].astype(np.float64) - v0
        e2 = t12[row, 8:11].astype(np.float64) - v0
        dd = d[hit]
        pv = np.cross(dd, e2)
        det = (e1 * pv).sum(-1)
        tv = cam["org"] - v0
        qv = np.cross(tv, e1)
        t = (e2 * qv).sum(-1) / det
        z64 = t / sw.max_depth
        gn = np.cross(e1, e2)
        gn /= np.linalg.norm(gn, axis=-1, keepdims=True)
        lam = np.maximum(0.0, -(dd * gn).sum(-1))
        cid = t12[row, 7].view(np.uint32)
        uc, inv = np.unique(cid, return_inverse=True)
        rc = np.array([O.random_color(int(c)) for c in uc], np.float64)[inv]
        col = np.clip(np.floor((.2 + rc * lam[:, None]) * 256), 0, 255)
        g8 = d8[r, sl][hit]
        lv = max(int(np.abs(((g8 >> (8 * k)) & 255) - col[:, k]).max()) for k in range(3))
        ez = np.abs(dz[r, sl][hit] - z64)
        worst_z = max(worst_z, float(ez.max()))
        worst_rel = max(worst_rel, float((ez / np.abs(z64)).max()))
        worst_lv = max(worst_lv, lv)
        hits += int(hit.sum())
    return {"sample": f"hit pixels of rows {rows[0]}..{rows[1]}, all ranks ({hits})",
            "max_abs_depth_gpu_vs_fp64": worst_z, "max_rel_depth_gpu_vs_fp64": worst_rel,
            "max_rgb8_levels_gpu_vs_fp64": worst_lv}


def host_threads():
    """Worker threads of the CPU baselines: every CPU this process may run
    on, at most 16 (one GPU's share of the host's cores), and the counts it
    was chosen from."""
    nproc = os.cpu_count() or 1
    allowed = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else nproc
    return min(16, allowed), {"nproc": nproc, "affinity_cpus": allowed,
                              "cap": "16 threads per GPU (the host's CPUs are shared per GPU)"}


def fp64_blobs128():
    """CPU leg: the 128^3 u8 two-brick volume at 512x512 (the CPU-runnable
    configuration): the GPU frame against the oracle's float32 and float64
    re-renders of both ranks composited in visibility order."""
    import torch
    import aloha_amd as A
    from aloha_amd import RankImages
    from aloha_amd import workloads as WL
    import oracle as O
    threads, _ = host_threads()
    wl = WL.blobs128()
    P, W, H = wl.num_ranks, wl.W, wl.H
    with A.Context(0, stream=torch.cuda.current_stream()) as ctx:
        ids = [ctx.upload_brick(d, v) for d, v in zip(wl.bricks, wl.voxels)]
        ctx.set_tf(wl.tf)
        ctx.set_camera(wl.cam["org"], wl.cam["dir_00"], wl.cam["dir_du"], wl.cam["dir_dv"])
        imgs = RankImages.alloc(P, W, H, rgba=True, rgba8=False, depth=False)
        out = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
        ctx.render_frame(wl.params(), W, H, [[ids[i] for i in r] for r in wl.rank_bricks], imgs,
                         wl.order, out, None)
        torch.cuda.synchronize()
        g = out.cpu().numpy()
    outs = [O.render_rank(wl.cam, wl.tf, [(wl.bricks[i], wl.voxels[i]) for i in wl.rank_bricks[r]],
                          W, H, dt=wl.dt, max_depth=wl.max_depth, want64=True, threads=threads)
            for r in range(P)]
    c32, _ = O.composite_over(np.stack([o["rgba"] for o in outs]), wl.order)
    c64 = O.composite_over_d(np.stack([o["rgba64"] for o in outs]), wl.order)
    mask = sum(o["nsamples"] for o in outs) > 0
    return _fp64_line(g, c32, c64, mask, "whole frame")


def cpu_baseline(wl, imgs, seconds):
    """The oracle (FP32 CPU restatement, OpenMP) on a bounded band of rows of
    every rank, on this host's cores; also the FP64 error of both paths on
    that band."""
    import oracle as O
    threads, host = host_threads()
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
                     f"samples, {el:.1f}s), FP32 oracle, OpenMP",
           "host": host}
    parity = {"rows": f"{r0[0]}..{r0[1]}", "max_abs_gpu_vs_cpu_fp32": max_ulp,
              "max_abs_gpu_vs_fp64": err_gpu, "max_abs_cpu_fp32_vs_fp64": err_cpu}
    return cpu, parity


if __name__ == "__main__":
    main()
