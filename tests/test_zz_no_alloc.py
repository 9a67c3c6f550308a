"""aloha_render_frame_noalloc -- the frame without caller-held rank images,
whose rank images live in the library's per-GPU workspace -- against
aloha_render_frame through the C ABI: the same outputs bit for bit on calls
in a row (the later ones smaller, so stale workspace contents would show),
and the host-side refusals (workspace too small, accumulation, heat map)."""
import numpy as np
import pytest

import aloha_amd as A
from aloha_amd import RankImages
from aloha_amd import workloads as WL

pytestmark = pytest.mark.gpu


def _upload(ctx, wl):
    ids = [ctx.upload_brick(d, v) for d, v in zip(wl.bricks, wl.voxels)]
    return [[ids[i] for i in r] for r in wl.rank_bricks]


def _both(ctx, wl, ranks, torch, **kw):
    ctx.set_tf(wl.tf)
    ctx.set_camera(wl.cam["org"], wl.cam["dir_00"], wl.cam["dir_du"], wl.cam["dir_dv"])
    p = wl.params(**kw)
    P, W, H = len(ranks), wl.W, wl.H
    imgs = RankImages.alloc(P, W, H)
    ref = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
    ref8 = torch.empty((H, W), dtype=torch.int32, device="cuda")
    ctx.render_frame(p, W, H, ranks, imgs, wl.order, ref, ref8)
    got = torch.full((H, W, 4), float("nan"), dtype=torch.float32, device="cuda")
    got8 = torch.full((H, W), -1, dtype=torch.int32, device="cuda")
    ctx.render_frame_noalloc(p, W, H, ranks, wl.order, got, got8)
    torch.cuda.synchronize()
    assert float(ref[..., 3].max()) > 0.1  # something was rendered
    assert torch.equal(got, ref)
    assert torch.equal(got8, ref8)
    rects = ctx.rank_rects(p, W, H, ranks)
    return sum(max(0, x1 - x0) * max(0, y1 - y0) for x0, y0, x1, y1 in rects)


def test_noalloc_frame_equals_render_frame_on_calls_in_a_row():
    import torch
    print("\n[case] noalloc frame: fbm 128^3 8 ranks 384x320, then blobs 2 ranks 160x120",
          flush=True)
    with A.Context(0, stream=torch.cuda.current_stream()) as ctx:
        big = WL.fbm512(num_ranks=8, W=384, H=320, size=128)
        small = WL.blobs128(num_ranks=2, W=160, H=120, size=48)
        rb, rs = _upload(ctx, big), _upload(ctx, small)
        n1 = _both(ctx, big, rb, torch)
        n2 = _both(ctx, small, rs, torch)
        assert n1 > n2 > 0
        # spp 2 with the crosshair (it changes only rank RGBA8, which this form
        # does not write) and a fixed jitter, then the plain frame again
        _both(ctx, big, rb, torch, spp=2, flags=A.FLAG_CROSSHAIR)
        _both(ctx, small, rs, torch, jitter_mode=A.JITTER_FIXED, jitter_value=0.25)


def test_noalloc_frame_refusals():
    import torch
    print("\n[case] noalloc frame: refusals on the host", flush=True)
    with A.Context(0, stream=torch.cuda.current_stream()) as ctx:
        wl = WL.blobs128(num_ranks=2, W=2048, H=2048, size=32)
        ranks = _upload(ctx, wl)
        ctx.set_tf(wl.tf)
        # camera inside the volume: both ranks' rectangles are the whole
        # 2048^2 image, 2 x 4 Mi pixels x 16 B = 128 MiB > the 64 MiB workspace
        c = wl.cam
        ctx.set_camera([16.0, 16.0, 16.0], c["dir_00"], c["dir_du"], c["dir_dv"])
        out = torch.empty((2048, 2048, 4), dtype=torch.float32, device="cuda")
        with pytest.raises(A.AlohaError) as e:
            ctx.render_frame_noalloc(wl.params(), 2048, 2048, ranks, wl.order, out, None)
        assert e.value.code == A.ERR_LIMIT
        for kw in ({"accum_id": 1}, {"flags": A.FLAG_HEATMAP}):
            with pytest.raises(A.AlohaError) as e:
                ctx.render_frame_noalloc(wl.params(**kw), 256, 256, ranks, wl.order, out, None)
            assert e.value.code == A.ERR_ARG
        with pytest.raises(ValueError):
            ctx.render_frame_noalloc(wl.params(), 256, 256, ranks, [0], out, None)
        # the context is still usable afterwards
        ctx.set_camera(c["org"], c["dir_00"], c["dir_du"], c["dir_dv"])
        ctx.render_frame_noalloc(wl.params(), 256, 256, ranks, wl.order, out, None)
        torch.cuda.synchronize()
