"""Helpers shared by the GPU parity tests: run a workload through the C ABI
(device) and through the oracle (host) on the same inputs."""
from __future__ import annotations

import numpy as np

import aloha_amd as A
import oracle as O
from aloha_amd import RankImages


def poison(imgs):
    """Fill freshly allocated output images with values no render writes
    (NaN colours and depths, all-ones words), so a pixel the kernels skip
    fails the comparison instead of passing on stale memory."""
    for name in ("rgba", "depth"):
        t = getattr(imgs, name, None)
        if t is not None:
            t.fill_(float("nan"))
    for name in ("rgba8", "nsamples"):
        t = getattr(imgs, name, None)
        if t is not None:
            t.fill_(-1)
    return imgs


def gpu_render(wl, ranks=None, P=None, *, accum=None, lanes=1, tf=None, **params):
    """Render every rank of `ranks` (brick-id lists; default: the workload's
    kd assignment) on the GPU with `lanes` lanes per ray.  Returns numpy
    arrays (P,H,W[,4])."""
    import torch
    ranks = wl.rank_bricks if ranks is None else ranks
    P = len(ranks)
    with A.Context(0, stream=torch.cuda.current_stream()) as ctx:
        ctx.set_march_schedule(lanes, lanes)
        ids = [ctx.upload_brick(d, v) for d, v in zip(wl.bricks, wl.voxels)]
        ctx.set_tf(wl.tf if tf is None else tf)
        ctx.set_camera(wl.cam["org"], wl.cam["dir_00"], wl.cam["dir_du"], wl.cam["dir_dv"])
        out = poison(RankImages.alloc(P, wl.W, wl.H, accum=True, nsamples=True))
        if accum is not None:
            out.accum.copy_(torch.from_numpy(accum))
        ctx.render_ranks(wl.params(**params), wl.W, wl.H, [[ids[i] for i in r] for r in ranks], out)
        torch.cuda.synchronize()
        res = {
            "rgba": out.rgba.cpu().numpy(),
            "rgba8": out.rgba8.cpu().numpy().view(np.uint32),
            "depth": out.depth.cpu().numpy(),
            "accum": out.accum.cpu().numpy(),
            "nsamples": out.nsamples.cpu().numpy().view(np.uint32),
        }
    return res


def oracle_render(wl, ranks=None, *, accum=None, want64=False, **params):
    ranks = wl.rank_bricks if ranks is None else ranks
    outs = []
    p = dict(dt=wl.dt, max_depth=wl.max_depth)
    p.update(params)
    flags = p.pop("flags", 0)
    heat_clock, heat_scale = p.pop("heat_clock", 0), p.pop("heat_map_scale", 1e-5)
    if flags & A.FLAG_HEATMAP:  # only a host-given clock delta is reproducible
        assert flags & A.FLAG_HEAT_CLOCK
        p.update(heat_clock=heat_clock, heat_scale=heat_scale)
    for r, bl in enumerate(ranks):
        o = O.render_rank(wl.cam, wl.tf, [(wl.bricks[i], wl.voxels[i]) for i in bl], wl.W, wl.H,
                          crosshair=bool(flags & A.FLAG_CROSSHAIR),
                          accum=None if accum is None else accum[r], want64=want64, **p)
        outs.append(o)
    keys = outs[0].keys()
    return {k: np.stack([o[k] for o in outs]) for k in keys if k != "touched"}


def compare(gpu, ref, tol=1e-4):
    """Max abs differences; asserts the path's tolerance (1e-4 float, 1/255 8-bit,
    exact sample counts and depths)."""
    d_rgba = float(np.abs(gpu["rgba"] - ref["rgba"]).max())
    c8 = lambda u, s: ((u >> s) & 255).astype(np.int32)
    d8 = max(int(np.abs(c8(gpu["rgba8"], s) - c8(ref["rgba8"], s)).max()) for s in (0, 8, 16, 24))
    d_depth = float(np.abs(gpu["depth"] - ref["depth"]).max())
    ns_bad = int((gpu["nsamples"] != ref["nsamples"]).sum())
    exact = float((gpu["rgba"] == ref["rgba"]).mean())
    info = dict(max_abs_rgba=d_rgba, max_rgba8=d8, max_depth=d_depth, nsamples_mismatch=ns_bad,
                exact_fraction=exact, samples=int(ref["nsamples"].sum()))
    assert d_rgba <= tol, info
    assert d8 <= 1, info
    assert d_depth == 0.0, info
    assert ns_bad == 0, info
    return info
