"""GPU checks of u8 brick storage: u8 voxels kept as bytes in HBM and decoded
at sample time through the table u/255.f give every output -- rank images,
the frame, the fused and batched forms, the footprint -- bit for bit as the
same bricks widened to float32 at upload (Volume::get's decode,
volume.h:78), in a quarter of the brick memory; updates keep the storage,
and the differentiable form refuses bricks it cannot differentiate."""
import numpy as np
import pytest

import aloha_amd as A
import oracle as O
from aloha_amd import RankImages
from aloha_amd import workloads as WL

from gpu_util import poison

pytestmark = pytest.mark.gpu


def _ctx(wl, keep_u8, torch):
    ctx = A.Context(0, stream=torch.cuda.current_stream())
    ctx.set_u8_storage(keep_u8)
    ids = [ctx.upload_brick(d, v) for d, v in zip(wl.bricks, wl.voxels)]
    ctx.set_tf(wl.tf)
    return ctx, ids


def _frame_outputs(ctx, wl, ranks, torch, **params):
    """Every volume entry point of one camera: rank images (all outputs),
    the frame with and without rank images, the fused form, the footprint."""
    W, H, P = wl.W, wl.H, len(ranks)
    order = wl.order if len(wl.order) == P else list(range(P))
    p = wl.params(**params)
    imgs = poison(RankImages.alloc(P, W, H, accum=True, nsamples=True))
    ctx.render_ranks(p, W, H, ranks, imgs)
    frame = torch.full((H, W, 4), float("nan"), device="cuda")
    frame8 = torch.full((H, W), -1, dtype=torch.int32, device="cuda")
    ctx.render_frame(p, W, H, ranks, None, order, frame, frame8)
    fused = torch.full((H, W, 4), float("nan"), device="cuda")
    if p.spp == 1 and p.flags == 0:
        ctx.render_composite(p, W, H, ranks, order, fused, None)
    torch.cuda.synchronize()
    out = {k: getattr(imgs, k).cpu().numpy().copy() for k in ("rgba", "rgba8", "depth", "accum",
                                                                  "nsamples")}
    out["frame"] = frame.cpu().numpy()
    out["frame8"] = frame8.cpu().numpy()
    out["fused"] = fused.cpu().numpy()
    out["touched"] = ctx.footprint_bricks(p, W, H, ranks)
    return out


@pytest.mark.parametrize("lanes", [1, 4])
def test_u8_storage_bit_identical_to_f32(lanes):
    import torch
    wl = WL.blobs128(num_ranks=2, W=160, H=144, size=64)
    print(f"\n[case] blobs64 u8 bricks, storage u8 vs f32, lanes {lanes}", flush=True)
    res = {}
    for keep in (True, False):
        ctx, ids = _ctx(wl, keep, torch)
        ctx.set_march_schedule(lanes, lanes)
        assert [ctx.brick_dtype(i) for i in ids] == [A.U8 if keep else A.F32] * len(ids)
        ranks = [[ids[i] for i in r] for r in wl.rank_bricks]
        ctx.set_camera(wl.cam["org"], wl.cam["dir_00"], wl.cam["dir_du"], wl.cam["dir_dv"])
        res[keep] = (_frame_outputs(ctx, wl, ranks, torch),
                     _frame_outputs(ctx, wl, ranks, torch, spp=2, jitter_mode=A.JITTER_RANDOM),
                     ctx.brick_bytes())
        # one rank over both bricks (the brick loop inside a rank)
        res[keep] += (_frame_outputs(ctx, wl, [[ids[0], ids[1]]], torch),)
        ctx.close()
    u8, f32 = res[True], res[False]
    for a, b in ((u8[0], f32[0]), (u8[1], f32[1]), (u8[3], f32[3])):
        for k in a:
            assert np.array_equal(a[k], b[k], equal_nan=True), k
    assert u8[0]["nsamples"].sum() > 0
    # and the oracle's rank images (float32 restatement of Volume::get + march)
    for r, br in enumerate(wl.rank_bricks):
        o = O.render_rank(wl.cam, wl.tf, [(wl.bricks[i], wl.voxels[i]) for i in br], wl.W, wl.H,
                          dt=wl.dt, max_depth=wl.max_depth)
        assert np.array_equal(u8[0]["nsamples"][r], o["nsamples"])
        assert float(np.abs(u8[0]["rgba"][r] - o["rgba"]).max()) <= 1e-4
    # a quarter of the padded voxel bytes (+ 16 bytes of read slack per brick)
    assert (u8[2] - 16 * len(wl.bricks)) * 4 == f32[2]


def test_u8_storage_batched_bit_identical():
    import torch
    wl = WL.batched(n_side=2, cell=32, smin=8, views=3, W=96, H=80)
    assert any(v.dtype == np.uint8 for v in wl.voxels) and any(v.dtype == np.float32
                                                               for v in wl.voxels)
    print("\n[case] batched 8 mixed u8/f32 bricks x 3 views, storage u8 vs f32", flush=True)
    outs = {}
    for keep in (True, False):
        ctx, ids = _ctx(wl, keep, torch)
        ranks = [[ids[i] for i in r] for r in wl.rank_bricks]
        V = len(wl.cams)
        out = torch.empty((V, wl.H, wl.W, 4), dtype=torch.float32, device="cuda")
        ns = torch.empty((V, wl.H, wl.W), dtype=torch.int32, device="cuda")
        b = A.Batch(ctx, wl.params(), wl.W, wl.H, wl.cams, ranks, wl.orders, out_rgba=out,
                    nsamples=ns)
        per = []
        for lanes in (1, 4):
            for blocks in (0, 16):
                out.fill_(-1.0)
                b.set_schedule(lanes, blocks)
                b.launch()
                torch.cuda.synchronize()
                per.append((out.cpu().numpy().copy(), ns.cpu().numpy().copy()))
        outs[keep] = per
        b.close()
        ctx.close()
    for (a, an), (c, cn) in zip(outs[True], outs[False]):
        assert np.array_equal(a, c) and np.array_equal(an, cn)
    assert outs[True][0][1].sum() > 0


def test_u8_brick_update_and_refusals():
    import torch
    wl = WL.blobs128(num_ranks=1, W=96, H=96, size=48)
    print("\n[case] u8 brick update, f32 update refused, differentiable form refused",
          flush=True)
    ctx, ids = _ctx(wl, True, torch)
    ranks = [[ids[i] for i in r] for r in wl.rank_bricks]
    ctx.set_camera(wl.cam["org"], wl.cam["dir_00"], wl.cam["dir_du"], wl.cam["dir_dv"])
    before = _frame_outputs(ctx, wl, ranks, torch)
    # brick 0's voxels inverted in place: the frame equals a fresh context's
    # upload of the inverted brick
    v0 = wl.voxels[0]
    swapped = (255 - v0).astype(np.uint8)
    ctx.brick_update(ids[0], torch.from_numpy(swapped).cuda())
    after = _frame_outputs(ctx, wl, ranks, torch)
    ref_ctx = A.Context(0, stream=torch.cuda.current_stream())
    rid = [ref_ctx.upload_brick(wl.bricks[0], swapped)] + [
        ref_ctx.upload_brick(d, v) for d, v in zip(wl.bricks[1:], wl.voxels[1:])]
    ref_ctx.set_tf(wl.tf)
    ref_ctx.set_camera(wl.cam["org"], wl.cam["dir_00"], wl.cam["dir_du"], wl.cam["dir_dv"])
    ref = _frame_outputs(ref_ctx, wl, [[rid[i] for i in r] for r in wl.rank_bricks], torch)
    ref_ctx.close()
    for k in ref:
        assert np.array_equal(after[k], ref[k], equal_nan=True), k
    assert not np.array_equal(after["frame"], before["frame"])
    # a float update of a brick kept as u8 is refused and changes nothing
    with pytest.raises(A.AlohaError, match="u8"):
        ctx.brick_update(ids[0], torch.zeros(v0.shape, dtype=torch.float32, device="cuda"))
    again = _frame_outputs(ctx, wl, ranks, torch)
    for k in after:
        assert np.array_equal(again[k], after[k], equal_nan=True), k
    # the differentiable form needs float32 bricks
    import ctypes as C
    from aloha_amd._lib import lib
    V, W, H = 1, 32, 32
    cams = np.concatenate([np.asarray(wl.cam[k], np.float32)
                           for k in ("org", "dir_00", "dir_du", "dir_dv")])
    rgba = torch.empty((V, H, W, 4), dtype=torch.float32, device="cuda")
    nsm = torch.empty((V, H, W), dtype=torch.int32, device="cuda")
    prm = wl.params()
    idarr = (C.c_int32 * 1)(ids[0])
    rc = lib().aloha_diff_forward(ctx._h, C.byref(prm), W, H, V,
                                  cams.ctypes.data_as(C.POINTER(C.c_float)), 1, idarr,
                                  C.c_void_p(rgba.data_ptr()), C.c_void_p(nsm.data_ptr()))
    assert rc == -1 and b"float32" in lib().aloha_last_error(ctx._h)
    ctx.close()
