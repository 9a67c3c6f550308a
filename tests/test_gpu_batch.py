"""GPU checks of the batched form (aloha_batch_*): every schedule the
autotuner can pick, and the captured hipGraph replayed several times, give
each view the same image, bit for bit, as the single-view fused frame -- and
the per-view plain frame (rank images + composite) through the oracle."""
import numpy as np
import pytest

import aloha_amd as A
import oracle as O
from aloha_amd import workloads as WL

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def wl():
    return WL.batched(n_side=2, cell=32, smin=8, views=4, W=96, H=80)


def _setup(wl, torch):
    ctx = A.Context(0, stream=torch.cuda.current_stream())
    ids = [ctx.upload_brick(d, v) for d, v in zip(wl.bricks, wl.voxels)]
    ctx.set_tf(wl.tf)
    ranks = [[ids[i] for i in r] for r in wl.rank_bricks]
    return ctx, ranks


def _single_views(ctx, wl, ranks, torch):
    refs = []
    for cam, order in zip(wl.cams, wl.orders):
        ctx.set_camera(cam["org"], cam["dir_00"], cam["dir_du"], cam["dir_dv"])
        out = torch.empty((wl.H, wl.W, 4), dtype=torch.float32, device="cuda")
        ctx.render_composite(wl.params(), wl.W, wl.H, ranks, order, out, None)
        refs.append(out)
    torch.cuda.synchronize()
    return torch.stack(refs)


def test_every_schedule_and_graph_replay_bit_exact(wl):
    import torch
    print("\n[case] batch 8 bricks x 4 views, all schedules + graph", flush=True)
    ctx, ranks = _setup(wl, torch)
    ref = _single_views(ctx, wl, ranks, torch)
    V = len(wl.cams)
    out = torch.empty((V, wl.H, wl.W, 4), dtype=torch.float32, device="cuda")
    b = A.Batch(ctx, wl.params(), wl.W, wl.H, wl.cams, ranks, wl.orders, out_rgba=out)
    for tile_w in (8, 16, 32):
        b.set_tile(tile_w)
        for lanes in (1, 2, 4, 8):
            for blocks in (0, 8, 64):
                out.fill_(-1.0)
                b.set_schedule(lanes, blocks)
                b.launch()
                torch.cuda.synchronize()
                assert torch.equal(out, ref), (tile_w, lanes, blocks)
    with pytest.raises(A.AlohaError):
        b.set_tile(12)
    best, table = b.autotune(reps=2)
    assert len(table) == 48 and best > 0
    assert {r[0] for r in table} == {8, 16, 32}
    assert b.schedule()["tile_w"] in (8, 16, 32)
    b.capture()
    assert b.schedule()["captured"]
    for rep in range(3):
        out.fill_(-1.0)
        torch.cuda.synchronize()
        b.launch()
        torch.cuda.synchronize()
        assert torch.equal(out, ref), rep
    b.close()
    ctx.close()


def test_batch_view_matches_oracle(wl):
    import torch
    print("\n[case] batch view 2 vs oracle rank images + composite", flush=True)
    ctx, ranks = _setup(wl, torch)
    V = len(wl.cams)
    out = torch.empty((V, wl.H, wl.W, 4), dtype=torch.float32, device="cuda")
    b = A.Batch(ctx, wl.params(), wl.W, wl.H, wl.cams, ranks, wl.orders, out_rgba=out)
    b.launch()
    torch.cuda.synchronize()
    v = 2
    imgs = [O.render_rank(wl.cams[v], wl.tf, [(wl.bricks[i], wl.voxels[i]) for i in r], wl.W,
                          wl.H, dt=wl.dt)["rgba"] for r in wl.rank_bricks]
    ref, _ = O.composite_over(np.stack(imgs), wl.orders[v])
    got = out[v].cpu().numpy()
    assert np.abs(got - ref).max() <= 1e-4
    assert ref[..., 3].max() > 0.3
    b.close()
    ctx.close()


def test_batch_every_output_spp_islands(wl):
    """RGBA, RGBA8 and sample counts of every view, with two samples per
    pixel and the island mask, equal the single-view fused frames bit for bit
    under every lanes-per-ray schedule."""
    import torch
    print("\n[case] batch 8 bricks x 4 views, spp 2, island 1 of 2, all outputs", flush=True)
    ctx, ranks = _setup(wl, torch)
    prm = wl.params(spp=2, num_islands=2, island_id=1)
    V, H, W = len(wl.cams), wl.H, wl.W
    ref = {"rgba": torch.empty((V, H, W, 4), dtype=torch.float32, device="cuda"),
           "rgba8": torch.empty((V, H, W), dtype=torch.int32, device="cuda"),
           "ns": torch.empty((V, H, W), dtype=torch.int32, device="cuda")}
    for v, (cam, order) in enumerate(zip(wl.cams, wl.orders)):
        ctx.set_camera(cam["org"], cam["dir_00"], cam["dir_du"], cam["dir_dv"])
        ctx.render_composite(prm, W, H, ranks, order, ref["rgba"][v], ref["rgba8"][v],
                             ref["ns"][v])
    torch.cuda.synchronize()
    assert int(ref["ns"].sum()) > 0 and float(ref["rgba"][..., 3].max()) > 0.3
    got = {k: torch.empty_like(t) for k, t in ref.items()}
    b = A.Batch(ctx, prm, W, H, wl.cams, ranks, wl.orders, out_rgba=got["rgba"],
                out_rgba8=got["rgba8"], nsamples=got["ns"])
    for tile_w, lanes in ((16, 1), (16, 2), (8, 4), (32, 8), (8, 1), (32, 4)):
        b.set_tile(tile_w)
        for blocks in (0, 64):
            for t in got.values():
                t.fill_(-1)
            b.set_schedule(lanes, blocks)
            b.launch()
            torch.cuda.synchronize()
            for k in ref:
                assert torch.equal(got[k], ref[k]), (k, tile_w, lanes, blocks)
    b.close()
    ctx.close()


def test_batch_multi_brick_and_empty_ranks(wl):
    """Rank slots of several bricks (the brick loop inside a slot) and an
    empty slot, with a different visibility order per view: every view
    equals its single-view fused frame bit for bit, persistent or not."""
    import torch
    print("\n[case] batch 8 bricks as slots of 3 / 0 / 5 bricks x 4 views", flush=True)
    ctx, ranks1 = _setup(wl, torch)
    ids = [i for r in ranks1 for i in r]
    assert len(ids) == 8
    ranks = [ids[:3], [], ids[3:]]
    orders = [[0, 1, 2], [2, 0, 1], [1, 2, 0], [2, 1, 0]][:len(wl.cams)]
    V, H, W = len(wl.cams), wl.H, wl.W
    ref = torch.empty((V, H, W, 4), dtype=torch.float32, device="cuda")
    ref_ns = torch.empty((V, H, W), dtype=torch.int32, device="cuda")
    for v, (cam, order) in enumerate(zip(wl.cams, orders)):
        ctx.set_camera(cam["org"], cam["dir_00"], cam["dir_du"], cam["dir_dv"])
        ctx.render_composite(wl.params(), W, H, ranks, order, ref[v], None, ref_ns[v])
    torch.cuda.synchronize()
    assert float(ref[..., 3].max()) > 0.3
    out = torch.empty_like(ref)
    ns = torch.empty_like(ref_ns)
    b = A.Batch(ctx, wl.params(), W, H, wl.cams, ranks, orders, out_rgba=out, nsamples=ns)
    for tile_w, lanes, blocks in ((16, 1, 0), (8, 4, 0), (32, 4, 64), (8, 8, 8)):
        out.fill_(float("nan"))
        ns.fill_(-1)
        b.set_tile(tile_w)
        b.set_schedule(lanes, blocks)
        b.launch()
        torch.cuda.synchronize()
        assert torch.equal(out, ref) and torch.equal(ns, ref_ns), (lanes, blocks)
    b.close()
    ctx.close()


def test_batch_follows_tf_and_brick_table_changes(wl):
    """A TF of another size (a new device buffer) and more bricks (a new
    device brick table) after the batch was made and captured: the next
    launch takes the current ones and equals the single-view frames."""
    import torch
    print("\n[case] batch after TF resize and brick upload", flush=True)
    ctx, ranks = _setup(wl, torch)
    V = len(wl.cams)
    out = torch.empty((V, wl.H, wl.W, 4), dtype=torch.float32, device="cuda")
    b = A.Batch(ctx, wl.params(), wl.W, wl.H, wl.cams, ranks, wl.orders, out_rgba=out)
    b.capture()
    b.launch()
    torch.cuda.synchronize()
    # a 3x longer TF through a larger allocation, and 40 extra bricks, which
    # move the context's brick table to a larger allocation
    tf2 = np.repeat(wl.tf[::-1], 3, axis=0).copy()
    ctx.set_tf(tf2)
    for _ in range(40):
        ctx.upload_brick(wl.bricks[0], wl.voxels[0])
    ref = _single_views(ctx, wl, ranks, torch)
    for rep in range(2):
        out.fill_(-1.0)
        torch.cuda.synchronize()
        b.launch()
        torch.cuda.synchronize()
        assert torch.equal(out, ref), rep
    assert b.schedule()["captured"]
    # a smaller TF in the same allocation changes only the entry count
    ctx.set_tf(wl.tf[:64])
    ref = _single_views(ctx, wl, ranks, torch)
    out.fill_(-1.0)
    b.launch()
    torch.cuda.synchronize()
    assert torch.equal(out, ref)
    b.close()
    ctx.close()


def test_batch_outliving_its_context_is_refused(wl):
    import torch
    print("\n[case] batch after its context is closed", flush=True)
    ctx, ranks = _setup(wl, torch)
    V = len(wl.cams)
    out = torch.empty((V, wl.H, wl.W, 4), dtype=torch.float32, device="cuda")
    b = A.Batch(ctx, wl.params(), wl.W, wl.H, wl.cams, ranks, wl.orders, out_rgba=out)
    b.launch()
    ctx.close()
    with pytest.raises(A.AlohaError):
        b.launch()
    b.close()
