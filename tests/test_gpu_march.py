"""GPU parity of the per-brick ray march (renderFrame/integrateBrick,
deviceCode.cpp:340-416, :256-298) against the CPU oracle, through the C ABI.

Bar: sample counts and depths identical, float RGBA within 1e-4 (the kernels
are built to match the oracle bit for bit, so the observed difference is
reported too), 8-bit within 1/255.
"""
import numpy as np
import pytest

import aloha_amd as A
from aloha_amd import workloads as WL

from gpu_util import compare, gpu_render, oracle_render, poison

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c0():
    return WL.blobs128()


@pytest.fixture(scope="module")
def c1():
    return WL.fbm512()


def _case(name):
    print(f"\n[case] {name}", flush=True)


def test_c0_two_ranks_bit_exact(c0):
    _case("blobs128 P=2 512^2 random jitter")
    g = gpu_render(c0)
    r = oracle_render(c0)
    info = compare(g, r)
    print(info)
    assert info["samples"] > 100000
    assert info["exact_fraction"] == 1.0


def test_c0_one_rank_brick_loop(c0):
    _case("blobs128 P=1 (brick loop)")
    ranks = [[0, 1]]
    info = compare(gpu_render(c0, ranks), oracle_render(c0, ranks))
    print(info)


def test_c0_fixed_jitter_dt2_crosshair(c0):
    _case("blobs128 fixed jitter dt=2 crosshair")
    kw = dict(jitter_mode=A.JITTER_FIXED, jitter_value=0.25, dt=2.0, flags=A.FLAG_CROSSHAIR)
    print(compare(gpu_render(c0, **kw), oracle_render(c0, **kw)))


def test_c0_heat_map_host_clock(c0):
    _case("blobs128 heat map with a host-given clock delta, spp=2, crosshair")
    kw = dict(flags=A.FLAG_HEATMAP | A.FLAG_HEAT_CLOCK | A.FLAG_CROSSHAIR, heat_clock=123457,
              heat_map_scale=2e-6, spp=2)
    g = gpu_render(c0, **kw)
    r = oracle_render(c0, **kw)
    info = compare(g, r)
    print(info)
    assert info["exact_fraction"] == 1.0
    assert (r["rgba"][..., 3] >= 0.25).all()  # every pixel carries the heat blend


def test_c0_spp_accum_islands(c0):
    _case("blobs128 spp=3 accum_id=2 islands 1/3")
    acc = np.random.default_rng(3).random((2, c0.H, c0.W, 4), dtype=np.float32)
    kw = dict(spp=3, accum_id=2, island_id=1, num_islands=3)
    g = gpu_render(c0, accum=acc, **kw)
    r = oracle_render(c0, accum=acc, **kw)
    print(compare(g, r))
    assert np.abs(g["accum"] - r["accum"]).max() <= 1e-4


def test_c1_eight_ranks(c1):
    _case("fbm512 P=8 1024^2")
    g = gpu_render(c1)
    r = oracle_render(c1)
    info = compare(g, r)
    print(info)
    assert info["samples"] > 1_000_000


def test_c1_one_rank_all_bricks(c1):
    _case("fbm512 P=1 1024^2 (8-brick loop)")
    ranks = [list(range(8))]
    print(compare(gpu_render(c1, ranks), oracle_render(c1, ranks)))


def test_small_bricks_and_u16():
    _case("tiny bricks 3^3 u16, camera inside volume")
    raw = (np.arange(6 * 5 * 4, dtype=np.uint16).reshape(4, 5, 6) * 997) % 65535
    wl = WL.make_volume_workload("tiny", raw.astype(np.uint16), 4, 2, 40, 30)
    # camera inside the volume, near-axis view direction (1e-5 clamp)
    wl.cam = {"org": np.array([2.5, 2.0, 1.5], np.float32),
              "dir_00": np.array([-0.3, -0.3, 1.0], np.float32),
              "dir_du": np.array([0.015, 0.0, 0.0], np.float32),
              "dir_dv": np.array([0.0, 0.02, 0.0], np.float32)}
    wl.max_depth = A.max_depth(wl.bounds, wl.cam["org"])
    print(compare(gpu_render(wl, dt=0.25), oracle_render(wl, dt=0.25)))


@pytest.mark.parametrize("lanes", [1, 2, 4, 8])
def test_many_ranks_odd_image(lanes):
    """The persistent per-rank launch's work split: 13 ranks over 27 uneven
    bricks, a 203x117 image (rectangle rows that are no multiple of a run,
    rectangles cut by the image border), every lanes-per-ray schedule --
    bit-exact against the oracle, and every pixel written once."""
    _case(f"blobs 72^3 in 27 bricks, 13 ranks, 203x117, {lanes} lanes per ray")
    raw = WL.synth_blobs_u8((72, 60, 54), 12, 0x5EED7, 3.0, 10.0, 0.3, 1.0)
    wl = WL.make_volume_workload("blobs_odd", raw, 27, 13, 203, 117, cam_seed=0x5EED8)
    g = gpu_render(wl, lanes=lanes)
    r = oracle_render(wl)
    info = compare(g, r)
    print(info)
    assert info["exact_fraction"] == 1.0
    assert info["samples"] > 10_000
    # the fused launch on the same split equals ranks + composite bit for bit
    import torch
    with A.Context(0, stream=torch.cuda.current_stream()) as ctx:
        ctx.set_march_schedule(lanes, lanes)
        ids = [ctx.upload_brick(d, v) for d, v in zip(wl.bricks, wl.voxels)]
        ctx.set_tf(wl.tf)
        ctx.set_camera(wl.cam["org"], wl.cam["dir_00"], wl.cam["dir_du"], wl.cam["dir_dv"])
        ranks = [[ids[i] for i in rb] for rb in wl.rank_bricks]
        imgs = poison(A.RankImages.alloc(wl.num_ranks, wl.W, wl.H))
        ref = torch.empty((wl.H, wl.W, 4), dtype=torch.float32, device="cuda")
        out = torch.empty_like(ref)
        ctx.render_ranks(wl.params(), wl.W, wl.H, ranks, imgs)
        ctx.composite_over(imgs.rgba, wl.order, wl.W, wl.H, ref, None)
        ctx.render_composite(wl.params(), wl.W, wl.H, ranks, wl.order, out, None)
        torch.cuda.synchronize()
        assert torch.equal(out, ref)


@pytest.mark.parametrize("which", ["c0", "c1", "c0spp", "c1one", "c0isl"])
def test_fused_render_composite_bit_exact(which, c0, c1):
    """The fused launch (rays set up per thread into LDS records, lane groups,
    in-tile blend) equals render_ranks + composite bit for bit, also with
    spp > 1 (records that start in a later spp sample), one rank over all
    bricks (re-selection inside the march) and the island mask."""
    import torch
    wl = c1 if which.startswith("c1") else c0
    _case(f"{wl.name} fused render+composite vs ranks+composite ({which})")
    kw = {"c0spp": dict(spp=3), "c1one": dict(spp=2),
          "c0isl": dict(spp=2, num_islands=3, island_id=1)}.get(which, {})
    params = wl.params(**kw)
    out = torch.empty((wl.H, wl.W, 4), dtype=torch.float32, device="cuda")
    out8 = torch.empty((wl.H, wl.W), dtype=torch.int32, device="cuda")
    ns = torch.empty((wl.H, wl.W), dtype=torch.int32, device="cuda")
    ref = torch.empty_like(out)
    ref8 = torch.empty_like(out8)
    with A.Context(0, stream=torch.cuda.current_stream()) as ctx:
        ids = [ctx.upload_brick(d, v) for d, v in zip(wl.bricks, wl.voxels)]
        ctx.set_tf(wl.tf)
        ctx.set_camera(wl.cam["org"], wl.cam["dir_00"], wl.cam["dir_du"], wl.cam["dir_dv"])
        ranks = [[ids[i] for i in r] for r in wl.rank_bricks]
        order = wl.order
        if which == "c1one":  # one rank marching every brick
            ranks, order = [ids], [0]
        imgs = poison(A.RankImages.alloc(len(ranks), wl.W, wl.H, nsamples=True))
        ctx.render_ranks(params, wl.W, wl.H, ranks, imgs)
        ctx.composite_over(imgs.rgba, order, wl.W, wl.H, ref, ref8)
        ctx.render_composite(params, wl.W, wl.H, ranks, order, out, out8, ns)
        torch.cuda.synchronize()
        assert torch.equal(out, ref)
        assert torch.equal(out8, ref8)
        assert int(ns.sum()) == int(imgs.nsamples.sum()) > 0


@pytest.mark.parametrize("which", ["c0", "c1", "odd", "heat"])
def test_sparse_composite_equals_dense(which, c0, c1):
    """The sparse composite over aloha_rank_rects' rectangles gives the dense
    composite of render_ranks' images bit for bit (float and RGBA8); with the
    heat map on, the rectangles are the whole image."""
    import torch
    if which == "odd":
        raw = WL.synth_blobs_u8((72, 60, 54), 12, 0x5EED7, 3.0, 10.0, 0.3, 1.0)
        wl = WL.make_volume_workload("blobs_odd", raw, 27, 13, 203, 117, cam_seed=0x5EED8)
    else:
        wl = c1 if which == "c1" else c0
    _case(f"{wl.name} sparse vs dense composite ({which})")
    params = wl.params(flags=A.FLAG_HEATMAP) if which == "heat" else wl.params()
    with A.Context(0, stream=torch.cuda.current_stream()) as ctx:
        ids = [ctx.upload_brick(d, v) for d, v in zip(wl.bricks, wl.voxels)]
        ctx.set_tf(wl.tf)
        ctx.set_camera(wl.cam["org"], wl.cam["dir_00"], wl.cam["dir_du"], wl.cam["dir_dv"])
        ranks = [[ids[i] for i in r] for r in wl.rank_bricks]
        imgs = poison(A.RankImages.alloc(wl.num_ranks, wl.W, wl.H))
        ctx.render_ranks(params, wl.W, wl.H, ranks, imgs)
        rects = ctx.rank_rects(params, wl.W, wl.H, ranks)
        dense = torch.empty((wl.H, wl.W, 4), dtype=torch.float32, device="cuda")
        dense8 = torch.empty((wl.H, wl.W), dtype=torch.int32, device="cuda")
        sparse, sparse8 = torch.empty_like(dense), torch.empty_like(dense8)
        ctx.composite_over(imgs.rgba, wl.order, wl.W, wl.H, dense, dense8)
        ctx.composite_over(imgs.rgba, wl.order, wl.W, wl.H, sparse, sparse8, rects=rects)
        torch.cuda.synchronize()
        assert torch.equal(sparse, dense) and torch.equal(sparse8, dense8)
        area = sum(max(0, x1 - x0) * max(0, y1 - y0) for x0, y0, x1, y1 in rects)
        if which == "heat":
            assert all(r == (0, 0, wl.W, wl.H) for r in rects)
        else:
            assert area < wl.num_ranks * wl.W * wl.H  # something was skipped
            # outside its rectangle every rank pixel is transparent black
            rgba = imgs.rgba.cpu().numpy()
            for s, (x0, y0, x1, y1) in enumerate(rects):
                m = np.ones((wl.H, wl.W), bool)
                m[max(y0, 0):y1, max(x0, 0):x1] = False
                assert not rgba[s][m].any()


@pytest.mark.parametrize("which", ["c0", "c1", "odd", "crosshair", "accum", "away"])
def test_render_frame_equals_ranks_then_composite(which, c0, c1):
    """render_frame (the per-rank launch also clears the outputs outside the
    rectangles' bounding box; the composite covers only the box) gives
    render_ranks' images and the dense composite of them bit for bit, for
    plain frames, the crosshair and accumulation (whole-image composite; the
    heat map's clock-based colours are not reproducible); two frames in a
    row on the same context; and the frame alone (no caller rank images)
    equal to them."""
    import torch
    if which == "odd":
        raw = WL.synth_blobs_u8((72, 60, 54), 12, 0x5EED7, 3.0, 10.0, 0.3, 1.0)
        wl = WL.make_volume_workload("blobs_odd", raw, 27, 13, 203, 117, cam_seed=0x5EED8)
    else:
        wl = c1 if which == "c1" else c0
    _case(f"{wl.name} render_frame vs render_ranks + composite ({which})")
    kw = {"crosshair": dict(flags=A.FLAG_CROSSHAIR), "accum": dict(accum_id=1, spp=2)}.get(which, {})
    params = wl.params(**kw)
    cam = dict(wl.cam)
    if which == "away":  # image plane turned off the volume: no rectangle, an empty box
        cam["dir_00"] = np.asarray(cam["dir_00"]) - 2.0 * wl.W * np.asarray(cam["dir_du"])
    with A.Context(0, stream=torch.cuda.current_stream()) as ctx:
        ctx.set_march_schedule(4, 4)
        ids = [ctx.upload_brick(d, v) for d, v in zip(wl.bricks, wl.voxels)]
        ctx.set_tf(wl.tf)
        ctx.set_camera(cam["org"], cam["dir_00"], cam["dir_du"], cam["dir_dv"])
        ranks = [[ids[i] for i in r] for r in wl.rank_bricks]
        accum = which == "accum"
        ref = A.RankImages.alloc(wl.num_ranks, wl.W, wl.H, accum=accum, nsamples=True)
        got = poison(A.RankImages.alloc(wl.num_ranks, wl.W, wl.H, accum=accum, nsamples=True))
        if accum:  # the same earlier frame in both
            prev = torch.rand_like(ref.accum)
            ref.accum.copy_(prev)
            got.accum.copy_(prev)
        ctx.render_ranks(params, wl.W, wl.H, ranks, ref)
        dense = torch.empty((wl.H, wl.W, 4), dtype=torch.float32, device="cuda")
        dense8 = torch.empty((wl.H, wl.W), dtype=torch.int32, device="cuda")
        ctx.composite_over(ref.rgba, wl.order, wl.W, wl.H, dense, dense8)
        out = torch.full_like(dense, float("nan"))
        out8 = torch.full_like(dense8, -1)
        for _ in range(2):
            if accum:
                got.accum.copy_(prev)
            ctx.render_frame(params, wl.W, wl.H, ranks, got, wl.order, out, out8)
            torch.cuda.synchronize()
            for name in ("rgba", "rgba8", "depth", "nsamples") + (("accum",) if accum else ()):
                assert torch.equal(getattr(got, name), getattr(ref, name)), name
            assert torch.equal(out, dense) and torch.equal(out8, dense8)
        # the frame only (no rank images from the caller: compact ones in the
        # context's workspace), twice in a row
        out_c = torch.full_like(dense, float("nan"))
        out_c8 = torch.full_like(dense8, -1)
        if accum:  # no accumulation buffer without rank images
            with pytest.raises(A.AlohaError):
                ctx.render_frame(params, wl.W, wl.H, ranks, None, wl.order, out_c, out_c8)
        else:
            for _ in range(2):
                ctx.render_frame(params, wl.W, wl.H, ranks, None, wl.order, out_c, out_c8)
                torch.cuda.synchronize()
                assert torch.equal(out_c, dense) and torch.equal(out_c8, dense8)
        if which == "away":
            rects = ctx.rank_rects(params, wl.W, wl.H, ranks)
            assert all(x0 >= x1 or y0 >= y1 for x0, y0, x1, y1 in rects)
            assert not out.any() and not out8.any()


def test_render_errors(c0):
    """Argument errors of the volume render entry points come back as
    AlohaError (the C status and message), output buffers that are host,
    strided or too small as ValueError from the Python layer, and the
    context stays usable."""
    import torch
    wl = c0
    _case("render entry point errors")
    with A.Context(0, stream=torch.cuda.current_stream()) as ctx:
        ids = [ctx.upload_brick(d, v) for d, v in zip(wl.bricks, wl.voxels)]
        ctx.set_tf(wl.tf)
        ctx.set_camera(wl.cam["org"], wl.cam["dir_00"], wl.cam["dir_du"], wl.cam["dir_dv"])
        ranks = [[ids[i] for i in r] for r in wl.rank_bricks]
        imgs = A.RankImages.alloc(wl.num_ranks, wl.W, wl.H)
        out = torch.empty((wl.H, wl.W, 4), dtype=torch.float32, device="cuda")
        with pytest.raises(A.AlohaError):  # a brick id that was never uploaded
            ctx.render_ranks(wl.params(), wl.W, wl.H, [[len(ids) + 5]], imgs)
        with pytest.raises(A.AlohaError):  # accumulation without the accum buffer
            ctx.render_ranks(wl.params(accum_id=1), wl.W, wl.H, ranks, imgs)
        no_rgba = A.RankImages.alloc(wl.num_ranks, wl.W, wl.H, rgba=False)
        with pytest.raises(A.AlohaError):  # render_frame composites the float images
            ctx.render_frame(wl.params(), wl.W, wl.H, ranks, no_rgba, wl.order, out)
        with pytest.raises(A.AlohaError):  # order is not a permutation of the slots
            ctx.render_frame(wl.params(), wl.W, wl.H, ranks, imgs, [0] * wl.num_ranks, out)
        with pytest.raises(A.AlohaError):
            ctx.render_frame(wl.params(spp=0), wl.W, wl.H, ranks, imgs, wl.order, out)
        with pytest.raises(ValueError):  # an output smaller than W x H: refused before launch
            ctx.render_frame(wl.params(), wl.W, wl.H, ranks, imgs, wl.order, out[:-1])
        with pytest.raises(ValueError):  # rank images for fewer slots than the table has
            ctx.render_ranks(wl.params(), wl.W, wl.H, ranks + [[]],
                             A.RankImages.alloc(wl.num_ranks, wl.W, wl.H))
        with pytest.raises(ValueError):  # a host tensor
            ctx.render_composite(wl.params(), wl.W, wl.H, ranks, wl.order, out.cpu())
        with pytest.raises(ValueError):  # fewer voxels than the brick's voxel range
            ctx.upload_brick(wl.bricks[0], wl.voxels[0][:-1])
        with pytest.raises(ValueError):  # TF entries that are not RGBA
            ctx.set_tf(wl.tf[:, :3])
        # still usable after the errors
        ctx.render_frame(wl.params(), wl.W, wl.H, ranks, imgs, wl.order, out)
        torch.cuda.synchronize()
        assert torch.isfinite(out).all()


@pytest.mark.parametrize("lanes", [2, 4, 8])
def test_lane_schedules_identical(c0, lanes):
    """Every lanes-per-ray schedule gives the 1-lane images bit for bit, with
    the epilogue exercised (spp, accumulation, islands, crosshair) and with a
    low-opacity TF whose rays run hundreds of samples (many group steps)."""
    _case(f"blobs128 P=2, {lanes} lanes per ray vs 1")
    acc = np.random.default_rng(4).random((2, c0.H, c0.W, 4), dtype=np.float32)
    kw = dict(spp=2, accum_id=1, island_id=1, num_islands=2, flags=A.FLAG_CROSSHAIR)
    one = gpu_render(c0, accum=acc, lanes=1, **kw)
    many = gpu_render(c0, accum=acc, lanes=lanes, **kw)
    for k in one:
        assert np.array_equal(one[k], many[k]), k
    thin = c0.tf.copy()
    thin[:, 3] *= 0.03
    one = gpu_render(c0, lanes=1, tf=thin)
    many = gpu_render(c0, lanes=lanes, tf=thin)
    for k in one:
        assert np.array_equal(one[k], many[k]), k
    assert one["nsamples"].max() > 100


def test_lane_schedules_brick_loop_and_fused(c1):
    """fbm512, one rank holding all 8 bricks (the brick loop) and the fused
    launch, for every schedule, against the oracle-checked 1-lane result."""
    import torch
    _case("fbm512 P=1 all bricks + fused, lanes 1/2/4/8")
    ranks = [list(range(len(c1.bricks)))]
    ref = gpu_render(c1, ranks, lanes=1)
    for lanes in (2, 4, 8):
        g = gpu_render(c1, ranks, lanes=lanes)
        for k in ref:
            assert np.array_equal(ref[k], g[k]), (lanes, k)
    with A.Context(0, stream=torch.cuda.current_stream()) as ctx:
        ids = [ctx.upload_brick(d, v) for d, v in zip(c1.bricks, c1.voxels)]
        ctx.set_tf(c1.tf)
        ctx.set_camera(c1.cam["org"], c1.cam["dir_00"], c1.cam["dir_du"], c1.cam["dir_dv"])
        rk = [[ids[i] for i in r] for r in c1.rank_bricks]
        outs = []
        for lanes in (1, 2, 4, 8):
            ctx.set_march_schedule(1, lanes)
            o = torch.zeros((c1.H, c1.W, 4), device="cuda")
            o8 = torch.zeros((c1.H, c1.W), dtype=torch.int32, device="cuda")
            ns = torch.zeros((c1.H, c1.W), dtype=torch.int32, device="cuda")
            ctx.render_composite(c1.params(), c1.W, c1.H, rk, c1.order, o, o8, ns)
            outs.append((o, o8, ns))
        torch.cuda.synchronize()
        for o, o8, ns in outs[1:]:
            assert torch.equal(o, outs[0][0]) and torch.equal(o8, outs[0][1])
            assert torch.equal(ns, outs[0][2])
        imgs = poison(A.RankImages.alloc(len(rk), c1.W, c1.H))
        (r_l, f_l), table = ctx.autotune_march(c1.params(), c1.W, c1.H, rk, imgs, c1.order,
                                               outs[0][0], outs[0][1], reps=2)
        print("autotune", r_l, f_l, table)
        assert r_l in (1, 2, 4, 8) and f_l in (1, 2, 4, 8) and len(table) == 4


def test_empty_rank_slots(c0):
    """Rank slots without bricks (more ranks than bricks) render transparent
    -- depth 1, no samples -- in the per-rank launch; the other slots equal
    the launch without the empty ones bit for bit; render_frame and the
    fused launch still equal the dense composite of the rank images."""
    import torch
    _case("blobs128: 4 rank slots, two of them empty, order 3 0 2 1")
    W, H = c0.W, c0.H
    params = c0.params()
    with A.Context(0, stream=torch.cuda.current_stream()) as ctx:
        ctx.set_march_schedule(4, 4)
        ids = [ctx.upload_brick(d, v) for d, v in zip(c0.bricks, c0.voxels)]
        ctx.set_tf(c0.tf)
        ctx.set_camera(c0.cam["org"], c0.cam["dir_00"], c0.cam["dir_du"], c0.cam["dir_dv"])
        ranks = [[ids[0]], [], [ids[1]], []]
        order = [3, 0, 2, 1]
        two = A.RankImages.alloc(2, W, H, nsamples=True)
        ctx.render_ranks(params, W, H, [[ids[0]], [ids[1]]], two)
        ref = poison(A.RankImages.alloc(4, W, H, nsamples=True))
        ctx.render_ranks(params, W, H, ranks, ref)
        dense = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
        dense8 = torch.empty((H, W), dtype=torch.int32, device="cuda")
        ctx.composite_over(ref.rgba, order, W, H, dense, dense8)
        got = poison(A.RankImages.alloc(4, W, H, nsamples=True))
        out = torch.full_like(dense, float("nan"))
        out8 = torch.full_like(dense8, -1)
        ctx.render_frame(params, W, H, ranks, got, order, out, out8)
        fout = torch.full_like(dense, float("nan"))
        fout8 = torch.full_like(dense8, -1)
        ctx.render_composite(params, W, H, ranks, order, fout, fout8)
        torch.cuda.synchronize()
        for s in (1, 3):
            assert not ref.rgba[s].any() and not ref.rgba8[s].any() and not ref.nsamples[s].any()
            assert bool((ref.depth[s] == 1.0).all())
        for name in ("rgba", "rgba8", "depth", "nsamples"):
            r, t = getattr(ref, name), getattr(two, name)
            assert torch.equal(r[0], t[0]) and torch.equal(r[2], t[1]), name
            assert torch.equal(getattr(got, name), r), name
        assert int(ref.nsamples.sum()) > 0
        assert torch.equal(out, dense) and torch.equal(out8, dense8)
        assert torch.equal(fout, dense) and torch.equal(fout8, dense8)
