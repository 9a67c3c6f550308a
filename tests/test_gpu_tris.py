"""GPU checks of the surface path (aloha_render_tris / _composite): per-rank
images against the oracle's brute-force closest hit (small meshes) and its
CPU walk of the exported BVH (large meshes), the fused depth composite
against per-rank images + aloha_composite_zmin, flags and edge cases."""
import numpy as np
import pytest

import aloha_amd as A
import oracle as O
from aloha_amd import workloads as WL

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    return torch


def _images(P, W, H):
    torch = _torch()
    return A.RankImages(rgba=torch.zeros((P, H, W, 4), dtype=torch.float32, device="cuda"),
                        rgba8=torch.zeros((P, H, W), dtype=torch.int32, device="cuda"),
                        depth=torch.zeros((P, H, W), dtype=torch.float32, device="cuda"))


def _u32(t):
    return t.cpu().numpy().view(np.uint32)


@pytest.fixture(scope="module")
def small():
    return WL.surface(num_spheres=40, rings=8, segments=12, num_clusters=4, num_ranks=4,
                      W=160, H=120, seed=21, dist_factor=1.0, name="small", rmin=0.06, rmax=0.15)


def _ctx(wl):
    ctx = A.Context()
    ctx.set_camera(wl.cam["org"], wl.cam["dir_00"], wl.cam["dir_du"], wl.cam["dir_dv"])
    geoms = ctx.upload_tris(wl.mesh, wl.ranks, wl.colors)
    return ctx, geoms


def _oracle_rank(wl, r, **kw):
    tri9 = wl.mesh.rank_tri9(wl.ranks[r])
    sizes = [l - f for f, l, _ in wl.mesh.clusters()]
    col = np.repeat(np.asarray(wl.colors[r], np.uint32), [sizes[c] for c in wl.ranks[r]])
    return O.tri_render(wl.cam, tri9, col, wl.W, wl.H, max_depth=wl.max_depth, **kw)


@pytest.mark.parametrize("variant", ["plain", "flags"])
def test_rank_images_match_brute_force(small, variant):
    print(f"\n[case] tris per-rank {variant} {small.W}x{small.H} P={len(small.ranks)}", flush=True)
    torch = _torch()
    kw = {} if variant == "plain" else {"spp": 2, "island_id": 1, "num_islands": 2, "crosshair": True}
    ctx, geoms = _ctx(small)
    P, W, H = len(geoms), small.W, small.H
    out = _images(P, W, H)
    stats = torch.zeros((P, H, W, 2), dtype=torch.int32, device="cuda")
    params = small.params(spp=kw.get("spp", 1), island_id=kw.get("island_id", 0),
                          num_islands=kw.get("num_islands", 1),
                          flags=A.FLAG_CROSSHAIR if kw.get("crosshair") else 0)
    ctx.render_tris(params, W, H, geoms, out, stats=stats)
    torch.cuda.synchronize()
    hits = 0
    for r in range(P):
        ref = _oracle_rank(small, r, **kw)
        hits += int((ref["prim"] >= 0).sum())
        assert np.array_equal(out.depth[r].cpu().numpy(), ref["depth"])
        assert np.array_equal(out.rgba[r].cpu().numpy(), ref["rgba"])
        assert np.array_equal(_u32(out.rgba8[r]), ref["rgba8"])
        st = stats[r].cpu().numpy()
        assert (st[..., 1][ref["prim"] >= 0] > 0).all()
    assert hits > 2000


def test_heat_map_host_clock(small):
    print(f"\n[case] tris heat map, host clock delta, {small.W}x{small.H}", flush=True)
    torch = _torch()
    ctx, geoms = _ctx(small)
    P, W, H = len(geoms), small.W, small.H
    out = _images(P, W, H)
    params = small.params(spp=2, flags=A.FLAG_HEATMAP | A.FLAG_HEAT_CLOCK, heat_clock=777777,
                          heat_map_scale=1e-6)
    ctx.render_tris(params, W, H, geoms, out)
    torch.cuda.synchronize()
    for r in range(P):
        ref = _oracle_rank(small, r, spp=2, heat_clock=777777, heat_scale=1e-6)
        assert np.array_equal(out.depth[r].cpu().numpy(), ref["depth"])
        assert np.array_equal(out.rgba[r].cpu().numpy(), ref["rgba"])
        assert np.array_equal(_u32(out.rgba8[r]), ref["rgba8"])
    ctx.close()


def test_fused_equals_ranks_plus_zmin(small):
    print("\n[case] tris fused composite vs ranks + zmin", flush=True)
    torch = _torch()
    ctx, geoms = _ctx(small)
    P, W, H = len(geoms), small.W, small.H
    for flags in (0, A.FLAG_CROSSHAIR):
        params = small.params(flags=flags)
        out = _images(P, W, H)
        ctx.render_tris(params, W, H, geoms, out)
        z8 = torch.zeros((H, W), dtype=torch.int32, device="cuda")
        zz = torch.zeros((H, W), dtype=torch.float32, device="cuda")
        ctx.composite_zmin(out.rgba8, out.depth, P, W, H, z8, zz)
        f8 = torch.zeros_like(z8)
        fz = torch.zeros_like(zz)
        ctx.render_tris_composite(params, W, H, geoms, f8, fz)
        torch.cuda.synchronize()
        assert torch.equal(f8, z8) and torch.equal(fz, zz)
        # and the oracle's depth composite of its own rank images
        refs = [_oracle_rank(small, r, crosshair=bool(flags)) for r in range(P)]
        o8, oz = O.composite_zmin(np.stack([x["rgba8"] for x in refs]),
                                  np.stack([x["depth"] for x in refs]))
        assert np.array_equal(_u32(f8), o8) and np.array_equal(fz.cpu().numpy(), oz)
        assert (oz < 1).sum() > 2000


def test_empty_rank_and_unsplit_mesh():
    print("\n[case] tris empty rank + unsplit mesh", flush=True)
    torch = _torch()
    dom = A.Box3f((0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    mesh = A.Mesh.spheres(6, 6, 8, dom, 0.05, 0.15, 0.1, seed=3)
    W, H = 64, 48
    cam = A.synth_camera(mesh.info()["bounds"], W, H, 45.0, 1.0, seed=4)
    md = A.max_depth(mesh.info()["bounds"], cam["org"])
    ctx = A.Context()
    ctx.set_camera(cam["org"], cam["dir_00"], cam["dir_du"], cam["dir_dv"])
    geoms = ctx.upload_tris(mesh, [[0], []], [[7], []])
    out = _images(2, W, H)
    ctx.render_tris(A.render_params(max_depth=md), W, H, geoms, out)
    torch.cuda.synchronize()
    v, t = mesh.arrays()
    ref = O.tri_render(cam, v[t].reshape(-1, 9), np.full(len(t), 7, np.uint32), W, H, max_depth=md)
    assert np.array_equal(_u32(out.rgba8[0]), ref["rgba8"])
    assert np.array_equal(out.depth[0].cpu().numpy(), ref["depth"])
    assert int(out.rgba8[1].abs().sum()) == 0 and bool((out.depth[1] == 1).all())
    assert (ref["prim"] >= 0).sum() > 100


def test_errors():
    print("\n[case] tris argument errors", flush=True)
    ctx = A.Context()
    mesh = A.Mesh.spheres(2, 4, 6, A.Box3f((0.0, 0.0, 0.0), (1.0, 1.0, 1.0)), 0.1, 0.2, 0.1, seed=1)
    with pytest.raises(A.AlohaError):
        ctx.upload_tris(mesh, [[1]], [[0]])           # unsplit mesh has cluster 0 only
    geoms = ctx.upload_tris(mesh, [[0]], [[0]])
    torch = _torch()
    out = _images(1, 8, 8)
    with pytest.raises(A.AlohaError):                 # no camera yet
        ctx.render_tris(A.render_params(), 8, 8, geoms, out)
    ctx.set_camera((0, 0, 3), (-0.5, -0.5, -1), (1 / 8, 0, 0), (0, 1 / 8, 0))
    with pytest.raises(A.AlohaError):
        ctx.render_tris(A.render_params(), 8, 8, [geoms[0] + 5], out)
    with pytest.raises(A.AlohaError):
        ctx.render_tris_composite(A.render_params(accum_id=1), 8, 8, geoms,
                                  torch.zeros((8, 8), dtype=torch.int32, device="cuda"))
    ctx.render_tris(A.render_params(), 8, 8, geoms, out)
    torch.cuda.synchronize()


def test_medium_mesh_against_bvh_walk():
    """0.3 M triangles, 8 ranks: every pixel of every rank image against the
    oracle's CPU walk of the same exported BVH and triangle arrays."""
    print("\n[case] tris 128 spheres x 2380, 8 ranks, 480x270 vs CPU BVH walk", flush=True)
    torch = _torch()
    wl = WL.surface(num_spheres=128, W=480, H=270, seed=33, name="medium", rmin=0.03, rmax=0.1)
    ctx, geoms = _ctx(wl)
    P, W, H = len(geoms), wl.W, wl.H
    out = _images(P, W, H)
    ctx.render_tris(wl.params(), W, H, geoms, out)
    f8 = torch.zeros((H, W), dtype=torch.int32, device="cuda")
    fz = torch.zeros((H, W), dtype=torch.float32, device="cuda")
    ctx.render_tris_composite(wl.params(), W, H, geoms, f8, fz)
    torch.cuda.synchronize()
    refs = []
    for r in range(P):
        nodes, tris, root = ctx.tris_export(geoms[r])
        ref = O.tri_render_bvh(wl.cam, nodes, tris, root, W, H, max_depth=wl.max_depth)
        refs.append(ref)
        assert np.array_equal(out.depth[r].cpu().numpy(), ref["depth"]), r
        assert np.array_equal(_u32(out.rgba8[r]), ref["rgba8"]), r
    o8, oz = O.composite_zmin(np.stack([x["rgba8"] for x in refs]), np.stack([x["depth"] for x in refs]))
    assert np.array_equal(_u32(f8), o8) and np.array_equal(fz.cpu().numpy(), oz)
    assert (oz < 1).sum() > 15000


def test_full_size_sampled_rows():
    """The 9.75 M-triangle scene at 1920x1080: fused frame = per-rank + zmin
    on every pixel; 12 sampled rows of each rank image against the CPU walk."""
    print("\n[case] tris 9.75M triangles 1920x1080, sampled rows", flush=True)
    torch = _torch()
    wl = WL.surface()
    ctx, geoms = _ctx(wl)
    P, W, H = len(geoms), wl.W, wl.H
    out = A.RankImages(rgba8=torch.zeros((P, H, W), dtype=torch.int32, device="cuda"),
                       depth=torch.zeros((P, H, W), dtype=torch.float32, device="cuda"))
    ctx.render_tris(wl.params(), W, H, geoms, out)
    z8 = torch.zeros((H, W), dtype=torch.int32, device="cuda")
    zz = torch.zeros((H, W), dtype=torch.float32, device="cuda")
    ctx.composite_zmin(out.rgba8, out.depth, P, W, H, z8, zz)
    f8 = torch.zeros_like(z8)
    fz = torch.zeros_like(zz)
    ctx.render_tris_composite(wl.params(), W, H, geoms, f8, fz)
    torch.cuda.synchronize()
    assert torch.equal(f8, z8) and torch.equal(fz, zz)
    assert int((fz < 1).sum()) > W * H // 6
    rows = np.linspace(0, H - 1, 12).astype(int)
    for r in range(P):
        nodes, tris, root = ctx.tris_export(geoms[r])
        d = out.depth[r].cpu().numpy()
        c = _u32(out.rgba8[r])
        for y in rows:
            ref = O.tri_render_bvh(wl.cam, nodes, tris, root, W, H, max_depth=wl.max_depth,
                                   rows=(int(y), int(y) + 1))
            assert np.array_equal(d[y], ref["depth"][y]), (r, y)
            assert np.array_equal(c[y], ref["rgba8"][y]), (r, y)


def test_headless_tri_renderer_api(tmp_path):
    """aloha_headless on a .tri file: maui::Renderer in TriMesh mode
    (round-robin clusters, colour ids, per-rank images, depth composite)
    against the oracle, 2 frames (accumulated), display rank and rank 1."""
    import subprocess
    from pathlib import Path
    print("\n[case] aloha_headless .tri, 4 ranks", flush=True)
    wl = WL.surface(num_spheres=30, rings=6, segments=10, num_clusters=6, num_ranks=4, W=96, H=72,
                    seed=41, rmin=0.06, rmax=0.15, name="headless")
    path = tmp_path / "s.tri"
    wl.mesh.write_tri(str(path))
    cam = [str(float(x)) for k in ("org", "dir_00", "dir_du", "dir_dv") for x in wl.cam[k]]
    exe = Path(A.__file__).resolve().parent / "bin" / "aloha_headless"
    refs = [_oracle_rank(wl, r, crosshair=True) for r in range(4)]  # the Renderer default
    o8, oz = O.composite_zmin(np.stack([x["rgba8"] for x in refs]), np.stack([x["depth"] for x in refs]))
    for rank in (0, 1):
        out = tmp_path / f"f{rank}"
        r = subprocess.run([str(exe), str(path), "-ranks", "4", "-rank", str(rank), "-size",
                            str(wl.W), str(wl.H), "-camera", *cam, "-frames", "2", "-o", str(out)],
                           capture_output=True, text=True, timeout=120)
        print(r.stdout, r.stderr)
        assert r.returncode == 0
        got8 = np.fromfile(str(out) + ".rgba8", np.uint32).reshape(wl.H, wl.W)
        gotz = np.fromfile(str(out) + ".depth", np.float32).reshape(wl.H, wl.W)
        assert np.array_equal(gotz, oz)
        # frames 0 and 1 accumulate the same colour: (c + c) * (1/2) == c for the pack
        if rank == 0:
            assert np.array_equal(got8, o8)
        else:
            assert np.array_equal(got8, refs[1]["rgba8"])
    assert (oz < 1).sum() > 1000
