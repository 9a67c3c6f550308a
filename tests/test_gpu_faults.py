"""GPU checks of the failure and edge paths through the C ABI:

* a transfer function with NaN entries: refused by aloha_tf_set; through the
  differentiable form (which takes the TF as a device tensor) the forward
  follows the oracle's float arithmetic (NaN colours where a sample reads a
  NaN entry, the same sample counts) and the adjoint propagates the NaN;
* zero-size and degenerate bricks: an empty voxel range is refused; bricks of
  one voxel, one voxel thick, or with a flat space range render exactly as
  the oracle renders them;
* allocation failure: every device / pinned allocation an entry point makes
  is failed in turn (aloha_debug_alloc_fault); the call returns
  ALOHA_ERR_ALLOC, leaves the context usable and as it was, and the same
  call made again gives exactly the result of a run without the fault;
* aloha_render_frame_noalloc makes no allocation at all.
"""
from types import SimpleNamespace

import numpy as np
import pytest

import aloha_amd as A
import oracle as O
from aloha_amd import workloads as WL
from aloha_amd._lib import ERR_ALLOC, ERR_ARG, ERR_STATE, lib

from gpu_util import compare, gpu_render, oracle_render

pytestmark = pytest.mark.gpu


def _case(name):
    print(f"\n[case] {name}", flush=True)


def _torch():
    import torch
    return torch


def _desc(vlo, n, space_lo=None, space_hi=None):
    """Brick of n voxels from vlo: cells = voxels, space = the cell box
    unless given."""
    d = A.BrickDesc()
    for a in range(3):
        d.cell_range.lower[a] = d.voxel_range.lower[a] = vlo[a]
        d.cell_range.upper[a] = d.voxel_range.upper[a] = vlo[a] + n[a]
        d.space_range.lower[a] = float(vlo[a]) if space_lo is None else space_lo[a]
        d.space_range.upper[a] = float(vlo[a] + n[a]) if space_hi is None else space_hi[a]
    d.value_range.lower, d.value_range.upper = 0.0, 1.0
    return d


def _brick_count(ctx):
    import ctypes as C
    n = C.c_int32()
    ctx._check(lib().aloha_brick_count(ctx._h, C.byref(n)))
    return n.value


# ------------------------------------------------------------------ NaN TF
def test_nan_tf_refused_by_tf_set():
    _case("aloha_tf_set with a NaN entry")
    tf = A.synth_tf(16, 4, 1)
    tf[5, 2] = np.nan
    with A.Context(0) as ctx:
        with pytest.raises(A.AlohaError) as e:
            ctx.set_tf(tf)
        assert e.value.code == ERR_ARG


def test_nan_tf_through_the_differentiable_form():
    torch = _torch()
    from aloha_amd.autograd import DifferentiableVolume
    _case("differentiable form with NaN TF entries: forward vs oracle, adjoint propagates")
    wl = WL.diff256(size=40, views=2, W=64, H=56)
    tf = wl.tf.copy()
    tf[60:70, 0] = np.nan   # a colour channel
    tf[90:95, 3] = np.nan   # opacity (the clamp takes the other operand of fminf)
    vol = DifferentiableVolume(wl.density.shape, wl.cams, wl.W, wl.H, dt=wl.dt, jitter_value=0.5)
    tft = torch.from_numpy(tf).cuda().requires_grad_(True)
    den = torch.from_numpy(wl.density).cuda().requires_grad_(True)
    rgba = vol.render(tft, den)
    _, ns = vol.forward_only(tft, den)
    torch.cuda.synchronize()
    g, gns = rgba.detach().cpu().numpy(), ns.cpu().numpy()
    n = wl.density.shape[::-1]
    blist = [({"cell": ((0, 0, 0), n), "voxel": ((0, 0, 0), n),
               "space": ((0.0, 0.0, 0.0), tuple(float(v) for v in n))}, wl.density)]
    nan_pixels = 0
    for v, cam in enumerate(wl.cams):
        r, rns = O.diff_render(cam, tf, blist, wl.W, wl.H, dt=wl.dt, jitter_mode=1,
                               jitter_value=0.5)
        assert np.array_equal(gns[v], rns.astype(np.int32)), v
        assert np.array_equal(np.isnan(g[v]), np.isnan(r)), v
        fin = ~np.isnan(r)
        assert np.abs(g[v][fin] - r[fin]).max() <= 1e-4
        nan_pixels += int(np.isnan(r).any(-1).sum())
    print(f"{nan_pixels} pixels with a NaN channel")
    assert nan_pixels > 0
    # the adjoint runs to the end and carries the NaN into the gradients (a
    # bin with a NaN contribution sums to NaN), not into integers
    rgba.sum().backward()
    torch.cuda.synchronize()
    assert torch.isnan(tft.grad).any() and torch.isnan(den.grad).any()
    vol.ctx.close()


# ------------------------------------------------------------------ degenerate bricks
def test_empty_voxel_range_refused():
    _case("brick with an empty voxel range")
    with A.Context(0) as ctx:
        with pytest.raises(A.AlohaError) as e:
            ctx.upload_brick(_desc((0, 0, 0), (4, 0, 4)), np.zeros((4, 1, 4), np.float32))
        assert e.value.code == ERR_ARG
        assert _brick_count(ctx) == 0


@pytest.mark.parametrize("shape", ["one_voxel", "one_thick", "flat_space"])
def test_degenerate_bricks_match_oracle(shape):
    _case(f"degenerate brick: {shape}")
    rng = np.random.default_rng(7)
    main = _desc((0, 0, 0), (12, 10, 8))
    bricks = [main]
    vox = [rng.random((8, 10, 12), dtype=np.float32)]
    if shape == "one_voxel":
        bricks.append(_desc((13, 4, 3), (1, 1, 1)))
        vox.append(np.full((1, 1, 1), 0.8, np.float32))
    elif shape == "one_thick":
        bricks.append(_desc((0, 0, 9), (12, 10, 1)))
        vox.append(rng.random((1, 10, 12), dtype=np.float32))
    else:  # voxels, but a space range of zero thickness in y: no ray can enter it
        bricks.append(_desc((0, 11, 0), (12, 3, 8), space_lo=(0.0, 12.0, 0.0),
                            space_hi=(12.0, 12.0, 8.0)))
        vox.append(rng.random((8, 3, 12), dtype=np.float32))
    bounds = A.Box3f((0, 0, 0), (14, 14, 10))
    W, H = 96, 80
    cam = A.synth_camera(bounds, W, H, 40.0, 2.0, 11)
    tf = A.synth_tf(256, 6, 5)
    tf[:, 3] *= 0.3
    wl = SimpleNamespace(bricks=bricks, voxels=vox, tf=tf, cam=cam, W=W, H=H, dt=0.25,
                         max_depth=A.max_depth(bounds, cam["org"]), rank_bricks=[[0], [1]])
    wl.params = lambda **kw: A.render_params(**{"dt": wl.dt, "max_depth": wl.max_depth, **kw})
    for ranks in ([[0], [1]], [[0, 1]]):
        info = compare(gpu_render(wl, ranks), oracle_render(wl, ranks))
        print(ranks, info)
        assert info["exact_fraction"] == 1.0


# ------------------------------------------------------------------ allocation faults
def _fault_each(make_state, op, result, expect_made=None):
    """make_state() -> state on a fresh context; op(state) makes the call
    under test; result(state) -> arrays to compare.  Every allocation op makes
    is failed in turn."""
    state = make_state()
    m0 = A.debug_alloc_fault(-1)
    op(state)
    n = A.debug_alloc_fault(-1) - m0
    want = result(state)
    state.close()
    if expect_made is not None:
        assert n == expect_made, n
    assert n > 0
    for k in range(n):
        state = make_state()
        A.debug_alloc_fault(k)
        with pytest.raises(A.AlohaError) as e:
            op(state)
        A.debug_alloc_fault(-1)
        assert e.value.code == ERR_ALLOC, (k, e.value)
        state.after_failure()
        op(state)  # again, without the fault
        got = result(state)
        state.close()
        for a, b in zip(want, got):
            assert np.array_equal(a, b, equal_nan=True), k
    return n


class _FrameState:
    """C0 workload on a fresh context; helpers for the fault tests."""
    wl = None

    def __init__(self, upload=True, tf=True):
        torch = _torch()
        if _FrameState.wl is None:
            _FrameState.wl = WL.blobs128()
        self.ctx = A.Context(0, stream=torch.cuda.current_stream())
        self.ids = []
        wl = self.wl
        self.ctx.set_camera(wl.cam["org"], wl.cam["dir_00"], wl.cam["dir_du"], wl.cam["dir_dv"])
        if tf:
            self.ctx.set_tf(wl.tf)
        if upload:
            self.ids = [self.ctx.upload_brick(d, v) for d, v in zip(wl.bricks, wl.voxels)]

    def render(self):
        torch = _torch()
        wl = self.wl
        imgs = A.RankImages.alloc(wl.num_ranks, wl.W, wl.H, rgba=True, rgba8=False, depth=False)
        out = torch.empty((wl.H, wl.W, 4), dtype=torch.float32, device="cuda")
        self.ctx.render_frame(wl.params(jitter_mode=A.JITTER_FIXED, jitter_value=0.5), wl.W, wl.H,
                              [[self.ids[i] for i in r] for r in wl.rank_bricks], imgs, wl.order,
                              out, None)
        torch.cuda.synchronize()
        return [out.cpu().numpy()]

    def after_failure(self):
        pass

    def close(self):
        self.ctx.close()


def test_alloc_fault_brick_upload():
    _case("allocation faults: brick upload (padded grid, staging, brick table)")

    class S(_FrameState):
        def __init__(self):
            super().__init__(upload=False)

        def after_failure(self):
            # the failed upload added nothing: only the bricks uploaded before it
            assert _brick_count(self.ctx) == len(self.ids)

    def op(s):
        wl = s.wl
        s.ids = []
        for d, v in zip(wl.bricks, wl.voxels):
            s.ids.append(s.ctx.upload_brick(d, v))

    n = _fault_each(S, op, lambda s: s.render())
    print(f"{n} allocations failed in turn")


def test_alloc_fault_tf_growth():
    _case("allocation faults: transfer-function table growth")

    class S(_FrameState):
        def __init__(self):
            super().__init__(tf=False)
            self.ctx.set_tf(self.wl.tf[:4])  # a small table first: the next one grows it

        def after_failure(self):
            with pytest.raises(A.AlohaError) as e:  # no table now: refused, not read
                self.render()
            assert e.value.code == ERR_STATE

    _fault_each(S, lambda s: s.ctx.set_tf(s.wl.tf), lambda s: s.render(), expect_made=1)


def test_alloc_fault_footprint():
    _case("allocation faults: footprint bitset")

    class S(_FrameState):
        pass

    wl_params = lambda s: s.wl.params(jitter_mode=A.JITTER_FIXED, jitter_value=0.5)

    def op(s):
        s.touched = s.ctx.footprint(wl_params(s), s.wl.W, s.wl.H,
                                    [[s.ids[i] for i in r] for r in s.wl.rank_bricks])

    _fault_each(S, op, lambda s: [np.array([s.touched])] + s.render())


def test_alloc_fault_adjoint_workspace():
    torch = _torch()
    _case("allocation faults: adjoint workspace (bins, ray offsets, pinned total, resampled "
          "values, records, tiles, TF partials)")
    from aloha_amd.autograd import DifferentiableVolume
    wl = WL.diff256(size=40, views=2, W=64, H=56)

    class S:
        def __init__(self):
            self.vol = DifferentiableVolume(wl.density.shape, wl.cams, wl.W, wl.H, dt=wl.dt,
                                            jitter_value=0.5)
            self.tf = torch.from_numpy(wl.tf).cuda().requires_grad_(True)
            self.den = torch.from_numpy(wl.density).cuda().requires_grad_(True)
            self.w = torch.from_numpy(np.random.default_rng(2).standard_normal(
                (len(wl.cams), wl.H, wl.W, 4)).astype(np.float32)).cuda()

        def after_failure(self):
            self.tf.grad = None
            self.den.grad = None

        def close(self):
            self.vol.ctx.close()

    def op(s):
        (s.vol.render(s.tf, s.den) * s.w).sum().backward()

    def result(s):
        torch.cuda.synchronize()
        return [s.tf.grad.cpu().numpy(), s.den.grad.cpu().numpy()]

    n = _fault_each(S, op, result)
    print(f"{n} allocations failed in turn")


def test_alloc_fault_batch_create():
    torch = _torch()
    _case("allocation faults: batch creation")
    bw = WL.batched(n_side=2, cell=48, smin=16, views=3, W=96, H=80)

    class S:
        def __init__(self):
            self.ctx = A.Context(0, stream=torch.cuda.current_stream())
            self.ids = [self.ctx.upload_brick(d, v) for d, v in zip(bw.bricks, bw.voxels)]
            self.ctx.set_tf(bw.tf)
            V = len(bw.cams)
            self.out = torch.empty((V, bw.H, bw.W, 4), dtype=torch.float32, device="cuda")
            self.b = None

        def after_failure(self):
            assert self.b is None

        def close(self):
            if self.b is not None:
                self.b.close()
            self.ctx.close()

    def op(s):
        s.b = A.Batch(s.ctx, bw.params(), bw.W, bw.H, bw.cams,
                      [[s.ids[i] for i in r] for r in bw.rank_bricks], bw.orders, out_rgba=s.out)

    def result(s):
        s.b.launch()
        torch.cuda.synchronize()
        return [s.out.cpu().numpy()]

    _fault_each(S, op, result)


def test_alloc_fault_tris_upload():
    torch = _torch()
    _case("allocation faults: triangle geometry upload")
    sw = WL.surface(num_spheres=20, rings=6, segments=8, num_clusters=3, num_ranks=3, W=64, H=48,
                    seed=3, name="tiny", rmin=0.08, rmax=0.2)

    class S:
        def __init__(self):
            self.ctx = A.Context(0, stream=torch.cuda.current_stream())
            self.ctx.set_camera(sw.cam["org"], sw.cam["dir_00"], sw.cam["dir_du"], sw.cam["dir_dv"])
            self.geoms = None

        def after_failure(self):
            assert self.geoms is None

        def close(self):
            self.ctx.close()

    def op(s):
        s.geoms = s.ctx.upload_tris(sw.mesh, sw.ranks, sw.colors)

    def result(s):
        imgs = A.RankImages.alloc(len(s.geoms), sw.W, sw.H, rgba=True, rgba8=True, depth=True)
        s.ctx.render_tris(sw.params(), sw.W, sw.H, s.geoms, imgs)
        torch.cuda.synchronize()
        # ids restart where the failed call began: the same ids, the same images
        return [np.array(s.geoms), imgs.rgba.cpu().numpy(), imgs.depth.cpu().numpy()]

    _fault_each(S, op, result, expect_made=2 * len(sw.ranks))


def test_noalloc_frame_makes_no_allocation():
    torch = _torch()
    _case("render_frame_noalloc with every allocation armed to fail")
    s = _FrameState()
    wl = s.wl
    out = torch.empty((wl.H, wl.W, 4), dtype=torch.float32, device="cuda")
    p = wl.params()
    ranks = [[s.ids[i] for i in r] for r in wl.rank_bricks]
    s.ctx.render_frame_noalloc(p, wl.W, wl.H, ranks, wl.order, out, None)  # first call: set-up
    torch.cuda.synchronize()
    m0 = A.debug_alloc_fault(0)  # the next allocation fails
    s.ctx.render_frame_noalloc(p, wl.W, wl.H, ranks, wl.order, out, None)
    torch.cuda.synchronize()
    assert A.debug_alloc_fault(-1) == m0
    s.close()
