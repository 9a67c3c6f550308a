"""GPU parity of the sort-last composite kernels (IceT ordered BLEND and
Z_BUFFER, waikikiRenderer.cpp:903-916, :1043-1077) against the oracle."""
import numpy as np
import pytest

import aloha_amd as A
import oracle as O
from aloha_amd import workloads as WL

from gpu_util import gpu_render

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def test_over_random_images_bit_exact():
    import torch
    print("\n[case] over random 8 ranks 333x257", flush=True)
    rng = np.random.default_rng(7)
    P, H, W = 8, 257, 333
    a = rng.random((P, H, W, 1), dtype=np.float32)
    imgs = (rng.random((P, H, W, 4), dtype=np.float32) * a).astype(np.float32)
    imgs[..., 3:] = a
    order = list(rng.permutation(P))
    ref, ref8 = O.composite_over(imgs, order)
    out = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
    out8 = torch.empty((H, W), dtype=torch.int32, device="cuda")
    with A.Context(0, stream=torch.cuda.current_stream()) as ctx:
        ctx.composite_over(_dev(imgs), order, W, H, out, out8)
        torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), ref)
    assert np.array_equal(out8.cpu().numpy().view(np.uint32), ref8)


def test_over_rgba8_and_zmin():
    import torch
    print("\n[case] over_rgba8 + zmin 5 ranks", flush=True)
    rng = np.random.default_rng(8)
    P, H, W = 5, 64, 100
    imgs8 = rng.integers(0, 2**32, (P, H, W), dtype=np.uint64).astype(np.uint32)
    order = [3, 1, 4, 0, 2]
    ref8 = O.composite_over_u8(imgs8, order)
    depth = rng.random((P, H, W), dtype=np.float32)
    depth[1, :5] = depth[0, :5]  # ties go to the lower rank
    refz8, refz = O.composite_zmin(imgs8, depth)
    out8 = torch.empty((H, W), dtype=torch.int32, device="cuda")
    z8 = torch.empty((H, W), dtype=torch.int32, device="cuda")
    z = torch.empty((H, W), dtype=torch.float32, device="cuda")
    with A.Context(0, stream=torch.cuda.current_stream()) as ctx:
        ctx.composite_over_rgba8(_dev(imgs8.view(np.int32)), order, W, H, out8)
        ctx.composite_zmin(_dev(imgs8.view(np.int32)), _dev(depth), P, W, H, z8, z)
        torch.cuda.synchronize()
    assert np.array_equal(out8.cpu().numpy().view(np.uint32), ref8)
    assert np.array_equal(z8.cpu().numpy().view(np.uint32), refz8)
    assert np.array_equal(z.cpu().numpy(), refz)


def test_c0_render_then_composite():
    import torch
    print("\n[case] blobs128 render P=2 + composite", flush=True)
    wl = WL.blobs128()
    g = gpu_render(wl)
    ref, ref8 = O.composite_over(g["rgba"], wl.order)
    out = torch.empty((wl.H, wl.W, 4), dtype=torch.float32, device="cuda")
    with A.Context(0, stream=torch.cuda.current_stream()) as ctx:
        ctx.composite_over(_dev(g["rgba"]), wl.order, wl.W, wl.H, out, None)
        torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), ref)
    assert ref[..., 3].max() > 0.5
