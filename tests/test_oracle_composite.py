"""Pins the oracle's 8-bit ordered blend (orc_composite_over_u8, the
reference's RGBA_UBYTE IceT path, waikikiRenderer.cpp:911-916) and its depth
composite (orc_composite_zmin, IceT Z_BUFFER, waikikiRenderer.cpp:903-909)
independently of the oracle's own code.

The 8-bit cases are worked by hand in exact arithmetic: a byte b decodes to
b/255, layers blend as premultiplied 'over' front to back onto transparent
black, and the result packs as min(255, floor(256 x)).  Each case is chosen
so that one plausible mistake changes a byte: blending back to front, a /256
decode, rounding instead of truncating in the pack, a missing clamp at 1.0,
or a non-premultiplied blend.  IceT's own ubyte arithmetic is not in the
reference (empty submodule): these pin this build's stated convention, not
IceT's bytes.

The depth composite is pinned against numpy's argmin, which returns the first
(lowest) index among equal minima: the tie rule this build fixes."""
from fractions import Fraction as F

import numpy as np
import pytest

import oracle as O


def rgba8(r, g, b, a):
    return r | g << 8 | b << 16 | a << 24


def _blend(layers):
    """layers: list of RGBA8 words, front first -> output word (one pixel)."""
    imgs = np.array([[w] for w in layers], np.uint32)  # P x 1 pixel
    return int(O.composite_over_u8(imgs, list(range(len(layers))))[0])


def _exact(layers):
    C = [F(0)] * 4
    for w in layers:
        c = [F((w >> (8 * ch)) & 255, 255) for ch in range(4)]
        rem = 1 - C[3]
        C = [C[ch] + rem * c[ch] for ch in range(4)]
    out = 0
    for ch in range(4):
        x = C[ch] * 256
        # keep every hand case away from a pack boundary, so float rounding
        # of the blend cannot decide the byte
        assert x >= 256 or x == int(x) or abs(x - round(x)) > F(1, 1000)
        out |= min(255, int(x)) << (8 * ch)
    return out


CASES = {
    # opaque front layer hides what is behind it; a byte b < 255 packs back
    # to b (floor(256 b / 255) = b), 255 packs to 255 through the clamp
    "opaque_front": ([rgba8(200, 100, 50, 255), rgba8(9, 9, 9, 128)], rgba8(200, 100, 50, 255)),
    # transparent layers in front leave the back layer byte for byte
    "transparent_front": ([0, 0, rgba8(48, 16, 32, 127)], rgba8(48, 16, 32, 127)),
    "all_255": ([0, rgba8(255, 255, 255, 255)], rgba8(255, 255, 255, 255)),
    # half-alpha front over an opaque back: G = 127/255 * 200/255 * 256 =
    # 99.9985 -> 99 (a rounding pack gives 100, a /256 decode gives 100);
    # A = 128/255 + 127/255 = 1 -> 256 -> clamped 255
    "half_over_opaque": ([rgba8(64, 0, 0, 128), rgba8(0, 200, 0, 255)], rgba8(64, 99, 0, 255)),
    # three layers, alphas .2 / .4 / 1: R = .2*256 = 51.2 -> 51, G = .8*.4*256
    # = 81.92 -> 81 (rounding: 82), B = .48*256 = 122.88 -> 122 (rounding: 123)
    "three_layers": ([rgba8(51, 0, 0, 51), rgba8(0, 102, 0, 102), rgba8(0, 0, 255, 255)],
                     rgba8(51, 81, 122, 255)),
    # back to front would give R = 0 here: order matters
    "order": ([rgba8(0, 0, 0, 255), rgba8(255, 0, 0, 255)], rgba8(0, 0, 0, 255)),
    # premultiplied: a front (R=100, A=100) adds its R unscaled, 100/255*256 =
    # 100.39 -> 100; a straight-alpha blend would scale it by A (39)
    "premultiplied": ([rgba8(100, 0, 0, 100), 0], rgba8(100, 0, 0, 100)),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_u8_blend_hand_cases(name):
    layers, want = CASES[name]
    assert _exact(layers) == want  # the hand-worked bytes themselves
    assert _blend(layers) == want


def test_u8_blend_random_pixels_match_exact_arithmetic():
    # random premultiplied bytes (rgb <= a), 4 layers: every pixel whose exact
    # result keeps clear of a pack boundary must agree with exact arithmetic
    rng = np.random.default_rng(11)
    P, n = 4, 3000
    a = rng.integers(0, 256, (P, n))
    rgb = (rng.random((P, n, 3)) * (a[..., None] + 1)).astype(np.int64).clip(0, 255)
    words = (rgb[..., 0] | rgb[..., 1] << 8 | rgb[..., 2] << 16 | a << 24).astype(np.uint32)
    order = [2, 0, 3, 1]
    got = O.composite_over_u8(words, order)
    checked = 0
    for i in range(n):
        C = [F(0)] * 4
        for r in order:
            w = int(words[r, i])
            rem = 1 - C[3]
            C = [C[ch] + rem * F((w >> (8 * ch)) & 255, 255) for ch in range(4)]
        xs = [C[ch] * 256 for ch in range(4)]
        if any(x < 256 and x != int(x) and abs(x - round(x)) < F(1, 10000) for x in xs):
            continue
        want = sum(min(255, int(x)) << (8 * ch) for ch, x in enumerate(xs))
        assert int(got[i]) == want, i
        checked += 1
    assert checked > 0.95 * n


def test_zmin_against_numpy_argmin_with_ties():
    rng = np.random.default_rng(12)
    P, n = 5, 4000
    depth = rng.random((P, n)).astype(np.float32)
    # ties: quantize a third of the pixels to a few levels so equal minima
    # across ranks are frequent; misses (depth 1) tie everywhere
    q = rng.random(n) < 0.33
    depth[:, q] = np.round(depth[:, q] * 3) / 3
    depth[:, rng.random(n) < 0.1] = 1.0
    imgs = rng.integers(0, 2 ** 32, (P, n), dtype=np.uint64).astype(np.uint32)
    out8, outz = O.composite_zmin(imgs, depth)
    win = np.argmin(depth, axis=0)  # first index among equal minima
    assert (np.sum(depth == depth.min(axis=0), axis=0) > 1).sum() > 500  # ties exercised
    assert np.array_equal(out8, imgs[win, np.arange(n)])
    assert np.array_equal(outz, depth.min(axis=0))
