"""Pins the CPU oracle (oracle/aloha_oracle.c) before it is trusted as the
checker of the HIP path.  The reference ships no tests, fixtures or golden
images (and its fakeOwl CPU backend is not vendored), so each building block
is pinned by what the mathematics fixes:

  * closed forms: constant and linear fields, TF texel centres and borders,
    the sample lattice, the alpha recurrence 1-(1-a)^n;
  * an independent brute-force march in float64 (explicit 8-corner weights,
    pure Python loops, tiny brick and image) that must agree with the
    oracle's FP64 data path to 1e-9 and its FP32 output to 2e-4;
  * invariants: premultiplied rgb <= alpha, 'over' associativity (rank images
    composited equal one multi-brick march when the jitter is fixed).
"""
import math

import numpy as np
import pytest

import aloha_amd as A
import oracle as O

f32 = np.float32


def _brick(vox_lo, dims, cell_hi=None):
    """dict brick descriptor with cell = voxel range (no ghost) unless given."""
    hi = tuple(vox_lo[a] + dims[a] for a in range(3))
    chi = cell_hi or hi
    return {"cell": (tuple(vox_lo), chi), "voxel": (tuple(vox_lo), hi),
            "space": (tuple(float(v) for v in vox_lo), tuple(float(v) for v in chi))}


# ---------------------------------------------------------------- TF
def test_tf_texel_centres_and_border():
    rng = np.random.default_rng(0)
    tf = rng.random((256, 4), dtype=np.float32)
    for k in [0, 1, 77, 254, 255]:
        v = f32((k + 0.5) / 256)
        assert np.array_equal(O.tf_probe(tf, v), tf[k])  # weight 0: exactly T[k]
    # border colour 0: half of the first/last entry at v = 0 / v = 1
    assert np.array_equal(O.tf_probe(tf, 0.0), f32(0.5) * tf[0])
    assert np.array_equal(O.tf_probe(tf, 1.0), f32(0.5) * tf[255])
    assert not O.tf_probe(tf, -0.5).any() and not O.tf_probe(tf, 1.5).any()
    # interior: linear interpolation between texel centres
    for v in rng.random(50):
        v = float(f32(v))
        x = v * 256 - 0.5
        i = math.floor(x)
        w = x - i
        e = lambda k: tf[k].astype(np.float64) if 0 <= k < 256 else np.zeros(4)
        want = (1 - w) * e(i) + w * e(i + 1)
        assert np.allclose(O.tf_probe(tf, f32(v)), want, atol=2e-6)


# ---------------------------------------------------------------- trilinear + gradient
def test_constant_field_is_exact():
    vox = np.full((5, 4, 6), 0.37, np.float32)
    b = _brick((0, 0, 0), (6, 4, 5))
    for p in [(0.0, 0.0, 0.0), (2.3, 1.7, 3.9), (5.9, 3.2, 4.5), (-0.4, 3.4, 0.1)]:
        s, s64 = O.sample_probe(b, vox, *p)
        assert s[0] == f32(0.37) and not s[1:].any()


def test_linear_field_value_gradient_and_clamp():
    nx, ny, nz = 9, 8, 7
    a, bb, c, d = 0.011, -0.023, 0.017, 0.4
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    vox = (a * x + bb * y + c * z + d).astype(np.float32)
    br = _brick((0, 0, 0), (nx, ny, nz))
    rng = np.random.default_rng(1)
    for _ in range(40):
        p = rng.uniform([0.6, 0.6, 0.6], [nx - 1.6, ny - 1.6, nz - 1.6])
        s, s64 = O.sample_probe(br, vox, *p.astype(np.float32))
        fv = a * p[0] + bb * p[1] + c * p[2] + d
        assert abs(s[0] - fv) < 1e-5 and abs(s64[0] - fv) < 1e-6
        assert np.allclose(s[1:], [a, bb, c], atol=1e-5)
        assert np.allclose(s64[1:], [a, bb, c], atol=1e-7)
    # x = 0.25: the x-0.5 tap clamps to voxel 0, so gx = f(0.75) - f(0) = 0.75a
    s, s64 = O.sample_probe(br, vox, 0.25, 3.0, 3.0)
    assert abs(s64[1] - 0.75 * a) < 1e-7
    # beyond the upper face every x tap clamps to voxel nx-1
    s, s64 = O.sample_probe(br, vox, nx - 0.5 + 0.7, 3.0, 3.0)
    assert abs(s64[1]) < 1e-12


def test_voxel_range_lower_is_the_texel_origin():
    rng = np.random.default_rng(2)
    vox = rng.random((4, 5, 6), dtype=np.float32)
    b0 = _brick((0, 0, 0), (6, 5, 4))
    b1 = _brick((10, 20, 30), (6, 5, 4))
    p = (2.3, 1.4, 2.2)
    s0, _ = O.sample_probe(b0, vox, *p)
    # the probe takes texel-space coordinates: same texel, same result
    s1, _ = O.sample_probe(b1, vox, *p)
    assert np.array_equal(s0, s1)
    # the value tap is a convex combination of the 8 corners
    corners = vox[2:4, 1:3, 2:4]
    assert corners.min() - 1e-7 <= s0[0] <= corners.max() + 1e-7


# ---------------------------------------------------------------- lattice, slab
def test_first_sample_lattice():
    rng = np.random.default_rng(3)
    for _ in range(200):
        t0 = f32(rng.uniform(0, 3000))
        dt = f32(rng.choice([0.25, 0.5, 1.0, 2.0]))
        j = f32(rng.random())
        t = O.first_sample_t(t0, dt, j)
        k = (t / dt) - j
        assert abs(k - round(k)) < 1e-3 * max(1.0, t / dt * 1e-4)  # on the lattice dt*(j+k)
        assert t >= t0 and t - dt < t0 + 1e-3


def test_slab_closed_form():
    hit, t0, t1 = O.slab([-10, 0.5, 0.5], [1, 1e-5, 1e-5], 0.0, 1e10, [0, 0, 0], [1, 1, 1])
    assert hit and t0 == 10.0 and t1 == 11.0
    hit, t0, t1 = O.slab([0.5, 0.5, 0.5], [0, 0, 1], 0.0, 1e10, [0, 0, 0], [1, 1, 1])
    assert hit and t0 == 0.0 and t1 == 0.5  # origin inside: clipped to tmin
    hit, _, _ = O.slab([-10, 5, 0.5], [1, 1e-5, 1e-5], 0.0, 1e10, [0, 0, 0], [1, 1, 1])
    assert not hit


# ---------------------------------------------------------------- RNG / pack
# TEA (Wheeler & Needham 1994) known-answer vectors at 32 cycles, blocks and
# keys read big-endian (the published test vectors of the cipher): an all-zero
# key and block, and a key whose four words differ, so a swapped key word, a
# wrong shift or a misplaced round sum fails.
TEA_KAT = [
    ((0x00000000, 0x00000000), (0, 0, 0, 0), (0x41EA3A0A, 0x94BAA940)),
    ((0x01020304, 0x05060708), (0, 0, 0, 0), (0x6A2F9CF3, 0xFCCF3C55)),
    ((0x00000000, 0x00000000), (0x01234567, 0x12345678, 0x23456789, 0x3456789A),
     (0x34E943B0, 0x900F5DCB)),
    ((0x01020304, 0x05060708), (0x01234567, 0x12345678, 0x23456789, 0x3456789A),
     (0x773DC179, 0x878A81C0)),
]
# OWL's LCG<4> key (the OptiX SDK's tea<N>, Zafar, Olano & Curtis 2010)
LCG_KEY = (0xA341316C, 0xC8013EA4, 0xAD90777D, 0x7E95761E)


def test_tea_known_answers():
    for block, key, want in TEA_KAT:
        assert O.tea(*block, key, 32) == want


def test_rng_seed_is_four_tea_rounds_of_pixel_and_accum_id():
    # Random random(pixelID, lp.accumID) (deviceCode.cpp:354): state = v0 of
    # TEA<4>(pixelID, accumID) under the LCG key
    for pixel, frame in [(0, 0), (1, 0), (0, 1), (512 * 300 + 17, 3), (2 ** 20 - 1, 7)]:
        assert O.rng_seed(pixel, frame) == O.tea(pixel, frame, LCG_KEY, 4)[0]
    # 4 rounds differ from the full cipher and from other round counts
    assert O.rng_seed(5, 2) != O.tea(5, 2, LCG_KEY, 32)[0]
    assert O.rng_seed(5, 2) != O.tea(5, 2, LCG_KEY, 3)[0]


def test_rng_draws_are_low_24_bits_of_the_numerical_recipes_lcg():
    for pixel, frame in [(0, 0), (777, 1), (1048575, 9)]:
        s = O.rng_seed(pixel, frame)
        want = []
        for _ in range(6):
            s = (1664525 * s + 1013904223) % 2 ** 32
            want.append((s & 0xFFFFFF) / 2.0 ** 24)
        assert O.rng_stream(pixel, frame, 6) == want


def test_rng_is_24bit_uniform():
    draws = np.array([d for p in range(200) for d in O.rng_stream(p, 3, 50)], np.float64)
    assert draws.min() >= 0 and draws.max() < 1
    assert np.all(draws * 2 ** 24 == np.floor(draws * 2 ** 24))
    assert abs(draws.mean() - 0.5) < 0.01 and abs(draws.var() - 1 / 12) < 0.005
    assert O.rng_stream(5, 0, 4) != O.rng_stream(6, 0, 4) != O.rng_stream(5, 1, 4)


def test_make_rgba():
    assert O.make_rgba([0, 0, 0, 0]) == 0
    assert O.make_rgba([1, 1, 1, 1]) == 0xFFFFFFFF
    assert O.make_rgba([0.5, 0, 0, 0]) == 128                 # R in the low byte
    assert O.make_rgba([0, 0, 0, 0.25]) == 64 << 24
    assert O.make_rgba([-1, 2, 0.999, 1 / 256]) == (0 | 255 << 8 | 255 << 16 | 1 << 24)


# ---------------------------------------------------------------- heat map
_HEAT_CLOCK = 1 << 20  # a power of two: heat_scale = t / clock is exact


@pytest.mark.parametrize("t, rgb", [(0.0, (0, 0, .5)), (.25, (0, .625, .625)), (.5, (0, .75, 0)),
                                    (.75, (.875, .875, 0)), (1.0, (1, 0, 0))])
def test_heat_map_ramp_closed_forms(t, rgb):
    # temperature_to_rgb (deviceCode.cpp:300-320): hue h = 4/6 (1 - t) through
    # hue_to_rgb's three ramps, value .5 + .5 t.  Over an empty pixel the blend
    # over((heat, .5), 0) leaves (heat, .5).
    cam = _cam_looking_down_z(4, 3)
    tf = np.zeros((4, 4), np.float32)
    o = O.render_rank(cam, tf, [], 4, 3, heat_clock=_HEAT_CLOCK, heat_scale=t / _HEAT_CLOCK)
    assert np.allclose(o["rgba"][..., :3], rgb, atol=1e-6)
    assert (o["rgba"][..., 3] == 0.5).all()
    # spp divides the clock before the ramp, and the spp sum after the blend
    o2 = O.render_rank(cam, tf, [], 4, 3, spp=2, heat_clock=_HEAT_CLOCK,
                       heat_scale=2 * t / _HEAT_CLOCK)
    assert np.allclose(o2["rgba"][..., :3], np.asarray(rgb) / 2, atol=1e-6)
    assert (o2["rgba"][..., 3] == 0.25).all()


def test_heat_map_blends_over_the_pixel():
    # accumColor = over((heat, .5), accumColor) = heat + .5 accumColor (:36-39, :402)
    vox = np.full((12, 6, 7), 0.5, np.float32)
    b = _brick((0, 0, 0), (7, 6, 12))
    tf = np.tile(np.array([[0.8, 0.4, 0.2, 0.3]], np.float32), (8, 1))
    cam = _cam_looking_down_z(6, 5, org=(3.0, 2.5, -6.0), fov=0.2)
    plain = O.render_rank(cam, tf, [(b, vox)], 6, 5, jitter_mode=1, jitter_value=0.5)
    hot = O.render_rank(cam, tf, [(b, vox)], 6, 5, jitter_mode=1, jitter_value=0.5,
                        heat_clock=_HEAT_CLOCK, heat_scale=0.5 / _HEAT_CLOCK)
    assert plain["rgba"][..., 3].min() > 0.1  # every ray crosses the brick
    heat = np.array([0, .75, 0], np.float32)
    want = np.empty_like(plain["rgba"])
    want[..., :3] = heat + f32(0.5) * plain["rgba"][..., :3]
    want[..., 3] = f32(0.5) + f32(0.5) * plain["rgba"][..., 3]
    assert np.allclose(hot["rgba"], want, atol=1e-6)
    assert np.array_equal(hot["depth"], plain["depth"])


# ---------------------------------------------------------------- composite
def test_over_associativity_and_opaque_front():
    rng = np.random.default_rng(4)
    a = rng.random((3, 4, 5, 1), dtype=np.float32)
    imgs = (rng.random((3, 4, 5, 4), dtype=np.float32) * a).astype(np.float32)
    imgs[..., 3:] = a
    full, _ = O.composite_over(imgs, [2, 0, 1])
    front, _ = O.composite_over(imgs[[2, 0]], [0, 1])
    both, _ = O.composite_over(np.stack([front, imgs[1]]), [0, 1])
    assert np.allclose(full, both, atol=1e-6)
    d = O.composite_over_d(imgs, [2, 0, 1])
    assert np.allclose(full, d, atol=1e-6)
    opaque = imgs.copy()
    opaque[0, ..., 3] = 1.0
    out, _ = O.composite_over(opaque, [0, 1, 2])
    assert np.array_equal(out, opaque[0])
    assert np.array_equal(O.composite_over(imgs[:1], [0])[0], imgs[0])


# ---------------------------------------------------------------- whole-ray closed forms
def _cam_looking_down_z(W, H, org=(3.0, 2.5, -6.0), fov=0.5):
    d00 = np.array([-fov, -fov, 1.0], np.float32)
    return {"org": np.array(org, np.float32), "dir_00": d00,
            "dir_du": np.array([2 * fov / W, 0, 0], np.float32),
            "dir_dv": np.array([0, 2 * fov / H, 0], np.float32)}


def test_constant_field_alpha_recurrence():
    vox = np.full((40, 6, 7), 0.5, np.float32)
    b = _brick((0, 0, 0), (7, 6, 40))
    tf = np.tile(np.array([[0.8, 0.4, 0.2, 0.3]], np.float32), (16, 1))
    cam = _cam_looking_down_z(9, 8)
    o = O.render_rank(cam, tf, [(b, vox)], 9, 8, dt=0.5, max_depth=100.0, jitter_mode=1,
                      jitter_value=0.0)
    ns = o["nsamples"]
    # alpha 0.3 per sample: 1-(0.7)^13 >= .99 > 1-(0.7)^12, so at most 13 samples
    assert ns.max() == 13 and (ns == 13).sum() > 10
    for n_, rgba in zip(ns.ravel(), o["rgba"].reshape(-1, 4)):
        alpha = 1 - 0.7 ** int(n_)
        assert abs(rgba[3] - alpha) < 1e-6
        # N = 0 (zero gradient) gives shade .1; colour is premultiplied
        assert np.allclose(rgba[:3], 0.1 * np.array([0.8, 0.4, 0.2]) * alpha, atol=1e-6)


def _brute_force(cam, tf, bricks, W, H, dt, jitter, max_depth, closed_lattice=False, amax=None):
    """float64 re-derivation of renderFrame for fixed jitter: float32 ray
    setup and sample lattice (the reference's control), float64 data with
    explicit corner weights.  Pure Python; tiny inputs only.

    closed_lattice / amax give the differentiable form instead: samples at
    t_k = dt*(j + (k0 + k)) in float32 (k0 the first lattice index at or
    past the entry), each sample's opacity clamped to amax."""
    n_tf = tf.shape[0]

    def tfl(v):
        x = v * n_tf - 0.5
        i = math.floor(x)
        w = x - i
        e = lambda k: tf[k].astype(np.float64) if 0 <= k < n_tf else np.zeros(4)
        return (1 - w) * e(i) + w * e(i + 1)

    def tri(vox, n, p):
        base = [math.floor(p[a]) for a in range(3)]
        fr = [p[a] - base[a] for a in range(3)]
        acc = 0.0
        for dz in (0, 1):
            for dy in (0, 1):
                for dx in (0, 1):
                    w = ((fr[0] if dx else 1 - fr[0]) * (fr[1] if dy else 1 - fr[1]) *
                         (fr[2] if dz else 1 - fr[2]))
                    ix = min(max(base[0] + dx, 0), n[0] - 1)
                    iy = min(max(base[1] + dy, 0), n[1] - 1)
                    iz = min(max(base[2] + dz, 0), n[2] - 1)
                    acc += w * float(vox[iz, iy, ix])
        return acc

    def slab(o, d, tmin, lo, hi):
        nr, fr = [], []
        for a in range(3):
            l = (f32(lo[a]) - o[a]) / d[a]
            h = (f32(hi[a]) - o[a]) / d[a]
            nr.append(min(l, h))
            fr.append(max(l, h))
        return max(f32(tmin), max(nr)), min(f32(1e10), min(fr))

    rgba = np.zeros((H, W, 4))
    ns = np.zeros((H, W), np.int64)
    depth = np.ones((H, W), np.float32)
    o = np.array(cam["org"], np.float32)
    for py in range(H):
        for px in range(W):
            su, sv = f32(px) + f32(0.5), f32(py) + f32(0.5)
            d = [(f32(cam["dir_00"][a]) + su * f32(cam["dir_du"][a])) + sv * f32(cam["dir_dv"][a])
                 for a in range(3)]
            ln = np.sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2])
            d = [f32(x / ln) for x in d]
            d = [f32(1e-5) if abs(x) < f32(1e-5) else x for x in d]
            c = np.zeros(4)
            z = f32(1.0)
            tmin = f32(0.0)
            while True:
                best = None
                for (desc, vox) in bricks:
                    t0, t1 = slab(o, d, tmin, desc["space"][0], desc["space"][1])
                    if not t0 < t1:
                        continue
                    if t0 > tmin:
                        rep, tn = t0, t0
                    elif t1 > tmin:
                        rep, tn = t1, f32(0)
                    else:
                        continue
                    if best is None or rep < best[0]:
                        best = (rep, tn, t0, t1, desc, vox)
                if best is None:
                    break
                _, tn, t0, t1, desc, vox = best
                tn = f32(tn + tmin)
                segs = np.floor((t0 - f32(dt) * f32(jitter)) / f32(dt))
                t = f32(dt) * (f32(jitter) + segs)
                k0 = segs
                if t < t0:
                    t = f32(t + f32(dt))
                    k0 = f32(segs + f32(1))
                k = 0
                if closed_lattice:
                    t = f32(dt) * (f32(jitter) + (f32(k0) + f32(k)))
                vl = desc["voxel"][0]
                n = [desc["voxel"][1][a] - vl[a] for a in range(3)]
                while t < t1 and c[3] < 0.99:
                    ns[py, px] += 1
                    p = [float(o[a]) + float(t) * float(d[a]) - vl[a] for a in range(3)]
                    g = []
                    for a in range(3):
                        pp, pm = list(p), list(p)
                        pp[a] += 0.5
                        pm[a] -= 0.5
                        g.append(tri(vox, n, pp) - tri(vox, n, pm))
                    e = [x + 1e-4 for x in g]
                    L = math.sqrt(sum(x * x for x in e))
                    shade = 0.1 + 0.9 * abs(sum(g[a] / L * float(d[a]) for a in range(3)))
                    s = tfl(tri(vox, n, p))
                    if amax is not None:
                        s[3] = min(s[3], float(f32(amax)))
                    w = (1 - c[3]) * s[3]
                    c[:3] += w * s[:3] * shade
                    c[3] += w
                    k += 1
                    if closed_lattice:
                        t = f32(dt) * (f32(jitter) + (f32(k0) + f32(k)))
                    else:
                        t = f32(t + f32(dt))
                z = min(z, f32(tn / f32(max_depth)))
                tmin = f32(t1 + f32(1e-20))
                if not c[3] < 0.99:
                    break
            rgba[py, px] = c
            depth[py, px] = z
    return rgba, ns, depth


@pytest.mark.parametrize("alpha_scale,dt", [(0.05, 0.5), (1.0, 0.7)])
def test_brute_force_two_bricks(alpha_scale, dt):
    rng = np.random.default_rng(5)
    raw = rng.random((6, 5, 8), dtype=np.float32)
    descs = A.split_volume((8, 5, 6), 2)
    bricks = []
    for dsc in descs:
        dd = dsc.as_dict()
        lo, hi = dd["voxel"]
        bricks.append(({"cell": dd["cell"], "voxel": dd["voxel"], "space": dd["space"]},
                       np.ascontiguousarray(raw[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]])))
    tf = rng.random((32, 4), dtype=np.float32)
    tf[:, 3] *= alpha_scale
    W, H = 7, 6
    cam = {"org": np.array([-4.0, 1.0, -3.0], np.float32),
           "dir_00": np.array([0.7, -0.1, 0.35], np.float32),
           "dir_du": np.array([0.05, 0.0, 0.02], np.float32),
           "dir_dv": np.array([0.0, 0.06, 0.01], np.float32)}
    o = O.render_rank(cam, tf, bricks, W, H, dt=dt, max_depth=50.0, jitter_mode=1,
                      jitter_value=0.3, want64=True)
    rgba, ns, depth = _brute_force(cam, tf, bricks, W, H, dt, 0.3, 50.0)
    assert ns.sum() > 20
    assert np.array_equal(o["nsamples"], ns)
    assert np.array_equal(o["depth"], depth)
    assert np.abs(o["rgba64"] - rgba).max() < 1e-9
    assert np.abs(o["rgba"] - rgba).max() < 2e-4
    # premultiplied: rgb <= alpha
    assert np.all(o["rgba"][..., :3] <= o["rgba"][..., 3:] + 1e-6)


@pytest.mark.parametrize("alpha_scale,dt", [(0.05, 0.5), (1.0, 0.7)])
def test_diff_forward64_against_brute_force(alpha_scale, dt):
    """orc_diff_render64 (the float64 forward the adjoint's finite differences
    use) and orc_diff_render (float) against the pure-Python float64 march of
    the differentiable form: closed-form lattice, opacity clamped to .9999."""
    rng = np.random.default_rng(6)
    raw = rng.random((6, 5, 8), dtype=np.float32)
    bricks = []
    for dsc in A.split_volume((8, 5, 6), 2):
        dd = dsc.as_dict()
        lo, hi = dd["voxel"]
        bricks.append(({"cell": dd["cell"], "voxel": dd["voxel"], "space": dd["space"]},
                       np.ascontiguousarray(raw[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]])))
    tf = rng.random((32, 4), dtype=np.float32)
    tf[:, 3] *= alpha_scale
    tf[4:9, 3] = 1.0  # opaque entries: the .9999 clamp is taken
    W, H = 7, 6
    cam = {"org": np.array([-4.0, 1.0, -3.0], np.float32),
           "dir_00": np.array([0.7, -0.1, 0.35], np.float32),
           "dir_du": np.array([0.05, 0.0, 0.02], np.float32),
           "dir_dv": np.array([0.0, 0.06, 0.01], np.float32)}
    kw = dict(dt=dt, jitter_mode=1, jitter_value=0.3)
    rgba32, ns = O.diff_render(cam, tf, bricks, W, H, **kw)
    rgba64 = O.diff_render64(cam, tf, bricks, W, H, **kw)
    ref, ref_ns, _ = _brute_force(cam, tf, bricks, W, H, dt, 0.3, 50.0, closed_lattice=True,
                                  amax=0.9999)
    assert ns.sum() > 20
    assert np.array_equal(ns, ref_ns)
    assert np.abs(rgba64 - ref).max() < 1e-9
    assert np.abs(rgba32 - rgba64).max() < 1e-5
    # the clamp is exercised: without it some pixel's colour changes
    unclamped, _, _ = _brute_force(cam, tf, bricks, W, H, dt, 0.3, 50.0, closed_lattice=True)
    assert np.abs(unclamped - ref).max() > 1e-6


def test_rank_images_composite_to_single_march():
    """over-associativity: per-brick rank images composited in kd order equal
    the single multi-brick march (fixed jitter, no early termination)."""
    from aloha_amd import workloads as WL
    wl = WL.blobs128(num_ranks=2, W=64, H=64, size=32)
    wl.tf[:, 3] *= 0.02
    kw = dict(dt=wl.dt, max_depth=wl.max_depth, jitter_mode=1, jitter_value=0.5)
    ranks = [O.render_rank(wl.cam, wl.tf, [(wl.bricks[i], wl.voxels[i]) for i in r], wl.W, wl.H,
                           **kw) for r in wl.rank_bricks]
    comp, _ = O.composite_over(np.stack([r["rgba"] for r in ranks]), wl.order)
    one = O.render_rank(wl.cam, wl.tf, list(zip(wl.bricks, wl.voxels)), wl.W, wl.H, **kw)
    assert one["nsamples"].sum() == sum(r["nsamples"].sum() for r in ranks) > 1000
    assert np.abs(comp - one["rgba"]).max() < 1e-5


def test_fp32_vs_fp64_error_is_small():
    from aloha_amd import workloads as WL
    wl = WL.blobs128(W=96, H=96)
    o = O.render_rank(wl.cam, wl.tf, [(wl.bricks[0], wl.voxels[0])], wl.W, wl.H, dt=wl.dt,
                      max_depth=wl.max_depth, want64=True)
    assert o["nsamples"].sum() > 1000
    assert np.abs(o["rgba"] - o["rgba64"]).max() < 1e-3
