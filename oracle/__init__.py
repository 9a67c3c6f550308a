"""ctypes view of the CPU oracle (oracle/aloha_oracle.c).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.  Build it with
`make -C oracle` (or `python -m aloha_amd.build`).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libaloha_oracle.so"


class OrcBrick(C.Structure):
    _fields_ = [("cell_lo", C.c_int32 * 3), ("cell_hi", C.c_int32 * 3),
                ("vox_lo", C.c_int32 * 3), ("vox_hi", C.c_int32 * 3),
                ("space_lo", C.c_float * 3), ("space_hi", C.c_float * 3),
                ("dtype", C.c_int32), ("pad_", C.c_int32), ("voxels", C.c_void_p)]


class OrcParams(C.Structure):
    _fields_ = [("org", C.c_float * 3), ("dir00", C.c_float * 3), ("du", C.c_float * 3),
                ("dv", C.c_float * 3), ("dt", C.c_float), ("max_depth", C.c_float),
                ("W", C.c_int32), ("H", C.c_int32), ("spp", C.c_int32), ("accum_id", C.c_int32),
                ("island_id", C.c_int32), ("num_islands", C.c_int32),
                ("jitter_mode", C.c_int32), ("jitter_value", C.c_float),
                ("crosshair", C.c_int32), ("tf_n", C.c_int32), ("tf", C.POINTER(C.c_float)),
                ("heat", C.c_int32), ("heat_scale", C.c_float), ("heat_clock", C.c_uint32),
                ("pad_", C.c_int32)]


class OrcDomain(C.Structure):
    _fields_ = [("cell_lo", C.c_int32 * 3), ("cell_hi", C.c_int32 * 3),
                ("vox_lo", C.c_int32 * 3), ("vox_hi", C.c_int32 * 3),
                ("space_lo", C.c_float * 3), ("space_hi", C.c_float * 3)]


class OrcKD(C.Structure):
    _fields_ = [("axis", C.c_int32), ("plane", C.c_float), ("child1", C.c_int32),
                ("child2", C.c_int32)]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} not built: run `make -C oracle`")
        _lib = C.CDLL(str(LIB_PATH))
        _lib.orc_make_rgba.restype = C.c_uint32
        _lib.orc_rng_seed.restype = C.c_uint32
        _lib.orc_rng_draw.restype = C.c_float
        _lib.orc_rng_seed.argtypes = [C.c_uint32, C.c_uint32]
        _lib.orc_tri_render.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_longlong,
                                        C.c_int, C.c_int] + [C.c_void_p] * 5 + [C.c_int]
        _lib.orc_rng_draw.argtypes = [C.POINTER(C.c_uint32)]
        _lib.orc_composite_over.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_longlong,
                                            C.c_void_p, C.c_void_p]
        _lib.orc_composite_over_d.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_longlong,
                                              C.c_void_p]
        _lib.orc_composite_over_u8.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_longlong,
                                               C.c_void_p]
        _lib.orc_composite_zmin.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_longlong,
                                            C.c_void_p, C.c_void_p]
    return _lib


def _i3(v):
    return (C.c_int32 * 3)(*[int(x) for x in v])


def _f3(v):
    return (C.c_float * 3)(*[float(x) for x in v])


def split_volume(dims, n):
    cap = 4 * n + 8
    out = (OrcDomain * cap)()
    k = lib().orc_split_volume(_i3(dims), int(n), out, cap)
    if k < 0:
        raise ValueError("orc_split_volume failed")
    return [{"cell": (tuple(d.cell_lo), tuple(d.cell_hi)), "voxel": (tuple(d.vox_lo), tuple(d.vox_hi)),
             "space": (tuple(d.space_lo), tuple(d.space_hi))} for d in out[:k]]


def volume_get(raw: np.ndarray, bpc: int, first, last):
    raw = np.ascontiguousarray(raw)
    ext = [int(last[a]) - int(first[a]) for a in range(3)]
    out = np.empty(ext[::-1], dtype=np.float32)
    vr = (C.c_float * 2)(np.finfo(np.float32).max, -np.finfo(np.float32).max)
    if lib().orc_volume_get(C.c_void_p(raw.ctypes.data), _i3(raw.shape[::-1]), bpc, _i3(first),
                            _i3(last), C.c_void_p(out.ctypes.data), vr) != 0:
        raise ValueError("bad bpc")
    return out, (vr[0], vr[1])


def partition_kd(domains: np.ndarray, P: int):
    """domains: (n,6) float32 (lower xyz, upper xyz).  Returns (per_rank, kd)."""
    d = np.ascontiguousarray(domains, dtype=np.float32)
    n = d.shape[0]
    counts = (C.c_int * P)()
    ids = (C.c_int * max(1, n))()
    cap = 4 * n + 8
    kd = (OrcKD * cap)()
    nk = lib().orc_partition_kd(C.c_void_p(d.ctypes.data), n, P, counts, ids, kd, cap)
    if nk < 0:
        raise ValueError("kd cap too small" if nk == -1 else "bricks cannot be separated")
    per, m = [], 0
    for r in range(P):
        per.append(list(ids[m:m + counts[r]]))
        m += counts[r]
    return per, [(k.axis, k.plane, k.child1, k.child2) for k in kd[:nk]]


def partition_rr(n, P):
    out = (C.c_int * max(1, n))()
    lib().orc_partition_rr(n, P, out)
    return list(out[:n])


def composite_order(kd, ref):
    arr = (OrcKD * max(1, len(kd)))(*[OrcKD(a, p, c1, c2) for a, p, c1, c2 in kd])
    cap = 4096
    out = (C.c_int * cap)()
    m = lib().orc_composite_order(arr, len(kd), _f3(ref), out, cap)
    return list(out[:m])


def _orc_brick(desc, voxels: np.ndarray) -> OrcBrick:
    b = OrcBrick()
    if isinstance(desc, dict):
        cell, vox, space = desc["cell"], desc["voxel"], desc["space"]
    else:
        cell = (tuple(desc.cell_range.lower), tuple(desc.cell_range.upper))
        vox = (tuple(desc.voxel_range.lower), tuple(desc.voxel_range.upper))
        space = (tuple(desc.space_range.lower), tuple(desc.space_range.upper))
    b.cell_lo, b.cell_hi = _i3(cell[0]), _i3(cell[1])
    b.vox_lo, b.vox_hi = _i3(vox[0]), _i3(vox[1])
    b.space_lo, b.space_hi = _f3(space[0]), _f3(space[1])
    b.dtype = _DT[voxels.dtype]
    b.voxels = voxels.ctypes.data
    assert voxels.shape[::-1] == tuple(vox[1][a] - vox[0][a] for a in range(3))
    return b


def sample_probe(desc, voxels: np.ndarray, x: float, y: float, z: float):
    """(value, gx, gy, gz) in float32 and float64 at texel-space (x,y,z)."""
    voxels = np.ascontiguousarray(voxels)
    b = _orc_brick(desc, voxels)
    out = (C.c_float * 4)()
    out64 = (C.c_double * 4)()
    lib().orc_sample_probe(C.byref(b), C.c_float(x), C.c_float(y), C.c_float(z), out, out64)
    return np.array(out[:], np.float32), np.array(out64[:], np.float64)


def tf_probe(tf: np.ndarray, v: float):
    tf = np.ascontiguousarray(tf, dtype=np.float32)
    out = (C.c_float * 4)()
    lib().orc_tf_probe(tf.ctypes.data_as(C.POINTER(C.c_float)), tf.shape[0], C.c_float(v), out)
    return np.array(out[:], np.float32)


def first_sample_t(t0: float, dt: float, jitter: float) -> float:
    f = lib().orc_first_sample_t
    f.restype = C.c_float
    f.argtypes = [C.c_float, C.c_float, C.c_float]
    return f(t0, dt, jitter)


def slab(o, d, tmin, tmax, lo, hi):
    t = (C.c_float * 2)()
    f = lib().orc_slab_probe
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_float, C.c_float, C.c_void_p, C.c_void_p, C.c_void_p]
    hit = f(_f3(o), _f3(d), tmin, tmax, _f3(lo), _f3(hi), t)
    return bool(hit), t[0], t[1]


def tea(v0: int, v1: int, key, rounds: int):
    """TEA encryption of the block (v0, v1) under the 4-word key, `rounds` cycles."""
    v = (C.c_uint32 * 2)(v0, v1)
    lib().orc_tea(v, (C.c_uint32 * 4)(*key), int(rounds))
    return int(v[0]), int(v[1])


def rng_seed(pixel: int, frame: int) -> int:
    return int(lib().orc_rng_seed(pixel, frame))


def rng_stream(pixel: int, frame: int, count: int):
    s = C.c_uint32(lib().orc_rng_seed(pixel, frame))
    return [lib().orc_rng_draw(C.byref(s)) for _ in range(count)]


def make_rgba(c) -> int:
    return lib().orc_make_rgba((C.c_float * 4)(*[float(x) for x in c]))


_DT = {np.dtype(np.uint8): 1, np.dtype(np.uint16): 2, np.dtype(np.float32): 4}


def render_rank(cam: dict, tf: np.ndarray, bricks, W: int, H: int, dt=0.5, max_depth=1.0,
                spp=1, accum_id=0, island_id=0, num_islands=1, jitter_mode=0, jitter_value=0.0,
                crosshair=False, rows=None, accum=None, want64=False, touched=False,
                threads=0, heat_clock=None, heat_scale=1e-5):
    """bricks: list of (desc, voxels) where desc has cell_range/voxel_range/
    space_range (aloha_amd.BrickDesc or dict) and voxels is (z,y,x) u8/u16/f32.
    Returns dict with rgba (H,W,4), rgba8 (H,W) u32, depth, nsamples, accum,
    and rgba64 / touched (list of per-brick u8 masks) when requested."""
    tf = np.ascontiguousarray(tf, dtype=np.float32)
    keep = [tf]
    br = (OrcBrick * max(1, len(bricks)))()
    for i, (d, v) in enumerate(bricks):
        v = np.ascontiguousarray(v)
        keep.append(v)
        if isinstance(d, dict):
            cell, vox, space = d["cell"], d["voxel"], d["space"]
        else:
            cell = (tuple(d.cell_range.lower), tuple(d.cell_range.upper))
            vox = (tuple(d.voxel_range.lower), tuple(d.voxel_range.upper))
            space = (tuple(d.space_range.lower), tuple(d.space_range.upper))
        br[i].cell_lo, br[i].cell_hi = _i3(cell[0]), _i3(cell[1])
        br[i].vox_lo, br[i].vox_hi = _i3(vox[0]), _i3(vox[1])
        br[i].space_lo, br[i].space_hi = _f3(space[0]), _f3(space[1])
        br[i].dtype = _DT[v.dtype]
        br[i].voxels = v.ctypes.data
        assert v.shape[::-1] == tuple(vox[1][a] - vox[0][a] for a in range(3))
    p = OrcParams()
    p.org, p.dir00, p.du, p.dv = (_f3(cam[k]) for k in ("org", "dir_00", "dir_du", "dir_dv"))
    p.dt, p.max_depth, p.W, p.H = dt, max_depth, W, H
    p.spp, p.accum_id, p.island_id, p.num_islands = spp, accum_id, island_id, num_islands
    p.jitter_mode, p.jitter_value, p.crosshair = jitter_mode, jitter_value, int(crosshair)
    p.tf_n = tf.shape[0]
    p.tf = tf.ctypes.data_as(C.POINTER(C.c_float))
    _heat(p, heat_clock, heat_scale)
    row0, row1 = (0, H) if rows is None else rows
    out = {
        "rgba": np.zeros((H, W, 4), np.float32),
        "rgba8": np.zeros((H, W), np.uint32),
        "depth": np.zeros((H, W), np.float32),
        "nsamples": np.zeros((H, W), np.uint32),
        "accum": np.zeros((H, W, 4), np.float32) if accum is None else np.ascontiguousarray(accum, np.float32),
    }
    if want64:
        out["rgba64"] = np.zeros((H, W, 4), np.float64)
    tptr = None
    if touched:
        masks = [np.zeros(np.prod(v.shape), np.uint8) for _, v in bricks]
        out["touched"] = masks
        tptr = (C.c_void_p * len(masks))(*[m.ctypes.data for m in masks])
    rc = lib().orc_render_rank(
        C.byref(p), br, len(bricks), row0, row1, C.c_void_p(out["rgba"].ctypes.data),
        C.c_void_p(out["rgba8"].ctypes.data), C.c_void_p(out["depth"].ctypes.data),
        C.c_void_p(out["accum"].ctypes.data), C.c_void_p(out["nsamples"].ctypes.data),
        C.c_void_p(out["rgba64"].ctypes.data) if want64 else None, tptr, int(threads))
    if rc != 0:
        raise ValueError(f"orc_render_rank failed ({rc})")
    return out


def _heat(p, heat_clock, heat_scale):
    """Heat map with every pixel's clock delta = heat_clock (None: off)."""
    if heat_clock is not None:
        p.heat, p.heat_scale, p.heat_clock = 1, heat_scale, int(heat_clock)


def composite_over(imgs: np.ndarray, order):
    """imgs: (P,H,W,4) float32 -> (rgba (H,W,4), rgba8 (H,W))."""
    imgs = np.ascontiguousarray(imgs, dtype=np.float32)
    P = imgs.shape[0]
    npix = int(np.prod(imgs.shape[1:3]))
    o = (C.c_int * P)(*order)
    out = np.zeros(imgs.shape[1:], np.float32)
    out8 = np.zeros(imgs.shape[1:3], np.uint32)
    lib().orc_composite_over(imgs.ctypes.data, o, P, npix, out.ctypes.data, out8.ctypes.data)
    return out, out8


def composite_over_d(imgs: np.ndarray, order):
    imgs = np.ascontiguousarray(imgs, dtype=np.float64)
    P = imgs.shape[0]
    npix = int(np.prod(imgs.shape[1:3]))
    o = (C.c_int * P)(*order)
    out = np.zeros(imgs.shape[1:], np.float64)
    lib().orc_composite_over_d(imgs.ctypes.data, o, P, npix, out.ctypes.data)
    return out


def composite_over_u8(imgs8: np.ndarray, order):
    imgs8 = np.ascontiguousarray(imgs8, dtype=np.uint32)
    P = imgs8.shape[0]
    npix = int(np.prod(imgs8.shape[1:]))
    o = (C.c_int * P)(*order)
    out8 = np.zeros(imgs8.shape[1:], np.uint32)
    lib().orc_composite_over_u8(imgs8.ctypes.data, o, P, npix, out8.ctypes.data)
    return out8


def composite_zmin(imgs8: np.ndarray, depths: np.ndarray):
    imgs8 = np.ascontiguousarray(imgs8, dtype=np.uint32)
    depths = np.ascontiguousarray(depths, dtype=np.float32)
    P = imgs8.shape[0]
    npix = int(np.prod(imgs8.shape[1:]))
    out8 = np.zeros(imgs8.shape[1:], np.uint32)
    outz = np.zeros(imgs8.shape[1:], np.float32)
    lib().orc_composite_zmin(imgs8.ctypes.data, depths.ctypes.data, P, npix, out8.ctypes.data,
                             outz.ctypes.data)
    return out8, outz


def _params(cam, tf, W, H, dt, jitter_mode, jitter_value):
    p = OrcParams()
    p.org, p.dir00, p.du, p.dv = (_f3(cam[k]) for k in ("org", "dir_00", "dir_du", "dir_dv"))
    p.dt, p.max_depth, p.W, p.H = dt, 1.0, W, H
    p.spp, p.accum_id, p.island_id, p.num_islands = 1, 0, 0, 1
    p.jitter_mode, p.jitter_value, p.crosshair = jitter_mode, jitter_value, 0
    p.tf_n = tf.shape[0]
    p.tf = tf.ctypes.data_as(C.POINTER(C.c_float))
    return p


def diff_render(cam, tf, bricks, W, H, dt=0.5, jitter_mode=0, jitter_value=0.0, threads=0):
    """Float restatement of the differentiable forward (closed-form lattice,
    opacity clamped to 0.9999).  Returns (rgba (H,W,4), nsamples (H,W))."""
    tf = np.ascontiguousarray(tf, dtype=np.float32)
    vox = [np.ascontiguousarray(v) for _, v in bricks]
    br = (OrcBrick * len(bricks))(*[_orc_brick(d, v) for (d, _), v in zip(bricks, vox)])
    p = _params(cam, tf, W, H, dt, jitter_mode, jitter_value)
    rgba = np.zeros((H, W, 4), np.float32)
    ns = np.zeros((H, W), np.uint32)
    rc = lib().orc_diff_render(C.byref(p), br, len(bricks), C.c_void_p(rgba.ctypes.data),
                               C.c_void_p(ns.ctypes.data), int(threads))
    if rc:
        raise ValueError("orc_diff_render failed")
    return rgba, ns


def diff_adjoint(cam, tf, bricks, W, H, dl_drgba, dt=0.5, jitter_mode=0, jitter_value=0.0,
                 threads=0):
    """float64 adjoint of diff_render: (dtf (n,4) f64, [dvox per brick (z,y,x) f64])."""
    tf = np.ascontiguousarray(tf, dtype=np.float32)
    vox = [np.ascontiguousarray(v) for _, v in bricks]
    br = (OrcBrick * len(bricks))(*[_orc_brick(d, v) for (d, _), v in zip(bricks, vox)])
    p = _params(cam, tf, W, H, dt, jitter_mode, jitter_value)
    g = np.ascontiguousarray(dl_drgba, dtype=np.float32)
    dtf = np.zeros(tf.shape, np.float64)
    dv = [np.zeros(v.shape, np.float64) for v in vox]
    ptrs = (C.c_void_p * len(dv))(*[a.ctypes.data for a in dv])
    rc = lib().orc_diff_adjoint(C.byref(p), br, len(bricks), C.c_void_p(g.ctypes.data),
                                C.c_void_p(dtf.ctypes.data), ptrs, int(threads))
    if rc:
        raise ValueError("orc_diff_adjoint failed")
    return dtf, dv


def diff_render64(cam, tf, bricks, W, H, dt=0.5, jitter_mode=0, jitter_value=0.0, threads=0):
    """float64 colour of the differentiable forward on its float control."""
    tf = np.ascontiguousarray(tf, dtype=np.float32)
    vox = [np.ascontiguousarray(v) for _, v in bricks]
    br = (OrcBrick * len(bricks))(*[_orc_brick(d, v) for (d, _), v in zip(bricks, vox)])
    p = _params(cam, tf, W, H, dt, jitter_mode, jitter_value)
    out = np.zeros((H, W, 4), np.float64)
    if lib().orc_diff_render64(C.byref(p), br, len(bricks), C.c_void_p(out.ctypes.data),
                               int(threads)):
        raise ValueError("orc_diff_render64 failed")
    return out


# ----------------------------------------------------------------------------
# surface path
# ----------------------------------------------------------------------------
def _tri_params(cam, W, H, max_depth, spp, island_id, num_islands, crosshair, heat_clock=None,
                heat_scale=1e-5):
    p = OrcParams()
    p.org, p.dir00, p.du, p.dv = (_f3(cam[k]) for k in ("org", "dir_00", "dir_du", "dir_dv"))
    p.dt, p.max_depth, p.W, p.H = 0.5, max_depth, W, H
    p.spp, p.accum_id, p.island_id, p.num_islands = spp, 0, island_id, num_islands
    p.jitter_mode, p.jitter_value, p.crosshair = 0, 0.0, int(crosshair)
    p.tf_n, p.tf = 0, None
    _heat(p, heat_clock, heat_scale)
    return p


def _tri_out(H, W):
    return {"rgba": np.zeros((H, W, 4), np.float32), "rgba8": np.zeros((H, W), np.uint32),
            "depth": np.zeros((H, W), np.float32), "prim": np.full((H, W), -1, np.int32),
            "t": np.zeros((H, W), np.float32)}


def _tri_ptrs(out):
    return [C.c_void_p(out[k].ctypes.data) for k in ("rgba", "rgba8", "depth", "prim", "t")]


def tri_render(cam, tri9: np.ndarray, colors: np.ndarray, W: int, H: int, max_depth=1.0, spp=1,
               island_id=0, num_islands=1, crosshair=False, rows=None, threads=0,
               heat_clock=None, heat_scale=1e-5):
    """One rank's surface image by brute-force closest hit (deviceCode.cpp:
    340-416 with ModelCH :176-188).  tri9: (n,9) float32 vertices in the
    rank's list order (primitive id = row); colors: (n,) colour ids.
    Returns rgba (H,W,4), rgba8, depth, prim (-1 = miss), t."""
    tri9 = np.ascontiguousarray(tri9, dtype=np.float32).reshape(-1, 9)
    colors = np.ascontiguousarray(colors, dtype=np.uint32).reshape(-1)
    assert len(colors) == len(tri9)
    p = _tri_params(cam, W, H, max_depth, spp, island_id, num_islands, crosshair, heat_clock,
                    heat_scale)
    out = _tri_out(H, W)
    row0, row1 = (0, H) if rows is None else rows
    rc = lib().orc_tri_render(C.byref(p), C.c_void_p(tri9.ctypes.data),
                              C.c_void_p(colors.ctypes.data), C.c_longlong(len(tri9)),
                              row0, row1, *_tri_ptrs(out), int(threads))
    if rc != 0:
        raise ValueError(f"orc_tri_render failed ({rc})")
    return out


def tri_render_bvh(cam, nodes: np.ndarray, tris: np.ndarray, root: int, W: int, H: int,
                   max_depth=1.0, spp=1, island_id=0, num_islands=1, crosshair=False, rows=None,
                   threads=0, heat_clock=None, heat_scale=1e-5):
    """The same image through a BVH in the library's device layout (nodes
    (n,32) 4-byte words, 4-wide, tris (m,12) float32), for meshes too large
    for brute force.  prim holds primitive ids."""
    nodes = np.ascontiguousarray(nodes).view(np.uint32).reshape(-1, 32)
    tris = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1, 12)
    p = _tri_params(cam, W, H, max_depth, spp, island_id, num_islands, crosshair, heat_clock,
                    heat_scale)
    out = _tri_out(H, W)
    row0, row1 = (0, H) if rows is None else rows
    f = lib().orc_tri_render_bvh
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_int, C.c_int] + \
        [C.c_void_p] * 5 + [C.c_int]
    rc = f(C.byref(p), C.c_void_p(nodes.ctypes.data if len(nodes) else None),
           C.c_void_p(tris.ctypes.data), int(root), row0, row1, *_tri_ptrs(out), int(threads))
    if rc != 0:
        raise ValueError(f"orc_tri_render_bvh failed ({rc})")
    return out


def random_color(idx: int):
    out = (C.c_float * 3)()
    lib().orc_random_color(C.c_uint32(idx), out)
    return tuple(out)


def mesh_split(verts: np.ndarray, tris: np.ndarray, bounds, n: int):
    """MeshSplitter, median strategy (chopSuey.cpp:60-198), restated in
    numpy: returns (tris reordered cluster by cluster, [(first, last, lo, hi)]).

    Per split: the cluster with most triangles (first of equals) is removed;
    axis = longest side of its domain (y if > x and >= z, z if > x and >= y,
    else x); primitive references (bbox centroids of the original triangles,
    one array for the whole mesh) are sorted over the cluster's index range
    by that coordinate; the plane is the centroid at position first +
    size_t(N / float(n)); triangles with min vertex < plane go left (order
    kept); L then R are appended when non-empty, until n clusters exist
    (at least one split).  A cluster coming back whole with an unchanged
    domain would repeat forever (the reference recurses without end): error.
    Ties in the sort are broken by primitive id."""
    v = np.asarray(verts, np.float32).reshape(-1, 3)
    t = np.array(tris, np.int32).reshape(-1, 3)
    N = len(t)
    tv = v[t]                                   # (N,3,3)
    cen = ((tv.min(axis=1) + tv.max(axis=1)) * np.float32(0.5)).astype(np.float32)
    ref_prim = np.arange(N)
    ref_cen = cen.copy()
    lo0 = np.array(bounds[0], np.float32)
    hi0 = np.array(bounds[1], np.float32)
    clusters = [(0, N, lo0.copy(), hi0.copy())]
    num = int(np.float32(N) / np.float32(n))
    while True:
        sizes = [c[1] - c[0] for c in clusters]
        k = int(np.argmax(sizes))
        first, last, lo, hi = clusters.pop(k)
        s = (hi - lo).astype(np.float32)
        axis = 0
        if s[1] > s[0] and s[1] >= s[2]:
            axis = 1
        elif s[2] > s[0] and s[2] >= s[1]:
            axis = 2
        o = np.lexsort((ref_prim[first:last], ref_cen[first:last, axis]))
        ref_prim[first:last] = ref_prim[first:last][o]
        ref_cen[first:last] = ref_cen[first:last][o]
        plane = ref_cen[min(first + num, N - 1), axis]
        mins = t[first:last]
        mv = v[mins][:, :, axis].min(axis=1)
        left = mv < plane
        t[first:last] = np.concatenate([mins[left], mins[~left]])
        split = first + int(left.sum())
        if split == last:
            split = first
        L = (first, split, lo.copy(), hi.copy())
        L[3][axis] = plane
        R = (split, last, lo.copy(), hi.copy())
        R[2][axis] = plane
        before = len(clusters)
        if L[1] - L[0] > 0:
            clusters.append(L)
        if R[1] - R[0] > 0:
            clusters.append(R)
        if len(clusters) < before + 2 and np.array_equal(clusters[-1][2], lo) \
                and np.array_equal(clusters[-1][3], hi):
            raise ValueError("cluster cannot be split")
        if len(clusters) >= n:
            break
    return t, clusters
