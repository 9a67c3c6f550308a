"""aloha_amd -- MI355X-native per-brick render-and-composite path of the Aloha
sort-last renderer (vtvamr/aloha).

The compute lives in libaloha_amd.so (hand-written HIP for gfx950 behind the C
ABI of include/aloha_amd.h).  This module is a thin Python view of that ABI
for tests, the benchmark and the autograd wrapper: host numpy arrays in,
device torch tensors for images.  It has no compute path of its own.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Sequence

import numpy as np

from ._lib import (ERR_ALLOC, ERR_ARG, ERR_HIP, ERR_IO, ERR_LIMIT, ERR_STATE,  # noqa: F401
                   NOALLOC_SCRATCH_BYTES)
from ._lib import (F32, FLAG_CROSSHAIR, FLAG_HEATMAP, FLAG_HEAT_CLOCK, JITTER_FIXED, JITTER_RANDOM, MAX_RANKS,
                   MAX_RANK_BRICKS, MAX_TF, U8, U16, AlohaError, Box3f, Box3i, BrickDesc, Images,
                   KDNode, Range1f, RenderParams, check, lib)

__all__ = [
    "AlohaError", "BrickDesc", "Box3i", "Box3f", "RenderParams", "Context", "Partitioner",
    "split_volume", "volume_extract", "synth_blobs_u8", "synth_blobs_f32", "synth_fbm_f32",
    "synth_tf", "synth_camera", "max_depth", "U8", "U16", "F32", "JITTER_RANDOM",
    "JITTER_FIXED", "FLAG_CROSSHAIR", "FLAG_HEATMAP", "FLAG_HEAT_CLOCK", "MAX_RANKS", "MAX_RANK_BRICKS", "MAX_TF",
    "vol_write", "vol_read", "Mesh", "mesh_color_ids",
]


def _i3(v) -> C.Array:
    return (C.c_int32 * 3)(*[int(x) for x in v])


def _f3(v) -> C.Array:
    return (C.c_float * 3)(*[float(x) for x in v])


# ----------------------------------------------------------------------------
# host-side bricking and partitioning
# ----------------------------------------------------------------------------
def split_volume(dims: Sequence[int], num_bricks: int) -> list[BrickDesc]:
    """chopSuey VolumeSplitter (chopSuey.cpp:245-323)."""
    cap = max(4 * num_bricks + 8, 16)
    out = (BrickDesc * cap)()
    n = C.c_int32()
    check(lib().aloha_split_volume(_i3(dims), int(num_bricks), out, cap, C.byref(n)))
    return [out[i] for i in range(n.value)]


def volume_extract(raw: np.ndarray, bpc: int, voxel_range: Box3i):
    """Volume::get (volume.h:46-98): raw is (z, y, x) u8/u16/f32; returns
    (float32 array (z,y,x) of the sub-box, (lo, hi))."""
    raw = np.ascontiguousarray(raw)
    dims = raw.shape[::-1]
    ext = [voxel_range.upper[a] - voxel_range.lower[a] for a in range(3)]
    out = np.empty(ext[::-1], dtype=np.float32)
    vr = Range1f(np.finfo(np.float32).max, -np.finfo(np.float32).max)
    check(lib().aloha_volume_extract(raw.ctypes.data, _i3(dims), int(bpc), C.byref(voxel_range),
                                     out.ctypes.data_as(C.POINTER(C.c_float)), C.byref(vr)))
    return out, (vr.lower, vr.upper)


class Partitioner:
    """Partitioner (waikikiRenderer.cpp:126-338)."""

    def __init__(self, handle: C.c_void_p, num_ranks: int):
        self._h = handle
        self._free = lib().aloha_partitioner_destroy  # held: module globals are gone at exit
        self.num_ranks = num_ranks

    @classmethod
    def kd(cls, domains: Sequence[Box3f], num_ranks: int) -> "Partitioner":
        arr = (Box3f * max(1, len(domains)))(*domains)
        h = C.c_void_p()
        check(lib().aloha_partitioner_create_kd(arr, len(domains), int(num_ranks), C.byref(h)))
        return cls(h, num_ranks)

    @classmethod
    def round_robin(cls, n: int, num_ranks: int) -> "Partitioner":
        h = C.c_void_p()
        check(lib().aloha_partitioner_create_rr(int(n), int(num_ranks), C.byref(h)))
        return cls(h, num_ranks)

    def __del__(self):
        if getattr(self, "_h", None):
            self._free(self._h)
            self._h = None

    def rank_bricks(self, rank: int) -> list[int]:
        cnt = C.c_int32()
        cap = 4096
        ids = (C.c_int32 * cap)()
        check(lib().aloha_partitioner_rank_bricks(self._h, int(rank), ids, cap, C.byref(cnt)))
        return list(ids[: cnt.value])

    def composite_order(self, reference: Sequence[float]) -> list[int]:
        cnt = C.c_int32()
        cap = 4096
        out = (C.c_int32 * cap)()
        check(lib().aloha_partitioner_composite_order(self._h, _f3(reference), out, cap,
                                                      C.byref(cnt)))
        return list(out[: cnt.value])

    def kd_nodes(self) -> list[tuple[int, float, int, int]]:
        cnt = C.c_int32()
        cap = 8192
        out = (KDNode * cap)()
        check(lib().aloha_partitioner_kd_nodes(self._h, out, cap, C.byref(cnt)))
        return [(n.axis, n.plane, n.child1, n.child2) for n in out[: cnt.value]]


# ----------------------------------------------------------------------------
# .vol files
# ----------------------------------------------------------------------------
def vol_write(path: str, cell_range: Box3i, voxel_range: Box3i, space_range: Box3f,
              bricks: Sequence[BrickDesc], voxels: Sequence[np.ndarray]) -> None:
    arrs = [np.ascontiguousarray(v, dtype=np.float32) for v in voxels]
    ptrs = (C.POINTER(C.c_float) * max(1, len(arrs)))(
        *[a.ctypes.data_as(C.POINTER(C.c_float)) for a in arrs])
    barr = (BrickDesc * max(1, len(bricks)))(*bricks)
    check(lib().aloha_vol_write(path.encode(), C.byref(cell_range), C.byref(voxel_range),
                                C.byref(space_range), len(bricks), barr, ptrs))


def vol_read(path: str):
    """Returns (cell_range, voxel_range, space_range, [BrickDesc], [voxels (z,y,x)])."""
    cr, vr, sr = Box3i(), Box3i(), Box3f()
    n = C.c_int32()
    cap = 65536
    descs = (BrickDesc * cap)()
    check(lib().aloha_vol_read_header(path.encode(), C.byref(cr), C.byref(vr), C.byref(sr), descs,
                                      cap, C.byref(n)))
    bricks, vox = [], []
    for i in range(n.value):
        d = descs[i]
        dims = d.voxel_dims()
        a = np.empty(dims[::-1], dtype=np.float32)
        check(lib().aloha_vol_read_brick(path.encode(), i,
                                          a.ctypes.data_as(C.POINTER(C.c_float)), a.size))
        bricks.append(d)
        vox.append(a)
    return cr, vr, sr, bricks, vox


# ----------------------------------------------------------------------------
# synthetic inputs
# ----------------------------------------------------------------------------
def synth_blobs_u8(dims, num_blobs=16, seed=0x5EED0, smin=4.0, smax=16.0, amin=0.3, amax=1.0):
    out = np.empty(tuple(dims)[::-1], dtype=np.uint8)
    check(lib().aloha_synth_blobs_u8(_i3(dims), num_blobs, seed, smin, smax, amin, amax,
                                     out.ctypes.data))
    return out


def synth_blobs_f32(dims, num_blobs=16, seed=0x5EED0, smin=4.0, smax=16.0, amin=0.3, amax=1.0):
    out = np.empty(tuple(dims)[::-1], dtype=np.float32)
    check(lib().aloha_synth_blobs_f32(_i3(dims), num_blobs, seed, smin, smax, amin, amax,
                                      out.ctypes.data))
    return out


def synth_fbm_f32(dims, octaves=5, base_period=64.0, seed=0x5EED1):
    out = np.empty(tuple(dims)[::-1], dtype=np.float32)
    check(lib().aloha_synth_fbm_f32(_i3(dims), octaves, base_period, seed, out.ctypes.data))
    return out


def synth_tf(n=256, num_points=6, seed=0x5EED2) -> np.ndarray:
    out = np.empty((n, 4), dtype=np.float32)
    check(lib().aloha_synth_tf(n, num_points, seed, out.ctypes.data))
    return out


def synth_camera(bounds: Box3f, W: int, H: int, fovy_deg=45.0, dist_factor=2.5, seed=0x5EED3):
    out = (C.c_float * 12)()
    check(lib().aloha_synth_camera(C.byref(bounds), W, H, fovy_deg, dist_factor, seed, out))
    v = np.frombuffer(out, dtype=np.float32).copy()
    return {"org": v[0:3], "dir_00": v[3:6], "dir_du": v[6:9], "dir_dv": v[9:12]}


def project_boxes(cam: dict, boxes, W: int, H: int):
    """Pixel rectangle (x0, y0, x1, y1) outside of which no primary ray of cam
    reaches any of boxes (pairs of lower/upper corners); host only."""
    arr = (Box3f * max(1, len(boxes)))()
    for k, (lo, hi) in enumerate(boxes):
        arr[k].lower[:] = [float(v) for v in lo]
        arr[k].upper[:] = [float(v) for v in hi]
    c = np.concatenate([np.asarray(cam[k], np.float32) for k in ("org", "dir_00", "dir_du",
                                                                  "dir_dv")])
    out = (C.c_int32 * 4)()
    check(lib().aloha_project_boxes(c.ctypes.data_as(C.POINTER(C.c_float)), arr, len(boxes), W,
                                    H, out))
    return tuple(out)


def max_depth(bounds: Box3f, org) -> float:
    """handleCameraEvent's maxDepth (waikikiRenderer.cpp:858-860), in float32:
    length(org + center) + length(center + upper)."""
    f = np.float32
    lo = np.array(bounds.lower, dtype=f)
    hi = np.array(bounds.upper, dtype=f)
    center = (lo + hi) * f(0.5)

    def length(v):
        return np.sqrt(f(v[0] * v[0] + v[1] * v[1]) + f(v[2] * v[2]), dtype=f)

    radius = length((center + hi).astype(f))
    return float(f(length((np.asarray(org, dtype=f) + center).astype(f)) + radius))


# ----------------------------------------------------------------------------
# device context
# ----------------------------------------------------------------------------
def _dev(t, nbytes: int, name: str):
    """Device pointer of a contiguous GPU tensor of at least nbytes (None
    passes through).  The kernels write through raw pointers, so a host,
    strided or short buffer is refused here rather than faulting the GPU."""
    if t is None:
        return None
    if not getattr(t, "is_cuda", False):
        raise ValueError(f"{name}: expected a GPU tensor")
    if not t.is_contiguous():
        raise ValueError(f"{name}: expected a contiguous tensor")
    have = t.numel() * t.element_size()
    if have < nbytes:
        raise ValueError(f"{name}: {have} bytes, the call writes or reads {nbytes}")
    return C.c_void_p(t.data_ptr())


@dataclass
class RankImages:
    """Device tensors of one render launch (P rank slots); any may be None."""
    rgba: object = None
    rgba8: object = None
    depth: object = None
    accum: object = None
    nsamples: object = None

    def as_c(self, P: int, W: int, H: int) -> Images:
        """The C struct for a launch of P rank slots at W x H (sizes checked)."""
        n = max(0, P) * max(0, W) * max(0, H)
        return Images(_dev(self.rgba, 16 * n, "rgba"), _dev(self.rgba8, 4 * n, "rgba8"),
                      _dev(self.depth, 4 * n, "depth"), _dev(self.accum, 16 * n, "accum"),
                      _dev(self.nsamples, 4 * n, "nsamples"))

    @classmethod
    def alloc(cls, P: int, W: int, H: int, device="cuda", rgba=True, rgba8=True, depth=True,
              accum=False, nsamples=False):
        import torch
        dev = torch.device(device)
        return cls(
            rgba=torch.empty((P, H, W, 4), dtype=torch.float32, device=dev) if rgba else None,
            rgba8=torch.empty((P, H, W), dtype=torch.int32, device=dev) if rgba8 else None,
            depth=torch.empty((P, H, W), dtype=torch.float32, device=dev) if depth else None,
            accum=torch.zeros((P, H, W, 4), dtype=torch.float32, device=dev) if accum else None,
            nsamples=torch.empty((P, H, W), dtype=torch.int32, device=dev) if nsamples else None,
        )


def debug_alloc_fault(after: int = -1) -> int:
    """Test hook: make the library's allocation `after` places further on fail
    as out of memory, once (after < 0: disarm).  Returns the allocations made
    so far."""
    made = C.c_int64()
    check(lib().aloha_debug_alloc_fault(int(after), C.byref(made)))
    return int(made.value)


def render_params(dt=0.5, max_depth=1.0, spp=1, accum_id=0, island_id=0, num_islands=1,
                  jitter_mode=JITTER_RANDOM, jitter_value=0.0, flags=0, heat_map_scale=1e-5,
                  heat_clock=0):
    return RenderParams(dt, max_depth, spp, accum_id, island_id, num_islands, jitter_mode,
                        jitter_value, flags, heat_map_scale, heat_clock)


def _rank_table(ranks: Sequence[Sequence[int]]):
    offs = [0]
    ids: list[int] = []
    for r in ranks:
        ids.extend(int(b) for b in r)
        offs.append(len(ids))
    return (C.c_int32 * len(offs))(*offs), (C.c_int32 * max(1, len(ids)))(*ids)


def _order(order: Sequence[int], P: int) -> C.Array:
    """Visibility order of P rank slots as a C array; the C side reads P
    entries, so a shorter or longer sequence is refused here."""
    if len(order) != P:
        raise ValueError(f"order: {len(order)} entries for {P} rank slots")
    return (C.c_int32 * max(1, P))(*[int(v) for v in order])


class Context:
    """Device context (bricks in HBM, TF, camera) on one GPU.

    stream: a torch.cuda.Stream or raw hipStream_t for all of the context's
    work.  Left out, the context launches on torch's current stream of the
    device at construction (so torch's allocations, fills and copies of the
    tensors it reads and writes are ordered with its kernels); without torch,
    on a non-blocking stream of its own."""

    def __init__(self, device: int = 0, stream=None):
        h = C.c_void_p()
        check(lib().aloha_ctx_create(int(device), C.byref(h)))
        self._h = h
        self._free = lib().aloha_ctx_destroy  # held: module globals are gone at exit
        self.device = device
        if stream is None:
            import sys
            torch = sys.modules.get("torch")
            if torch is not None and torch.cuda.is_available():
                stream = torch.cuda.current_stream(device)
        if stream is not None:
            self.set_stream(stream)

    def close(self):
        if getattr(self, "_h", None):
            self._free(self._h)
            self._h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        check(rc, self._h)

    def set_stream(self, stream) -> None:
        """stream: a torch.cuda.Stream or a raw hipStream_t int (0: the null stream)."""
        handle = int(getattr(stream, "cuda_stream", stream) or 0)
        self._check(lib().aloha_set_stream(self._h, C.c_void_p(handle) if handle else None))

    def sync(self) -> None:
        self._check(lib().aloha_sync(self._h))

    def upload_brick(self, desc: BrickDesc, voxels, dtype: int | None = None) -> int:
        """voxels: numpy (z,y,x) u8/u16/f32 on the host, or a cuda torch tensor,
        holding at least the brick's voxel-range count of elements (checked:
        the copy reads through a raw pointer)."""
        bid = C.c_int32()
        lo, hi = desc.voxel_range.lower, desc.voxel_range.upper
        count = 1
        for a in range(3):
            count *= max(0, hi[a] - lo[a])

        def need(code, nbytes):
            want = count * {U8: 1, U16: 2, F32: 4}.get(code, 4)
            if nbytes < want:
                raise ValueError(f"voxels: {nbytes} bytes, the brick's voxel range needs {want}")

        if hasattr(voxels, "data_ptr") and getattr(voxels, "is_cuda", False):
            import torch
            t = voxels.contiguous()
            code = {torch.uint8: U8, torch.int16: U16, torch.float32: F32}[t.dtype]
            need(dtype or code, t.numel() * t.element_size())
            self._check(lib().aloha_brick_upload(self._h, C.byref(desc), C.c_void_p(t.data_ptr()),
                                                 dtype or code, 1, C.byref(bid)))
            return bid.value
        a = np.ascontiguousarray(voxels)
        code = {np.dtype(np.uint8): U8, np.dtype(np.uint16): U16,
                np.dtype(np.float32): F32}[a.dtype]
        need(dtype or code, a.nbytes)
        self._check(lib().aloha_brick_upload(self._h, C.byref(desc), a.ctypes.data, dtype or code,
                                             0, C.byref(bid)))
        return bid.value

    def brick_bytes(self) -> int:
        v = C.c_uint64()
        self._check(lib().aloha_brick_bytes(self._h, C.byref(v)))
        return v.value

    def set_u8_storage(self, keep_u8: bool) -> None:
        """Storage of u8 bricks uploaded from now on: bytes (True; decoded
        at sample time to the same floats, a quarter of the memory) or
        float32 (False, the default)."""
        self._check(lib().aloha_set_u8_storage(self._h, 1 if keep_u8 else 0))

    def brick_update(self, brick_id: int, voxels) -> None:
        """Replace an uploaded brick's voxels (same extent) from a cuda
        tensor (uint8, int16 as u16 bits, or float32), asynchronously on the
        context's stream; a brick kept as u8 takes uint8 only."""
        import torch
        if not (hasattr(voxels, "data_ptr") and getattr(voxels, "is_cuda", False)):
            raise ValueError("voxels: a cuda tensor")
        t = voxels.contiguous()
        code = {torch.uint8: U8, torch.int16: U16, torch.float32: F32}[t.dtype]
        n = C.c_int32()
        self._check(lib().aloha_brick_count(self._h, C.byref(n)))
        if not 0 <= brick_id < n.value:
            raise ValueError(f"brick_id {brick_id} not uploaded")
        self._check(lib().aloha_brick_update(self._h, brick_id, C.c_void_p(t.data_ptr()), code))
        if t.data_ptr() != voxels.data_ptr():
            # a contiguous copy made here: the context's stream reads it
            # before it goes out of scope
            self._check(lib().aloha_sync(self._h))

    def brick_dtype(self, brick_id: int) -> int:
        """How the brick is stored in HBM: U8 or F32."""
        v = C.c_int32()
        self._check(lib().aloha_brick_dtype(self._h, brick_id, C.byref(v)))
        return v.value

    def set_tf(self, tf: np.ndarray) -> None:
        tf = np.ascontiguousarray(tf, dtype=np.float32)
        if tf.ndim != 2 or tf.shape[1] != 4:
            raise ValueError(f"tf: expected (n, 4) RGBA entries, got {tf.shape}")
        self._check(lib().aloha_tf_set(self._h, tf.ctypes.data, tf.shape[0]))

    def set_camera(self, org, dir_00, dir_du, dir_dv) -> None:
        self._check(lib().aloha_camera_set(self._h, _f3(org), _f3(dir_00), _f3(dir_du),
                                           _f3(dir_dv)))

    def render_ranks(self, params: RenderParams, W: int, H: int,
                     ranks: Sequence[Sequence[int]], out: RankImages) -> None:
        offs, ids = _rank_table(ranks)
        imgs = out.as_c(len(ranks), W, H)
        self._check(lib().aloha_render_ranks(self._h, C.byref(params), W, H, len(ranks), offs, ids,
                                             C.byref(imgs)))

    def render_frame(self, params: RenderParams, W: int, H: int, ranks: Sequence[Sequence[int]],
                     images: RankImages | None, order: Sequence[int], out_rgba=None,
                     out_rgba8=None) -> None:
        """One frame: render_ranks into `images` (float RGBA required) and the
        ordered 'over' of their RGBA into the outputs, as render_ranks followed
        by composite_over with rank_rects' rectangles (bit for bit), in two
        launches: the per-rank launch also clears the outputs outside the
        rectangles' bounding box, the composite covers the box.  images None:
        the frame only, the rank images kept compact (only inside their
        rectangles) in the context's workspace."""
        offs, ids = _rank_table(ranks)
        o = _order(order, len(ranks))
        n = max(0, W) * max(0, H)
        imgs = C.byref(images.as_c(len(ranks), W, H)) if images is not None else None
        self._check(lib().aloha_render_frame(self._h, C.byref(params), W, H, len(ranks), offs, ids,
                                             imgs, o, _dev(out_rgba, 16 * n, "out_rgba"),
                                             _dev(out_rgba8, 4 * n, "out_rgba8")))

    def render_frame_noalloc(self, params: RenderParams, W: int, H: int,
                             ranks: Sequence[Sequence[int]], order: Sequence[int],
                             out_rgba=None, out_rgba8=None) -> None:
        """render_frame's outputs without caller-held rank images: the ranks'
        images go to a workspace the library holds on the GPU (compact, only
        each rank's rectangle).  Allocates nothing on the device."""
        offs, ids = _rank_table(ranks)
        o = _order(order, len(ranks))
        n = max(0, W) * max(0, H)
        self._check(lib().aloha_render_frame_noalloc(
            self._h, C.byref(params), W, H, len(ranks), offs, ids, o,
            _dev(out_rgba, 16 * n, "out_rgba"), _dev(out_rgba8, 4 * n, "out_rgba8")))

    def render_composite(self, params: RenderParams, W: int, H: int,
                         ranks: Sequence[Sequence[int]], order: Sequence[int], out_rgba=None,
                         out_rgba8=None, nsamples=None) -> None:
        """Fused raygen + ordered 'over' composite (one launch)."""
        offs, ids = _rank_table(ranks)
        o = _order(order, len(ranks))
        n = max(0, W) * max(0, H)
        self._check(lib().aloha_render_composite(self._h, C.byref(params), W, H, len(ranks), offs,
                                                 ids, o, _dev(out_rgba, 16 * n, "out_rgba"),
                                                 _dev(out_rgba8, 4 * n, "out_rgba8"),
                                                 _dev(nsamples, 4 * n, "nsamples")))

    def set_march_schedule(self, ranks_lanes: int, fused_lanes: int) -> None:
        self._check(lib().aloha_set_march_schedule(self._h, ranks_lanes, fused_lanes))

    def march_schedule(self) -> tuple[int, int]:
        r, f = C.c_int32(), C.c_int32()
        self._check(lib().aloha_get_march_schedule(self._h, C.byref(r), C.byref(f)))
        return r.value, f.value

    def autotune_march(self, params: RenderParams, W: int, H: int, ranks: Sequence[Sequence[int]],
                       out: RankImages, order: Sequence[int] | None = None, out_rgba=None,
                       out_rgba8=None, reps: int = 5):
        """Pick the fastest lanes-per-ray schedule for these launches; returns
        the chosen (ranks_lanes, fused_lanes) and the timing table."""
        offs, ids = _rank_table(ranks)
        o = _order(order, len(ranks)) if order is not None else None
        imgs = out.as_c(len(ranks), W, H)
        n = max(0, W) * max(0, H)
        table = (C.c_float * 12)()
        rows = C.c_int32()
        self._check(lib().aloha_autotune_march(self._h, C.byref(params), W, H, len(ranks), offs, ids,
                                               o, C.byref(imgs), _dev(out_rgba, 16 * n, "out_rgba"),
                                               _dev(out_rgba8, 4 * n, "out_rgba8"), reps,
                                               table, C.byref(rows)))
        t = [(int(table[3 * i]), float(table[3 * i + 1]), float(table[3 * i + 2]))
             for i in range(rows.value)]
        return self.march_schedule(), t

    def footprint(self, params: RenderParams, W: int, H: int,
                  ranks: Sequence[Sequence[int]]) -> int:
        offs, ids = _rank_table(ranks)
        v = C.c_uint64()
        self._check(lib().aloha_render_footprint(self._h, C.byref(params), W, H, len(ranks), offs,
                                                 ids, C.byref(v)))
        return v.value

    def footprint_bricks(self, params: RenderParams, W: int, H: int,
                         ranks: Sequence[Sequence[int]]) -> np.ndarray:
        """Distinct voxels read per brick id (all uploaded bricks)."""
        offs, ids = _rank_table(ranks)
        n = C.c_int32()
        self._check(lib().aloha_brick_count(self._h, C.byref(n)))
        out = (C.c_uint64 * max(1, n.value))()
        self._check(lib().aloha_render_footprint_bricks(self._h, C.byref(params), W, H, len(ranks),
                                                        offs, ids, out, n.value))
        return np.frombuffer(out, dtype=np.uint64)[:n.value].copy()

    def composite_over(self, rank_rgba, order: Sequence[int], W: int, H: int, out_rgba=None,
                       out_rgba8=None, rects=None) -> None:
        """Ordered 'over' of the rank images.  rects (P x 4 x0, y0, x1, y1, e.g.
        from rank_rects): the sparse form, reading each rank only inside its
        rectangle (the images must be transparent outside it)."""
        o = (C.c_int32 * len(order))(*order)
        n = max(0, W) * max(0, H)
        outs = (_dev(out_rgba, 16 * n, "out_rgba"), _dev(out_rgba8, 4 * n, "out_rgba8"))
        src = _dev(rank_rgba, 16 * n * len(order), "rank_rgba")
        if rects is None:
            self._check(lib().aloha_composite_over(self._h, src, len(order), o, W, H, *outs))
        else:
            flat = [int(v) for r in rects for v in r]
            if len(flat) != 4 * len(order):
                raise ValueError("rects: need 4 integers per rank")
            rc = (C.c_int32 * len(flat))(*flat)
            self._check(lib().aloha_composite_over_sparse(self._h, src, len(order), o, rc, W, H,
                                                          *outs))

    def rank_rects(self, params: RenderParams, W: int, H: int,
                   ranks: Sequence[Sequence[int]]) -> list[tuple[int, int, int, int]]:
        """The pixel rectangles outside of which render_ranks with these
        arguments writes transparent black (x0, y0, x1, y1 per rank)."""
        offs, ids = _rank_table(ranks)
        out = (C.c_int32 * (4 * len(ranks)))()
        self._check(lib().aloha_rank_rects(self._h, C.byref(params), W, H, len(ranks), offs, ids,
                                           out))
        return [tuple(out[4 * s:4 * s + 4]) for s in range(len(ranks))]

    def composite_over_rgba8(self, rank_rgba8, order: Sequence[int], W: int, H: int,
                             out_rgba8) -> None:
        o = (C.c_int32 * len(order))(*order)
        n = max(0, W) * max(0, H)
        self._check(lib().aloha_composite_over_rgba8(
            self._h, _dev(rank_rgba8, 4 * n * len(order), "rank_rgba8"), len(order), o, W, H,
            _dev(out_rgba8, 4 * n, "out_rgba8")))

    def composite_zmin(self, rank_rgba8, rank_depth, P: int, W: int, H: int, out_rgba8=None,
                       out_depth=None) -> None:
        n = max(0, W) * max(0, H)
        self._check(lib().aloha_composite_zmin(
            self._h, _dev(rank_rgba8, 4 * n * max(0, P), "rank_rgba8"),
            _dev(rank_depth, 4 * n * max(0, P), "rank_depth"), P, W, H,
            _dev(out_rgba8, 4 * n, "out_rgba8"), _dev(out_depth, 4 * n, "out_depth")))


    # ---- surface path ---------------------------------------------------
    def upload_tris(self, mesh: "Mesh", ranks: Sequence[Sequence[int]],
                    colors: Sequence[Sequence[int]] | None = None) -> list[int]:
        """Upload each rank's clusters (flattened, with a BVH); colors: per
        rank, the colour id of each listed cluster (default mesh_color_ids)."""
        P = len(ranks)
        if colors is None:
            colors = mesh_color_ids(ranks, mesh.num_clusters)
        offs, ids = _rank_table(ranks)
        col = np.ascontiguousarray(np.concatenate([np.asarray(c, dtype=np.uint32) for c in colors])
                                   if sum(len(c) for c in colors) else np.zeros(1, np.uint32))
        out = (C.c_int32 * P)()
        self._check(lib().aloha_tris_upload(self._h, mesh._h, P, offs, ids, col.ctypes.data, out))
        return list(out)

    def tris_info(self, geom: int) -> dict:
        nt, nn, dp, nb = C.c_int64(), C.c_int64(), C.c_int32(), C.c_uint64()
        self._check(lib().aloha_tris_info(self._h, geom, C.byref(nt), C.byref(nn), C.byref(dp),
                                          C.byref(nb)))
        return {"ntris": nt.value, "nnodes": nn.value, "depth": dp.value, "bytes": nb.value}

    def tris_export(self, geom: int):
        """(nodes (n,32) raw 4-byte words as uint32 (4-wide BVH nodes), tris
        (ntris,12) float32, root)."""
        info = self.tris_info(geom)
        nodes = np.zeros((max(info["nnodes"], 1), 32), dtype=np.uint32)
        tris = np.zeros((max(info["ntris"], 1), 12), dtype=np.float32)
        root = C.c_int32()
        self._check(lib().aloha_tris_export(self._h, geom, nodes.ctypes.data, tris.ctypes.data,
                                            C.byref(root)))
        return nodes[:info["nnodes"]], tris[:info["ntris"]], root.value

    def render_tris(self, params: RenderParams, W: int, H: int, geoms: Sequence[int],
                    out: RankImages, stats=None) -> None:
        g = (C.c_int32 * len(geoms))(*geoms)
        imgs = out.as_c(len(geoms), W, H)
        self._check(lib().aloha_render_tris(self._h, C.byref(params), W, H, len(geoms), g,
                                            C.byref(imgs),
                                            _dev(stats, 8 * len(geoms) * max(0, W) * max(0, H),
                                                 "stats")))

    def render_tris_composite(self, params: RenderParams, W: int, H: int, geoms: Sequence[int],
                              out_rgba8=None, out_depth=None, stats=None) -> None:
        g = (C.c_int32 * len(geoms))(*geoms)
        n = max(0, W) * max(0, H)
        self._check(lib().aloha_render_tris_composite(self._h, C.byref(params), W, H, len(geoms), g,
                                                      _dev(out_rgba8, 4 * n, "out_rgba8"),
                                                      _dev(out_depth, 4 * n, "out_depth"),
                                                      _dev(stats, 8 * n, "stats")))


# ----------------------------------------------------------------------------
# triangle meshes (host)
# ----------------------------------------------------------------------------
def mesh_color_ids(ranks: Sequence[Sequence[int]], num_clusters: int) -> list[list[int]]:
    """ClusterID each rank gives its i-th geometry for randomColor
    (waikikiRenderer.cpp:554-557): rank * (numClusters / numRanks) + i, with
    integer division (the increment at :588 changes a dead loop-local copy)."""
    per = num_clusters // len(ranks)
    return [[r * per + i for i in range(len(cl))] for r, cl in enumerate(ranks)]


class Mesh:
    """Host triangle mesh (aloha_mesh): vertices, triangles, clusters."""

    def __init__(self, handle: C.c_void_p):
        self._h = handle
        self._free = lib().aloha_mesh_destroy  # held: module globals are gone at exit

    @classmethod
    def from_arrays(cls, verts: np.ndarray, tris: np.ndarray) -> "Mesh":
        v = np.ascontiguousarray(verts, dtype=np.float32).reshape(-1, 3)
        t = np.ascontiguousarray(tris, dtype=np.int32).reshape(-1, 3)
        h = C.c_void_p()
        check(lib().aloha_mesh_create(v.ctypes.data, len(v), t.ctypes.data, len(t), C.byref(h)))
        return cls(h)

    @classmethod
    def spheres(cls, num_spheres: int, rings: int, segments: int, domain: Box3f,
                rmin_frac=0.005, rmax_frac=0.03, disp_frac=0.1, seed=0x5EED4) -> "Mesh":
        h = C.c_void_p()
        check(lib().aloha_mesh_synth_spheres(num_spheres, rings, segments, C.byref(domain),
                                             rmin_frac, rmax_frac, disp_frac, seed, C.byref(h)))
        return cls(h)

    @classmethod
    def read_obj(cls, path: str) -> "Mesh":
        h = C.c_void_p()
        check(lib().aloha_mesh_read_obj(str(path).encode(), C.byref(h)))
        return cls(h)

    @classmethod
    def read_tri(cls, path: str) -> "Mesh":
        h = C.c_void_p()
        check(lib().aloha_tri_read(str(path).encode(), C.byref(h)))
        return cls(h)

    def write_tri(self, path: str) -> None:
        check(lib().aloha_tri_write(str(path).encode(), self._h))

    def close(self):
        if getattr(self, "_h", None):
            self._free(self._h)
            self._h = None

    __del__ = close

    def info(self) -> dict:
        nv, nt, nc = C.c_int64(), C.c_int64(), C.c_int32()
        b = Box3f()
        check(lib().aloha_mesh_info(self._h, C.byref(nv), C.byref(nt), C.byref(b), C.byref(nc)))
        return {"nverts": nv.value, "ntris": nt.value, "bounds": b, "nclusters": nc.value}

    @property
    def num_clusters(self) -> int:
        return max(1, self.info()["nclusters"])

    def arrays(self):
        i = self.info()
        v = np.empty((i["nverts"], 3), dtype=np.float32)
        t = np.empty((i["ntris"], 3), dtype=np.int32)
        check(lib().aloha_mesh_get(self._h, v.ctypes.data, t.ctypes.data))
        return v, t

    def split(self, num_clusters: int) -> int:
        n = C.c_int32()
        check(lib().aloha_mesh_split(self._h, num_clusters, C.byref(n)))
        return n.value

    def clusters(self):
        """[(first, last, Box3f)] triangle ranges and split domains."""
        n = C.c_int32()
        cap = max(1, self.info()["nclusters"])
        r = np.zeros((cap, 2), dtype=np.int64)
        b = (Box3f * cap)()
        check(lib().aloha_mesh_clusters(self._h, r.ctypes.data, b, cap, C.byref(n)))
        return [(int(r[i, 0]), int(r[i, 1]), b[i]) for i in range(n.value)]

    def build_bvh(self, cluster_ids: Sequence[int], colors: Sequence[int] | None = None):
        """Host-only build of one rank's device layout (aloha_tris_build_host):
        (nodes (n,32) uint32 words of 4-wide nodes, tris (m,12) float32, root,
        depth of the 4-wide tree)."""
        ids = (C.c_int32 * max(1, len(cluster_ids)))(*cluster_ids)
        col = np.ascontiguousarray(colors if colors is not None else np.zeros(len(cluster_ids)),
                                   dtype=np.uint32)
        if col.size == 0:
            col = np.zeros(1, np.uint32)
        nt, nn, dp, root = C.c_int64(), C.c_int64(), C.c_int32(), C.c_int32()
        check(lib().aloha_tris_build_host(self._h, len(cluster_ids), ids, col.ctypes.data,
                                          C.byref(nt), C.byref(nn), C.byref(dp), None, None,
                                          C.byref(root)))
        nodes = np.zeros((max(nn.value, 1), 32), np.uint32)
        tris = np.zeros((max(nt.value, 1), 12), np.float32)
        check(lib().aloha_tris_build_host(self._h, len(cluster_ids), ids, col.ctypes.data,
                                          C.byref(nt), C.byref(nn), C.byref(dp), nodes.ctypes.data,
                                          tris.ctypes.data, C.byref(root)))
        return nodes[:nn.value], tris[:nt.value], root.value, dp.value

    def rank_tri9(self, cluster_ids: Sequence[int]) -> np.ndarray:
        """A rank's flattened triangles (ntris, 9) in list order, as the
        upload flattens them (primitive id = row)."""
        v, t = self.arrays()
        cl = self.clusters() or [(0, len(t), None)]
        rows = [t[cl[c][0]:cl[c][1]] for c in cluster_ids]
        idx = np.concatenate(rows) if rows else np.zeros((0, 3), np.int32)
        return v[idx].reshape(-1, 9)


class Batch:
    """Many views of one rank table in one launch (aloha_batch_*).  Output
    tensors are (V, H, W, 4) float32 / (V, H, W) int32; each view's image
    equals Context.render_composite for that view's camera and order."""

    def __init__(self, ctx: Context, params: RenderParams, W: int, H: int, cams,
                 ranks: Sequence[Sequence[int]], orders, out_rgba=None, out_rgba8=None,
                 nsamples=None):
        self.ctx = ctx
        V = len(cams)
        cam = np.concatenate([np.concatenate([c["org"], c["dir_00"], c["dir_du"], c["dir_dv"]])
                              for c in cams]).astype(np.float32)
        offs, ids = _rank_table(ranks)
        ords = np.ascontiguousarray(np.asarray(orders, dtype=np.int32).reshape(V, len(ranks)))
        n = V * max(0, W) * max(0, H)
        h = C.c_void_p()
        ctx._check(lib().aloha_batch_create(ctx._h, C.byref(params), W, H, V,
                                            cam.ctypes.data_as(C.POINTER(C.c_float)), len(ranks),
                                            offs, ids, ords.ctypes.data_as(C.POINTER(C.c_int32)),
                                            _dev(out_rgba, 16 * n, "out_rgba"),
                                            _dev(out_rgba8, 4 * n, "out_rgba8"),
                                            _dev(nsamples, 4 * n, "nsamples"), C.byref(h)))
        self._h = h
        self._free = lib().aloha_batch_destroy  # held: module globals are gone at exit
        self._keep = (out_rgba, out_rgba8, nsamples)

    def close(self):
        if getattr(self, "_h", None):
            self._free(self._h)
            self._h = None

    __del__ = close

    def set_schedule(self, lanes_per_ray: int, persistent_blocks: int) -> None:
        self.ctx._check(lib().aloha_batch_set_schedule(self._h, lanes_per_ray, persistent_blocks))

    def set_tile(self, tile_w: int) -> None:
        """Tiles of 64 pixels: 8 (8x8), 16 (16x4) or 32 (32x2) wide."""
        self.ctx._check(lib().aloha_batch_set_tile(self._h, tile_w))

    def schedule(self):
        t, b, c, w = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        self.ctx._check(lib().aloha_batch_get_schedule(self._h, C.byref(t), C.byref(b), C.byref(c)))
        self.ctx._check(lib().aloha_batch_get_tile(self._h, C.byref(w)))
        return {"lanes_per_ray": t.value, "persistent_blocks": b.value, "tile_w": w.value,
                "captured": bool(c.value)}

    def autotune(self, reps: int = 3):
        """Times every schedule; returns (best ms, [(tile_w, lanes, blocks, ms), ...])."""
        best = C.c_float()
        cap = 64
        table = (C.c_float * (4 * cap))()
        n = C.c_int32()
        self.ctx._check(lib().aloha_batch_autotune(self._h, reps, C.byref(best), table, cap,
                                                   C.byref(n)))
        rows = [(int(table[4 * i]), int(table[4 * i + 1]), int(table[4 * i + 2]),
                 float(table[4 * i + 3])) for i in range(min(n.value, cap))]
        return best.value, rows

    def capture(self) -> None:
        self.ctx._check(lib().aloha_batch_capture(self._h))

    def launch(self) -> None:
        self.ctx._check(lib().aloha_batch_launch(self._h))
