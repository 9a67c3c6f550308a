"""ctypes binding of libaloha_amd.so (the C ABI declared in include/aloha_amd.h).

The library is loaded from this package directory.  A missing library is an
error: there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libaloha_amd.so"

OK = 0
ERR_ARG, ERR_HIP, ERR_ALLOC, ERR_STATE, ERR_IO, ERR_LIMIT = -1, -2, -3, -4, -5, -6
U8, U16, F32 = 1, 2, 4
JITTER_RANDOM, JITTER_FIXED = 0, 1
FLAG_CROSSHAIR, FLAG_HEATMAP, FLAG_HEAT_CLOCK = 1, 2, 4
MAX_RANKS, MAX_RANK_BRICKS, MAX_TF = 64, 256, 4096
NOALLOC_SCRATCH_BYTES = 64 << 20


class Box3i(C.Structure):
    _fields_ = [("lower", C.c_int32 * 3), ("upper", C.c_int32 * 3)]


class Box3f(C.Structure):
    _fields_ = [("lower", C.c_float * 3), ("upper", C.c_float * 3)]


class Range1f(C.Structure):
    _fields_ = [("lower", C.c_float), ("upper", C.c_float)]


class BrickDesc(C.Structure):
    _fields_ = [("cell_range", Box3i), ("voxel_range", Box3i), ("space_range", Box3f),
                ("value_range", Range1f)]

    def voxel_dims(self):
        return tuple(self.voxel_range.upper[a] - self.voxel_range.lower[a] for a in range(3))

    def as_dict(self):
        return {
            "cell": (tuple(self.cell_range.lower), tuple(self.cell_range.upper)),
            "voxel": (tuple(self.voxel_range.lower), tuple(self.voxel_range.upper)),
            "space": (tuple(self.space_range.lower), tuple(self.space_range.upper)),
            "value": (self.value_range.lower, self.value_range.upper),
        }


class RenderParams(C.Structure):
    _fields_ = [("dt", C.c_float), ("max_depth", C.c_float), ("spp", C.c_int32),
                ("accum_id", C.c_int32), ("island_id", C.c_int32), ("num_islands", C.c_int32),
                ("jitter_mode", C.c_int32), ("jitter_value", C.c_float), ("flags", C.c_int32),
                ("heat_map_scale", C.c_float), ("heat_clock", C.c_uint32)]


class Images(C.Structure):
    _fields_ = [("rgba", C.c_void_p), ("rgba8", C.c_void_p), ("depth", C.c_void_p),
                ("accum", C.c_void_p), ("nsamples", C.c_void_p)]


class KDNode(C.Structure):
    _fields_ = [("axis", C.c_int32), ("plane", C.c_float), ("child1", C.c_int32),
                ("child2", C.c_int32)]


assert C.sizeof(BrickDesc) == 80 and C.sizeof(KDNode) == 16

_P = C.c_void_p
_I = C.c_int32
_IP = C.POINTER(C.c_int32)
_FP = C.POINTER(C.c_float)

# (name, restype, argtypes)
_SIGS = [
    ("aloha_last_error", C.c_char_p, [_P]),
    ("aloha_version", C.c_char_p, []),
    ("aloha_debug_alloc_fault", C.c_int, [C.c_int64, C.POINTER(C.c_int64)]),
    ("aloha_split_volume", C.c_int, [_IP, _I, C.POINTER(BrickDesc), _I, _IP]),
    ("aloha_volume_extract", C.c_int, [_P, _IP, _I, C.POINTER(Box3i), _FP, C.POINTER(Range1f)]),
    ("aloha_partitioner_create_kd", C.c_int, [C.POINTER(Box3f), _I, _I, C.POINTER(_P)]),
    ("aloha_partitioner_create_rr", C.c_int, [_I, _I, C.POINTER(_P)]),
    ("aloha_partitioner_destroy", None, [_P]),
    ("aloha_partitioner_rank_bricks", C.c_int, [_P, _I, _IP, _I, _IP]),
    ("aloha_partitioner_composite_order", C.c_int, [_P, _FP, _IP, _I, _IP]),
    ("aloha_partitioner_kd_nodes", C.c_int, [_P, _P, _I, _IP]),
    ("aloha_vol_write", C.c_int, [C.c_char_p, C.POINTER(Box3i), C.POINTER(Box3i), C.POINTER(Box3f), _I,
                                  C.POINTER(BrickDesc), C.POINTER(_FP)]),
    ("aloha_vol_read_header", C.c_int, [C.c_char_p, C.POINTER(Box3i), C.POINTER(Box3i), C.POINTER(Box3f),
                                        C.POINTER(BrickDesc), _I, _IP]),
    ("aloha_vol_read_brick", C.c_int, [C.c_char_p, _I, _FP, C.c_size_t]),
    ("aloha_synth_blobs_u8", C.c_int, [_IP, _I, C.c_uint64, C.c_float, C.c_float, C.c_float, C.c_float, _P]),
    ("aloha_synth_blobs_f32", C.c_int, [_IP, _I, C.c_uint64, C.c_float, C.c_float, C.c_float, C.c_float, _P]),
    ("aloha_synth_fbm_f32", C.c_int, [_IP, _I, C.c_float, C.c_uint64, _P]),
    ("aloha_synth_tf", C.c_int, [_I, _I, C.c_uint64, _P]),
    ("aloha_synth_camera", C.c_int, [C.POINTER(Box3f), _I, _I, C.c_float, C.c_float, C.c_uint64, _FP]),
    ("aloha_ctx_create", C.c_int, [_I, C.POINTER(_P)]),
    ("aloha_ctx_destroy", C.c_int, [_P]),
    ("aloha_set_stream", C.c_int, [_P, _P]),
    ("aloha_sync", C.c_int, [_P]),
    ("aloha_brick_upload", C.c_int, [_P, C.POINTER(BrickDesc), _P, _I, _I, _IP]),
    ("aloha_brick_count", C.c_int, [_P, _IP]),
    ("aloha_brick_bytes", C.c_int, [_P, C.POINTER(C.c_uint64)]),
    ("aloha_set_u8_storage", C.c_int, [_P, _I]),
    ("aloha_brick_dtype", C.c_int, [_P, _I, _IP]),
    ("aloha_tf_set", C.c_int, [_P, _P, _I]),
    ("aloha_camera_set", C.c_int, [_P, _FP, _FP, _FP, _FP]),
    ("aloha_render_ranks", C.c_int, [_P, C.POINTER(RenderParams), _I, _I, _I, _IP, _IP, C.POINTER(Images)]),
    ("aloha_render_composite", C.c_int, [_P, C.POINTER(RenderParams), _I, _I, _I, _IP, _IP, _IP,
                                         _P, _P, _P]),
    ("aloha_render_frame", C.c_int, [_P, C.POINTER(RenderParams), _I, _I, _I, _IP, _IP,
                                     C.POINTER(Images), _IP, _P, _P]),
    ("aloha_tf_set_device", C.c_int, [_P, _P, _I]),
    ("aloha_brick_update", C.c_int, [_P, _I, _P, _I]),
    ("aloha_diff_forward", C.c_int, [_P, C.POINTER(RenderParams), _I, _I, _I, _FP, _I, _IP, _P, _P]),
    ("aloha_diff_backward", C.c_int, [_P, C.POINTER(RenderParams), _I, _I, _I, _FP, _I, _IP, _P, _P,
                                      _P, _P, C.POINTER(C.c_void_p)]),
    ("aloha_batch_create", C.c_int, [_P, C.POINTER(RenderParams), _I, _I, _I, _FP, _I, _IP, _IP, _IP,
                                     _P, _P, _P, C.POINTER(_P)]),
    ("aloha_batch_destroy", C.c_int, [_P]),
    ("aloha_batch_set_schedule", C.c_int, [_P, _I, _I]),
    ("aloha_batch_get_schedule", C.c_int, [_P, _IP, _IP, _IP]),
    ("aloha_batch_set_tile", C.c_int, [_P, _I]),
    ("aloha_batch_get_tile", C.c_int, [_P, _IP]),
    ("aloha_batch_autotune", C.c_int, [_P, _I, _FP, _FP, _I, _IP]),
    ("aloha_batch_capture", C.c_int, [_P]),
    ("aloha_batch_launch", C.c_int, [_P]),
    ("aloha_set_march_schedule", C.c_int, [_P, _I, _I]),
    ("aloha_get_march_schedule", C.c_int, [_P, _IP, _IP]),
    ("aloha_autotune_march", C.c_int, [_P, C.POINTER(RenderParams), _I, _I, _I, _IP, _IP, _IP,
                                       C.POINTER(Images), _P, _P, _I, _FP, _IP]),
    ("aloha_render_footprint", C.c_int, [_P, C.POINTER(RenderParams), _I, _I, _I, _IP, _IP,
                                         C.POINTER(C.c_uint64)]),
    ("aloha_render_footprint_bricks", C.c_int, [_P, C.POINTER(RenderParams), _I, _I, _I, _IP,
                                                _IP, C.POINTER(C.c_uint64), _I]),
    ("aloha_composite_over", C.c_int, [_P, _P, _I, _IP, _I, _I, _P, _P]),
    ("aloha_composite_over_sparse", C.c_int, [_P, _P, _I, _IP, _IP, _I, _I, _P, _P]),
    ("aloha_rank_rects", C.c_int, [_P, C.POINTER(RenderParams), _I, _I, _I, _IP, _IP, _IP]),
    ("aloha_render_frame_noalloc", C.c_int, [_P, C.POINTER(RenderParams), _I, _I, _I, _IP, _IP, _IP, _P, _P]),
    ("aloha_project_boxes", C.c_int, [_FP, C.POINTER(Box3f), _I, _I, _I, _IP]),
    ("aloha_composite_over_rgba8", C.c_int, [_P, _P, _I, _IP, _I, _I, _P]),
    ("aloha_composite_zmin", C.c_int, [_P, _P, _P, _I, _I, _I, _P, _P]),
    ("aloha_mesh_create", C.c_int, [_P, C.c_int64, _P, C.c_int64, C.POINTER(_P)]),
    ("aloha_mesh_destroy", None, [_P]),
    ("aloha_mesh_info", C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(Box3f), _IP]),
    ("aloha_mesh_get", C.c_int, [_P, _P, _P]),
    ("aloha_mesh_clusters", C.c_int, [_P, _P, _P, _I, _IP]),
    ("aloha_mesh_synth_spheres", C.c_int, [_I, _I, _I, C.POINTER(Box3f), C.c_float, C.c_float, C.c_float,
                                           C.c_uint64, C.POINTER(_P)]),
    ("aloha_mesh_split", C.c_int, [_P, _I, _IP]),
    ("aloha_mesh_read_obj", C.c_int, [C.c_char_p, C.POINTER(_P)]),
    ("aloha_tri_write", C.c_int, [C.c_char_p, _P]),
    ("aloha_tri_read", C.c_int, [C.c_char_p, C.POINTER(_P)]),
    ("aloha_tris_upload", C.c_int, [_P, _P, _I, _IP, _IP, _P, _IP]),
    ("aloha_tris_build_host", C.c_int, [_P, _I, _IP, _P, C.POINTER(C.c_int64), C.POINTER(C.c_int64), _IP,
                                        _P, _P, _IP]),
    ("aloha_tris_info", C.c_int, [_P, _I, C.POINTER(C.c_int64), C.POINTER(C.c_int64), _IP,
                                  C.POINTER(C.c_uint64)]),
    ("aloha_tris_export", C.c_int, [_P, _I, _P, _P, _IP]),
    ("aloha_render_tris", C.c_int, [_P, C.POINTER(RenderParams), _I, _I, _I, _IP, C.POINTER(Images), _P]),
    ("aloha_render_tris_composite", C.c_int, [_P, C.POINTER(RenderParams), _I, _I, _I, _IP, _P, _P, _P]),
]

# every symbol include/aloha_amd.h declares
DECLARED = [s[0] for s in _SIGS]


class AlohaError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def load(path: Path = LIB_PATH) -> C.CDLL:
    if not path.exists():
        raise ImportError(
            f"{path} is not built; run `python -m aloha_amd.build` (there is no CPU fallback)")
    # One HIP runtime per process: torch ships its own libamdhip64 (NEEDED as
    # "libamdhip64.so", soname libamdhip64.so.7).  Loaded first, it satisfies
    # this library's libamdhip64.so.7 dependency; loaded after, a second HIP
    # and HSA runtime would come up next to ours and fail to find the device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(str(path))
    for name, res, args in _SIGS:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        _lib = load()
    return _lib


def check(rc: int, ctx=None) -> None:
    if rc != OK:
        msg = lib().aloha_last_error(ctx)
        raise AlohaError(rc, msg.decode() if msg else "")
