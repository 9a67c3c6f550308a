"""The C-ABI library: loads, exports every symbol include/aloha_amd.h
declares, validates arguments, and its host-side seeded inputs are
deterministic.  No GPU calls (a context on a machine without a GPU must fail
with a status code, not crash)."""
import re
from pathlib import Path

import numpy as np
import pytest

import aloha_amd as A
from aloha_amd import _lib

ROOT = Path(__file__).resolve().parent.parent


def _declared(header: Path):
    text = header.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(aloha_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported_and_bound():
    lib = _lib.lib()
    for header in sorted((ROOT / "include").glob("*.h")):
        for name in _declared(header):
            assert hasattr(lib, name), f"{header.name}: {name} not exported"
    assert sorted(_lib.DECLARED) == _declared(ROOT / "include" / "aloha_amd.h")
    assert b"gfx950" in lib.aloha_version()


def test_context_without_gpu_is_an_error_not_a_crash():
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is visible")
    except ImportError:
        pass
    with pytest.raises(A.AlohaError) as e:
        A.Context(0)
    assert e.value.code == _lib.ERR_HIP


def test_host_argument_validation():
    with pytest.raises(A.AlohaError):
        A.split_volume((0, 4, 4), 2)
    with pytest.raises(A.AlohaError):
        A.volume_extract(np.zeros((4, 4, 4), np.uint8), 3, A.Box3i((0, 0, 0), (2, 2, 2)))
    with pytest.raises(A.AlohaError):
        A.volume_extract(np.zeros((4, 4, 4), np.uint8), 1, A.Box3i((0, 0, 0), (5, 2, 2)))
    with pytest.raises(A.AlohaError):
        A.Partitioner.kd([], 0)
    p = A.Partitioner.round_robin(4, 2)
    with pytest.raises(A.AlohaError):
        p.rank_bricks(2)
    with pytest.raises(A.AlohaError):
        A.synth_tf(0)


def test_synthetic_inputs_are_seeded():
    a = A.synth_blobs_u8((20, 16, 12), 5, 7)
    assert a.shape == (12, 16, 20) and a.max() == 255
    assert np.array_equal(a, A.synth_blobs_u8((20, 16, 12), 5, 7))
    assert not np.array_equal(a, A.synth_blobs_u8((20, 16, 12), 5, 8))
    f = A.synth_fbm_f32((24, 20, 16), 3, 8.0, 11)
    assert f.min() == 0.0 and f.max() == 1.0
    assert np.array_equal(f, A.synth_fbm_f32((24, 20, 16), 3, 8.0, 11))
    tf = A.synth_tf(256, 6, 3)
    assert tf.shape == (256, 4) and tf.min() >= 0 and tf.max() <= 1
    # piecewise linear: at most 5 slope changes per channel (6 control points)
    slope = np.diff(tf[:, 0].astype(np.float64))
    assert (np.abs(np.diff(slope)) > 1e-5).sum() <= 2 * 5


def test_camera_convention():
    b = A.Box3f((0, 0, 0), (128, 64, 32))
    cam = A.synth_camera(b, 300, 200, 45.0, 2.5, 42)
    c = np.array([64, 32, 16], np.float64)
    diag = np.linalg.norm([128, 64, 32])
    org = cam["org"].astype(np.float64)
    assert abs(np.linalg.norm(org - c) - 2.5 * diag) < 1e-3 * diag
    # the image centre ray points at the centre of the bounds
    mid = cam["dir_00"] + 150 * cam["dir_du"] + 100 * cam["dir_dv"]
    to_c = (c - org) / np.linalg.norm(c - org)
    assert np.dot(mid / np.linalg.norm(mid), to_c) > 1 - 1e-6
    # vertical field of view 45 degrees, square pixels
    top = cam["dir_00"] + 150 * cam["dir_du"] + 200 * cam["dir_dv"]
    bottom = cam["dir_00"] + 150 * cam["dir_du"]
    ang = np.arccos(np.dot(top, bottom) / np.linalg.norm(top) / np.linalg.norm(bottom))
    assert abs(np.degrees(ang) - 45.0) < 1e-3
    assert abs(np.linalg.norm(cam["dir_du"]) - np.linalg.norm(cam["dir_dv"])) < 1e-7


def test_max_depth_formula():
    # maxDepth = length(org + center) + length(center + upper) (sic)
    b = A.Box3f((0, 0, 0), (2, 2, 2))
    md = A.max_depth(b, [3.0, 0.0, 0.0])
    assert abs(md - (np.linalg.norm([4, 1, 1]) + np.linalg.norm([3, 3, 3]))) < 1e-5


def test_device_buffer_checks_on_host():
    """The Python layer passes None through and refuses host tensors before
    any call reaches the C ABI (the kernels write through raw pointers)."""
    import torch
    assert A._dev(None, 1 << 20, "out") is None
    with pytest.raises(ValueError, match="GPU tensor"):
        A._dev(torch.zeros(16), 4, "out")
    with pytest.raises(ValueError, match="GPU tensor"):
        A.RankImages(rgba=torch.zeros((1, 2, 2, 4))).as_c(1, 2, 2)
    assert A.RankImages().as_c(4, 8, 8).rgba is None
