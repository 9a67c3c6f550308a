"""aloha_project_boxes: the pixel rectangle outside of which the per-rank
launch writes the miss result without marching.  Checked against the
oracle's full march (every pixel) -- no pixel outside the rectangle may reach
the brick -- including a narrow field of view where the 1e-5 direction clamp
of generateRay (deviceCode.cpp:103-105) moves rays by several pixels."""
import math

import numpy as np
import pytest

import aloha_amd as A
import oracle as O


def _axis_camera(W, H, fov_deg, org=(0.0, 0.0, -100.0)):
    h = math.tan(math.radians(fov_deg) / 2)
    return {"org": np.array(org, np.float32),
            "dir_00": np.array([-h, -h, 1.0], np.float32),
            "dir_du": np.array([2 * h / W, 0, 0], np.float32),
            "dir_dv": np.array([0, 2 * h / H, 0], np.float32)}


def _reached(cam, desc, W, H):
    vox_lo, vox_hi = desc["voxel"]
    dims = [vox_hi[a] - vox_lo[a] for a in range(3)]
    vox = np.full(dims[::-1], 0.5, np.float32)
    tf = np.tile(np.array([[0.8, 0.4, 0.2, 0.3]], np.float32), (16, 1))
    o = O.render_rank(cam, tf, [(desc, vox)], W, H, dt=0.5, max_depth=1000.0)
    return (o["depth"] < 1) | (o["nsamples"] > 0) | (o["rgba"][..., 3] > 0)


def _check(cam, desc, W, H):
    x0, y0, x1, y1 = A.project_boxes(cam, [desc["space"]], W, H)
    hit = _reached(cam, desc, W, H)
    ys, xs = np.nonzero(hit)
    assert len(xs) > 0
    assert xs.min() >= x0 and xs.max() < x1 and ys.min() >= y0 and ys.max() < y1
    return (x0, y0, x1, y1), xs, ys


def test_narrow_fov_near_axis_clamp_stays_inside():
    # fov 0.5 deg at 2048 px: a pixel spans 4.3e-6 rad, while the clamp sends
    # every ray with |d.x| < 1e-5 along d.x = +1e-5 -- up to 4.7 pixels away.
    # The brick starts at x = 0.95e-3, just left of where those rays cross
    # z in [0, 10], so clamped rays of pixels left of its outline reach it.
    W = H = 2048
    cam = _axis_camera(W, H, 0.5)
    desc = {"cell": ((0, -5, 0), (6, 6, 11)), "voxel": ((0, -5, 0), (6, 6, 11)),
            "space": ((0.00095, -5.0, 0.0), (5.0, 5.0, 10.0))}
    (x0, _, _, _), xs, _ = _check(cam, desc, W, H)
    # the outline's left edge (the far face's x / z as a pixel coordinate)
    h = math.tan(math.radians(0.25))
    left = (0.00095 / 110.0 + h) / (2 * h / W) - 0.5
    # rays from well left of the outline reach the brick: a 2-pixel margin
    # around the outline would have cut them off
    assert xs.min() < left - 3
    assert x0 <= xs.min()


@pytest.mark.parametrize("fov", [45.0, 5.0, 0.5])
def test_oblique_views_match_full_march(fov):
    W, H = 256, 192
    rng = np.random.default_rng(int(fov * 10))
    desc = {"cell": ((2, 3, 1), (9, 8, 7)), "voxel": ((2, 3, 1), (9, 8, 7)),
            "space": ((2.0, 3.0, 1.0), (9.0, 8.0, 7.0))}
    centre = np.array([5.5, 5.5, 4.0])
    for _ in range(3):
        d = rng.standard_normal(3)
        d /= np.linalg.norm(d)
        org = centre - 40 * d
        fwd = d
        up = np.cross(fwd, [0.3, 1.0, 0.1])
        up /= np.linalg.norm(up)
        right = np.cross(up, fwd)
        h = math.tan(math.radians(fov) / 2)
        # aim slightly off the centre so the brick sits near the image middle
        cam = {"org": org.astype(np.float32),
               "dir_00": (fwd - h * right - h * H / W * up).astype(np.float32),
               "dir_du": (2 * h * right / W).astype(np.float32),
               "dir_dv": (2 * h * H / W * up / H).astype(np.float32)}
        _check(cam, desc, W, H)


def test_whole_image_when_a_brick_reaches_the_camera_plane():
    cam = _axis_camera(64, 64, 45.0, org=(0.5, 0.5, 0.5))
    assert A.project_boxes(cam, [((0, 0, 0), (1, 1, 1))], 64, 64) == (0, 0, 64, 64)
    assert A.project_boxes(cam, [], 64, 64) == (0, 0, 0, 0)
