"""Seeded synthetic volume workloads in the reference's data layout.

Each workload is a raw volume bricked by chopSuey's VolumeSplitter, bricks
assigned to virtual ranks by partitionKD, a piecewise-linear TF and an orbit
camera -- everything Renderer's constructor and event handlers would hold
(waikikiRenderer.cpp:344-846, :853-1013) after loading a .vol file.

  blobs128   128^3 u8 Gaussian-blob volume, 2 bricks, 512^2 (the CPU-mode case)
  fbm512     512^3 f32 fBm volume, 8 bricks (2x2x2), 1024^2 (the GPU headline)
  surface    4096 displaced spheres (9.75 M triangles), 8 clusters / ranks, 1920x1080
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import (Box3f, BrickDesc, Partitioner, max_depth, render_params, split_volume,
               synth_blobs_f32, synth_blobs_u8, synth_camera, synth_fbm_f32, synth_tf)


@dataclass
class VolumeWorkload:
    name: str
    dims: tuple
    raw: np.ndarray                      # (z, y, x) u8 or f32
    bricks: list                         # BrickDesc per brick (chopSuey order)
    voxels: list                         # raw-dtype (z,y,x) array per brick
    tf: np.ndarray                       # (n, 4) float32
    cam: dict                            # org, dir_00, dir_du, dir_dv
    W: int
    H: int
    dt: float
    num_ranks: int
    rank_bricks: list = field(default_factory=list)   # per rank: brick ids
    order: list = field(default_factory=list)         # composite order (rank ids)
    max_depth: float = 1.0

    @property
    def bounds(self) -> Box3f:
        # the .vol header spaceRange = [0, dims] is Renderer::modelBounds
        b = Box3f()
        for a in range(3):
            b.lower[a] = 0.0
            b.upper[a] = float(self.dims[a])
        return b

    def params(self, **kw):
        base = dict(dt=self.dt, max_depth=self.max_depth)
        base.update(kw)
        return render_params(**base)

    def brick_voxels_total(self) -> int:
        return int(sum(v.size for v in self.voxels))


def brick_subvolumes(raw: np.ndarray, bricks) -> list:
    out = []
    for d in bricks:
        lo, hi = d.voxel_range.lower, d.voxel_range.upper
        out.append(np.ascontiguousarray(raw[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]]))
    return out


def make_volume_workload(name: str, raw: np.ndarray, num_bricks: int, num_ranks: int, W: int,
                         H: int, dt: float = 0.5, tf_seed: int = 0x5EED2,
                         cam_seed: int = 0x5EED3, tf_n: int = 256) -> VolumeWorkload:
    dims = tuple(raw.shape[::-1])
    bricks = split_volume(dims, num_bricks)
    voxels = brick_subvolumes(raw, bricks)
    wl = VolumeWorkload(name=name, dims=dims, raw=raw, bricks=bricks, voxels=voxels,
                        tf=synth_tf(tf_n, 6, tf_seed), cam={}, W=W, H=H, dt=dt,
                        num_ranks=num_ranks)
    wl.cam = synth_camera(wl.bounds, W, H, 45.0, 2.5, cam_seed)
    part = Partitioner.kd([d.space_range for d in bricks], num_ranks)
    wl.rank_bricks = [part.rank_bricks(r) for r in range(num_ranks)]
    wl.order = part.composite_order(wl.cam["org"])
    wl.max_depth = max_depth(wl.bounds, wl.cam["org"])
    return wl


def blobs128(num_ranks: int = 2, W: int = 512, H: int = 512, seed: int = 0x5EED0,
             size: int = 128) -> VolumeWorkload:
    raw = synth_blobs_u8((size, size, size), 16, seed, 4.0, 16.0, 0.3, 1.0)
    return make_volume_workload("blobs128", raw, 2, num_ranks, W, H)


def fbm512(num_ranks: int = 8, W: int = 1024, H: int = 1024, seed: int = 0x5EED1,
           size: int = 512) -> VolumeWorkload:
    raw = synth_fbm_f32((size, size, size), 5, size / 8.0, seed)
    return make_volume_workload("fbm512", raw, 8, num_ranks, W, H)


@dataclass
class DiffWorkload:
    """Differentiable-form workload: a density grid (blob field), a TF,
    V orbit views, and target images made from a perturbed TF (+-20% per
    entry) and perturbed density (+5% seeded noise)."""
    density: np.ndarray   # (z, y, x) f32
    tf: np.ndarray        # (n, 4) f32
    cams: list
    W: int
    H: int
    target_tf: np.ndarray
    target_density: np.ndarray
    dt: float = 0.5


def diff256(size: int = 256, views: int = 8, W: int = 512, H: int = 512, seed: int = 0x5EED4,
            tf_alpha_scale: float = 0.25) -> DiffWorkload:
    dims = (size, size, size)
    density = synth_blobs_f32(dims, 16, seed, 4.0 * size / 128, 16.0 * size / 128, 0.3, 1.0)
    tf = synth_tf(256, 6, seed + 1)
    # a semi-transparent TF keeps gradients flowing through many samples
    tf[:, 3] *= tf_alpha_scale
    b = Box3f()
    for a in range(3):
        b.lower[a], b.upper[a] = 0.0, float(size)
    cams = [synth_camera(b, W, H, 45.0, 2.5, seed + 10 + v) for v in range(views)]
    rng = np.random.default_rng(seed + 2)
    target_tf = np.clip(tf * (1.0 + 0.2 * rng.uniform(-1, 1, tf.shape)), 0, 1).astype(np.float32)
    target_density = (density * (1.0 + 0.05 * rng.standard_normal(density.shape))).astype(np.float32)
    return DiffWorkload(density, tf, cams, W, H, target_tf, target_density)


@dataclass
class BatchWorkload:
    """Uneven bricks on a lattice of cells (one brick per cell, at the cell's
    lower corner) seen from V views; one virtual rank per brick, kd-assigned
    over the cell boxes so middle planes never cut a brick."""
    bricks: list
    voxels: list
    tf: np.ndarray
    cams: list
    orders: list
    rank_bricks: list
    W: int
    H: int
    dt: float = 0.5

    def params(self, **kw):
        base = dict(dt=self.dt, max_depth=1.0)
        base.update(kw)
        return render_params(**base)


def batched(n_side: int = 4, cell: int = 256, smin: int = 32, views: int = 16, W: int = 768,
            H: int = 768, seed: int = 0x5EED5) -> BatchWorkload:
    rng = np.random.default_rng(seed)
    bricks, voxels, cells = [], [], []
    for i in range(n_side ** 3):
        c = (i % n_side, (i // n_side) % n_side, i // (n_side * n_side))
        dims = tuple(int(v) for v in rng.integers(smin, cell + 1, 3))
        lo = tuple(c[a] * cell for a in range(3))
        d = BrickDesc()
        for a in range(3):
            d.cell_range.lower[a] = d.voxel_range.lower[a] = lo[a]
            d.cell_range.upper[a] = d.voxel_range.upper[a] = lo[a] + dims[a]
            d.space_range.lower[a] = float(lo[a])
            d.space_range.upper[a] = float(lo[a] + dims[a])
        if rng.random() < 0.5:
            v = synth_blobs_u8(dims, 16, seed + 100 + i, 4.0, 16.0, 0.3, 1.0)
        else:
            v = synth_blobs_f32(dims, 16, seed + 100 + i, 4.0, 16.0, 0.3, 1.0)
        d.value_range.lower, d.value_range.upper = 0.0, 1.0
        bricks.append(d)
        voxels.append(v)
        cb = Box3f()
        for a in range(3):
            cb.lower[a], cb.upper[a] = float(lo[a]), float(lo[a] + cell)
        cells.append(cb)
    P = len(bricks)
    part = Partitioner.kd(cells, P)
    rank_bricks = [part.rank_bricks(r) for r in range(P)]
    bounds = Box3f()
    for a in range(3):
        bounds.lower[a], bounds.upper[a] = 0.0, float(n_side * cell)
    cams = [synth_camera(bounds, W, H, 45.0, 2.5, seed + 10 + v) for v in range(views)]
    orders = [part.composite_order(c["org"]) for c in cams]
    return BatchWorkload(bricks, voxels, synth_tf(256, 6, seed + 1), cams, orders, rank_bricks,
                         W, H)


@dataclass
class SurfaceWorkload:
    """Triangle-mesh scene as Renderer holds it after loading a .tri file in
    TriMesh mode (waikikiRenderer.cpp:390-593): clusters from MeshSplitter,
    blocks of clusters per rank (partitionRoundRobin), colour ids per
    rank-local geometry, camera and maxDepth."""
    name: str
    mesh: object                          # aloha_amd.Mesh (split)
    ranks: list                           # cluster ids per rank
    colors: list                          # colour id per listed cluster
    cam: dict
    W: int
    H: int
    max_depth: float
    ntris: int

    def params(self, **kw):
        return render_params(max_depth=self.max_depth, **kw)


def surface(num_spheres: int = 4096, rings: int = 35, segments: int = 35, num_clusters: int = 8,
            num_ranks: int = 8, W: int = 1920, H: int = 1080, seed: int = 0x5EED6,
            dist_factor: float = 1.2, name: str = "spheres", rmin: float = 0.005,
            rmax: float = 0.03) -> SurfaceWorkload:
    """Seeded displaced spheres in the unit cube (radius 0.005-0.03, radial
    noise 10 % of the radius), split into clusters, one block per rank.
    Defaults: 4096 spheres x 2380 triangles = 9.75 M triangles, 8 clusters,
    8 ranks, 1920x1080, camera 1.2 diagonals from the centre."""
    from . import Mesh, mesh_color_ids
    dom = Box3f((0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    mesh = Mesh.spheres(num_spheres, rings, segments, dom, rmin, rmax, 0.1, seed=seed)
    n = mesh.split(num_clusters)
    rr = Partitioner.round_robin(n, num_ranks)
    ranks = [rr.rank_bricks(r) for r in range(num_ranks)]
    info = mesh.info()
    b = info["bounds"]
    cam = synth_camera(b, W, H, 45.0, dist_factor, seed=seed + 1)
    return SurfaceWorkload(name, mesh, ranks, mesh_color_ids(ranks, n), cam, W, H,
                           max_depth(b, cam["org"]), info["ntris"])
