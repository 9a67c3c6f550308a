"""torch.autograd wrapper of the differentiable form (aloha_diff_forward /
aloha_diff_backward in the C ABI): images of a density grid through a
transfer function, with gradients with respect to both.

Used to check the hand-written adjoint against finite differences and to
drive optimisation; all compute runs in the HIP kernels.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import Box3f, BrickDesc, Context, F32, render_params
from ._lib import lib

__all__ = ["DifferentiableVolume"]


def _box_brick(shape_zyx) -> BrickDesc:
    """One brick covering a (z,y,x) grid: cells = voxels, space [0, n]."""
    nz, ny, nx = shape_zyx
    d = BrickDesc()
    for a, n in enumerate((nx, ny, nz)):
        d.cell_range.lower[a] = d.voxel_range.lower[a] = 0
        d.cell_range.upper[a] = d.voxel_range.upper[a] = n
        d.space_range.lower[a] = 0.0
        d.space_range.upper[a] = float(n)
    d.value_range.lower, d.value_range.upper = 0.0, 1.0
    return d


class _DiffRender(torch.autograd.Function):
    @staticmethod
    def forward(ctx, tf, density, vol):
        rgba, ns = vol._forward(tf, density)
        ctx.save_for_backward(tf, density, rgba, ns)
        ctx.vol = vol
        return rgba

    @staticmethod
    def backward(ctx, grad):
        tf, density, rgba, ns = ctx.saved_tensors
        vol = ctx.vol
        vol._set_inputs(tf, density)  # the adjoint re-marches these inputs
        dtf = torch.zeros(tf.shape, dtype=torch.float32, device=rgba.device)
        dd = torch.zeros(density.shape, dtype=torch.float32, device=rgba.device)
        vol._backward(rgba, ns, grad.contiguous(), dtf, dd)
        return dtf, dd, None


class DifferentiableVolume:
    """A density grid (one brick) rendered from V cameras at W x H.

    render(tf, density) -> (V, H, W, 4) premultiplied RGBA, differentiable in
    tf (n, 4) and density (z, y, x); both float32 tensors on the GPU."""

    def __init__(self, shape_zyx, cams, W: int, H: int, dt: float = 0.5,
                 jitter_value: float | None = 0.5):
        self.ctx = Context(0, stream=torch.cuda.current_stream())
        self.shape = tuple(int(v) for v in shape_zyx)
        self.desc = _box_brick(shape_zyx)
        self.brick = self.ctx.upload_brick(self.desc, np.zeros(shape_zyx, np.float32))
        self.V = len(cams)
        self.cams = np.concatenate([np.concatenate([c["org"], c["dir_00"], c["dir_du"],
                                                    c["dir_dv"]]) for c in cams]).astype(np.float32)
        self.W, self.H = W, H
        self.params = (render_params(dt=dt, jitter_mode=1, jitter_value=jitter_value)
                       if jitter_value is not None else render_params(dt=dt))
        self._ids = (C.c_int32 * 1)(self.brick)
        self._cams = self.cams.ctypes.data_as(C.POINTER(C.c_float))

    @property
    def bounds(self) -> Box3f:
        return self.desc.space_range

    def render(self, tf: torch.Tensor, density: torch.Tensor) -> torch.Tensor:
        return _DiffRender.apply(tf, density, self)

    def forward_only(self, tf, density):
        with torch.no_grad():
            return self._forward(tf, density)

    def _set_inputs(self, tf, density):
        # the TF and the grid are read through raw device pointers
        if not (tf.is_cuda and density.is_cuda):
            raise ValueError("tf and density must be GPU tensors")
        if tf.dim() != 2 or tf.shape[1] != 4:
            raise ValueError(f"tf: expected (n, 4), got {tuple(tf.shape)}")
        if tuple(density.shape) != self.shape:
            raise ValueError(f"density: expected {self.shape}, got {tuple(density.shape)}")
        # the context already holds these exact tensors (the backward re-marches
        # the forward's inputs): no new copy.  A tensor changed in place since
        # (an optimizer step) has a new version counter; the last inputs are
        # kept referenced, so no other tensor can take their addresses.
        key = (tf.data_ptr(), tf._version, tuple(tf.shape), density.data_ptr(),
               density._version)
        if key == getattr(self, "_last_key", None):
            return
        tf_c = tf.detach().contiguous().float()
        den_c = density.detach().contiguous().float()
        self.ctx._check(lib().aloha_tf_set_device(self.ctx._h, C.c_void_p(tf_c.data_ptr()),
                                                  tf_c.shape[0]))
        self.ctx._check(lib().aloha_brick_update(self.ctx._h, self.brick,
                                                 C.c_void_p(den_c.data_ptr()), F32))
        self._last_tf = tf_c  # keep alive until the stream has consumed them
        self._last_density = den_c
        self._last_key = key
        self._last_src = (tf, density)

    def _forward(self, tf, density):
        self._set_inputs(tf, density)
        rgba = torch.empty((self.V, self.H, self.W, 4), dtype=torch.float32, device="cuda")
        ns = torch.empty((self.V, self.H, self.W), dtype=torch.int32, device="cuda")
        self.ctx._check(lib().aloha_diff_forward(
            self.ctx._h, C.byref(self.params), self.W, self.H, self.V, self._cams, 1, self._ids,
            C.c_void_p(rgba.data_ptr()), C.c_void_p(ns.data_ptr())))
        return rgba, ns

    def _backward(self, rgba, ns, grad, dtf, dd):
        ptrs = (C.c_void_p * 1)(dd.data_ptr())
        self.ctx._check(lib().aloha_diff_backward(
            self.ctx._h, C.byref(self.params), self.W, self.H, self.V, self._cams, 1, self._ids,
            C.c_void_p(rgba.data_ptr()), C.c_void_p(ns.data_ptr()), C.c_void_p(grad.data_ptr()),
            C.c_void_p(dtf.data_ptr()), ptrs))
