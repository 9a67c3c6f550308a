"""Pins the oracle's float64 adjoint of the differentiable forward
(orc_diff_adjoint) by central finite differences of its float64 forward
(orc_diff_render64): every TF entry and every voxel of a small brick, loss
sum(w * C) with random weights w.  Transparent enough that no ray reaches
the alpha cut-off, so the float control (sample set) does not move."""
import numpy as np

import oracle as O


def _setup(seed=0):
    rng = np.random.default_rng(seed)
    vox = rng.random((5, 6, 7), dtype=np.float32)
    b = {"cell": ((0, 0, 0), (7, 6, 5)), "voxel": ((0, 0, 0), (7, 6, 5)),
         "space": ((0.0, 0.0, 0.0), (7.0, 6.0, 5.0))}
    tf = rng.random((16, 4), dtype=np.float32)
    tf[:, 3] *= 0.1
    cam = {"org": np.array([3.5, 3.0, -8.0], np.float32),
           "dir_00": np.array([-0.35, -0.3, 1.0], np.float32),
           "dir_du": np.array([0.1, 0.0, 0.0], np.float32),
           "dir_dv": np.array([0.0, 0.1, 0.0], np.float32)}
    W, H = 7, 6
    w = rng.standard_normal((H, W, 4)).astype(np.float32)
    return b, vox, tf, cam, W, H, w


def test_adjoint_matches_finite_differences():
    b, vox, tf, cam, W, H, w = _setup()
    kw = dict(dt=0.5, jitter_mode=1, jitter_value=0.3)
    rgba, ns = O.diff_render(cam, tf, [(b, vox)], W, H, **kw)
    assert ns.sum() > 50 and rgba[..., 3].max() < 0.9
    dtf, (dvox,) = O.diff_adjoint(cam, tf, [(b, vox)], W, H, w, **kw)
    wd = w.astype(np.float64)

    def loss_tf(t):
        return float((O.diff_render64(cam, t, [(b, vox)], W, H, **kw) * wd).sum())

    # the forward takes float32 TF entries: step them in float64 through a
    # float32-exact perturbation of a power of two
    eps = 2.0 ** -12
    worst = 0.0
    for i in range(tf.shape[0]):
        for c in range(4):
            tp, tm = tf.copy(), tf.copy()
            tp[i, c] += eps
            tm[i, c] -= eps
            fd = (loss_tf(tp) - loss_tf(tm)) / (float(tp[i, c]) - float(tm[i, c]))
            worst = max(worst, abs(fd - dtf[i, c]) / (1e-3 + abs(fd)))
    assert worst < 2e-3, worst

    def loss_v(v):
        return float((O.diff_render64(cam, tf, [(b, v)], W, H, **kw) * wd).sum())

    rng = np.random.default_rng(1)
    idx = [tuple(rng.integers(0, s) for s in vox.shape) for _ in range(60)]
    worst = 0.0
    checked = 0
    for ix in idx:
        vp, vm = vox.copy(), vox.copy()
        vp[ix] += eps
        vm[ix] -= eps
        fd = (loss_v(vp) - loss_v(vm)) / (float(vp[ix]) - float(vm[ix]))
        if abs(fd) + abs(dvox[ix]) == 0:
            continue
        checked += 1
        worst = max(worst, abs(fd - dvox[ix]) / (1e-3 + abs(fd)))
    assert checked > 30 and worst < 2e-2, worst


def test_diff_forward_close_to_plain_forward():
    # the closed-form lattice moves samples by at most an ulp of t
    b, vox, tf, cam, W, H, _ = _setup(2)
    rgba, ns = O.diff_render(cam, tf, [(b, vox)], W, H, dt=0.5, jitter_mode=1, jitter_value=0.3)
    ref = O.render_rank(cam, tf, [(b, vox)], W, H, dt=0.5, jitter_mode=1, jitter_value=0.3)
    same = ns == ref["nsamples"]
    assert same.mean() > 0.9
    assert np.abs(rgba[same] - ref["rgba"][same]).max() < 1e-5
