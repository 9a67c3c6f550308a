"""GPU checks of the differentiable form (aloha_diff_forward / _backward):
the forward against the oracle's float restatement, the hand-written adjoint
against the oracle's float64 adjoint and against central finite differences
through the torch.autograd wrapper, and a short Adam fit."""
import numpy as np
import pytest

import oracle as O
from aloha_amd import workloads as WL

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small():
    return WL.diff256(size=48, views=2, W=80, H=72)


def _vol(wl):
    from aloha_amd.autograd import DifferentiableVolume
    return DifferentiableVolume(wl.density.shape, wl.cams, wl.W, wl.H, dt=wl.dt, jitter_value=0.5)


def _brick(wl):
    n = wl.density.shape[::-1]
    return {"cell": ((0, 0, 0), n), "voxel": ((0, 0, 0), n),
            "space": ((0.0, 0.0, 0.0), tuple(float(v) for v in n))}


def test_forward_matches_oracle(small):
    import torch
    print("\n[case] diff forward 48^3, 2 views 80x72", flush=True)
    vol = _vol(small)
    tf = torch.from_numpy(small.tf).cuda()
    den = torch.from_numpy(small.density).cuda()
    rgba, ns = vol.forward_only(tf, den)
    torch.cuda.synchronize()
    rgba, ns = rgba.cpu().numpy(), ns.cpu().numpy()
    for v, cam in enumerate(small.cams):
        ref, rns = O.diff_render(cam, small.tf, [(_brick(small), small.density)], small.W, small.H,
                                 dt=small.dt, jitter_mode=1, jitter_value=0.5)
        assert np.array_equal(ns[v], rns.astype(np.int32))
        assert np.abs(rgba[v] - ref).max() <= 1e-4
    assert ns.sum() > 10000


def test_adjoint_matches_fp64_oracle(small):
    import torch
    print("\n[case] diff backward vs fp64 oracle", flush=True)
    vol = _vol(small)
    tf = torch.from_numpy(small.tf).cuda().requires_grad_(True)
    den = torch.from_numpy(small.density).cuda().requires_grad_(True)
    rgba = vol.render(tf, den)
    w = torch.from_numpy(np.random.default_rng(0).standard_normal(tuple(rgba.shape)).astype(np.float32)).cuda()
    (rgba * w).sum().backward()
    torch.cuda.synchronize()
    g_tf, g_den = tf.grad.cpu().numpy(), den.grad.cpu().numpy()
    ref_tf = np.zeros_like(g_tf, dtype=np.float64)
    ref_den = np.zeros_like(g_den, dtype=np.float64)
    wn = w.cpu().numpy()
    for v, cam in enumerate(small.cams):
        dtf, (dvox,) = O.diff_adjoint(cam, small.tf, [(_brick(small), small.density)], small.W,
                                      small.H, wn[v], dt=small.dt, jitter_mode=1, jitter_value=0.5)
        ref_tf += dtf
        ref_den += dvox
    e_tf = np.linalg.norm(g_tf - ref_tf) / np.linalg.norm(ref_tf)
    e_den = np.linalg.norm(g_den - ref_den) / np.linalg.norm(ref_den)
    print(f"relative L2 error dTF {e_tf:.2e} dDensity {e_den:.2e}")
    assert e_tf < 1e-3 and e_den < 1e-3


def test_adjoint_matches_finite_differences(small):
    """Central differences (eps 1e-3) of a float64 L2 image loss through the
    float32 forward, on 32 TF entries and the 32 voxels with the largest
    gradients.  A perturbation that changes any pixel's sample count (an
    alpha<.99 or t<t1 branch) is flagged and excluded; the loss is only
    piecewise smooth, so at most 1 in 8 of the remaining voxel entries may
    sit across a kink (TF control point, |N.d| = 0).  The float64 oracle
    adjoint (test above) covers every voxel, faces included."""
    import torch
    print("\n[case] diff backward vs central differences", flush=True)
    vol = _vol(small)
    tf0 = torch.from_numpy(small.tf).cuda()
    den0 = torch.from_numpy(small.density).cuda()
    with torch.no_grad():
        target, _ = vol.forward_only(torch.from_numpy(small.target_tf).cuda(), den0)
        target = target.clone()

    def loss(rgba):
        return ((rgba.double() - target.double()) ** 2).sum()

    tf = tf0.clone().requires_grad_(True)
    den = den0.clone().requires_grad_(True)
    loss(vol.render(tf, den)).backward()
    g_tf, g_den = tf.grad.clone(), den.grad.clone()
    eps = 1e-3

    def fd(kind, idx):
        outs, counts = [], []
        for sgn in (1, -1):
            t, d = tf0.clone(), den0.clone()
            (t if kind == "tf" else d)[idx] += sgn * eps
            rgba, ns = vol.forward_only(t, d)
            outs.append(float(loss(rgba)))
            counts.append(ns.clone())
        return (outs[0] - outs[1]) / (2 * eps), not torch.equal(counts[0], counts[1])

    rng = np.random.default_rng(5)
    nz = (g_tf.abs() > 1e-3 * float(g_tf.abs().max())).nonzero().cpu().numpy()
    tf_idx = [tuple(nz[i]) for i in rng.choice(len(nz), 32, replace=False)]
    # density entries: interior voxels where the field has a real gradient.
    # Where |grad density| ~ 0 (flat regions, clamped brick faces) the normal
    # N = g/|g+1e-4| bends on the 1e-4 scale, far below the 1e-3 step.
    dn = small.density.astype(np.float64)
    gm = np.sqrt(sum(np.gradient(dn, axis=a) ** 2 for a in range(3)))
    inner = np.zeros_like(gm, dtype=bool)
    inner[3:-3, 3:-3, 3:-3] = True
    gabs = g_den.abs().cpu().numpy()
    cand = np.argwhere(inner & (gm > 0.02))
    # the 256 such voxels with the largest adjoint magnitude (face voxels can
    # hold far larger values, so a fraction of the global maximum excludes all)
    cand = cand[np.argsort(-gabs[tuple(cand.T)], kind="stable")[:256]]
    den_idx = [tuple(int(v) for v in cand[i]) for i in rng.choice(len(cand), 32, replace=False)]
    report = {}
    for kind, picks, g in (("tf", tf_idx, g_tf), ("den", den_idx, g_den)):
        gmax = float(g.abs().max())
        errs, flagged = [], 0
        for idx in picks:
            f, flipped = fd(kind, idx)
            if flipped:
                flagged += 1
                continue
            a = float(g[idx])
            errs.append(abs(f - a) / (abs(f) + 1e-3 * gmax))
            if errs[-1] > 5e-2:
                print(f"  {kind}{idx}: fd {f:.6g} adjoint {a:.6g}")
        errs = np.array(errs)
        report[kind] = (len(errs), flagged, float(np.median(errs)), int((errs > 5e-2).sum()))
    print("FD (compared, flagged, median rel err, >5% outliers):", report)
    n_tf, _, med_tf, bad_tf = report["tf"]
    n_den, _, med_den, bad_den = report["den"]
    assert n_tf >= 16 and bad_tf == 0 and med_tf < 1e-2
    assert n_den >= 8 and bad_den <= n_den // 8 and med_den < 2e-2


def test_adam_fit_reduces_loss(small):
    import torch
    print("\n[case] 40 Adam steps", flush=True)
    vol = _vol(small)
    den0 = torch.from_numpy(small.density).cuda()
    with torch.no_grad():
        target, _ = vol.forward_only(torch.from_numpy(small.target_tf).cuda(),
                                     torch.from_numpy(small.target_density).cuda())
        target = target.clone()
    tf = torch.from_numpy(small.tf).cuda().requires_grad_(True)
    den = den0.clone().requires_grad_(True)
    opt = torch.optim.Adam([tf, den], lr=2e-3)
    losses = []
    for _ in range(40):
        opt.zero_grad()
        l = ((vol.render(tf, den) - target) ** 2).sum()
        l.backward()
        opt.step()
        with torch.no_grad():
            tf.clamp_(0, 1)
        losses.append(float(l.detach()))
    print(f"loss {losses[0]:.4f} -> {losses[-1]:.4f}")
    # lr 2e-3 is the stable step here (5e-3 already diverges on this case);
    # 40 steps take the loss down ~21 % (the same fit through the oracle's
    # float64 adjoint on the CPU: 1.5058 -> 1.1888), still falling at the end
    assert losses[-1] < 0.85 * losses[0]
    assert losses[-1] == min(losses[-5:])


def test_multi_brick_forward_and_adjoint():
    """One rank marching 4 chopSuey bricks (ghost voxels shared between
    neighbours) through the C ABI: forward sample counts identical and RGBA
    within 1e-4 of the oracle; dTF and every brick's dDensity within 1e-3
    (relative L2) of the oracle's float64 adjoint."""
    import ctypes as C
    import torch
    import aloha_amd as A
    from aloha_amd._lib import lib
    print("\n[case] diff forward + adjoint, 40^3 in 4 bricks, 2 views 64x56", flush=True)
    wl = WL.diff256(size=40, views=2, W=64, H=56)
    dims = wl.density.shape[::-1]
    bricks = A.split_volume(dims, 4)
    vox = WL.brick_subvolumes(wl.density, bricks)
    V, W, H = len(wl.cams), wl.W, wl.H
    with A.Context(0, stream=torch.cuda.current_stream()) as ctx:
        ids = [ctx.upload_brick(d, v) for d, v in zip(bricks, vox)]
        ctx.set_tf(wl.tf)
        cams = np.concatenate([np.concatenate([c["org"], c["dir_00"], c["dir_du"], c["dir_dv"]])
                               for c in wl.cams]).astype(np.float32)
        cp = cams.ctypes.data_as(C.POINTER(C.c_float))
        idarr = (C.c_int32 * len(ids))(*ids)
        prm = A.render_params(dt=wl.dt, jitter_mode=1, jitter_value=0.5)
        rgba = torch.full((V, H, W, 4), float("nan"), device="cuda")
        ns = torch.full((V, H, W), -1, dtype=torch.int32, device="cuda")
        ctx._check(lib().aloha_diff_forward(ctx._h, C.byref(prm), W, H, V, cp, len(ids), idarr,
                                            C.c_void_p(rgba.data_ptr()), C.c_void_p(ns.data_ptr())))
        w = torch.from_numpy(np.random.default_rng(1).standard_normal((V, H, W, 4))
                             .astype(np.float32)).cuda()
        dtf = torch.zeros(tuple(wl.tf.shape), dtype=torch.float32, device="cuda")
        dd = [torch.zeros(v.shape, dtype=torch.float32, device="cuda") for v in vox]
        ptrs = (C.c_void_p * len(dd))(*[t.data_ptr() for t in dd])
        ctx._check(lib().aloha_diff_backward(ctx._h, C.byref(prm), W, H, V, cp, len(ids), idarr,
                                             C.c_void_p(rgba.data_ptr()), C.c_void_p(ns.data_ptr()),
                                             C.c_void_p(w.data_ptr()), C.c_void_p(dtf.data_ptr()),
                                             ptrs))
        torch.cuda.synchronize()
        g_rgba, g_ns = rgba.cpu().numpy(), ns.cpu().numpy()
        g_tf, g_dd = dtf.cpu().numpy(), [t.cpu().numpy() for t in dd]
    blist = list(zip(bricks, vox))
    ref_tf = np.zeros(wl.tf.shape, np.float64)
    ref_dd = [np.zeros(v.shape, np.float64) for v in vox]
    wn = w.cpu().numpy()
    for v, cam in enumerate(wl.cams):
        r, rns = O.diff_render(cam, wl.tf, blist, W, H, dt=wl.dt, jitter_mode=1, jitter_value=0.5)
        assert np.array_equal(g_ns[v], rns.astype(np.int32)), v
        assert np.abs(g_rgba[v] - r).max() <= 1e-4, v
        t, dv = O.diff_adjoint(cam, wl.tf, blist, W, H, wn[v], dt=wl.dt, jitter_mode=1,
                               jitter_value=0.5)
        ref_tf += t
        for i in range(len(vox)):
            ref_dd[i] += dv[i]
    assert g_ns.sum() > 5000
    e_tf = np.linalg.norm(g_tf - ref_tf) / np.linalg.norm(ref_tf)
    errs = [np.linalg.norm(g - r) / np.linalg.norm(r) for g, r in zip(g_dd, ref_dd)]
    print(f"relative L2 error dTF {e_tf:.2e}, dDensity per brick {['%.2e' % e for e in errs]}")
    assert e_tf < 1e-3 and max(errs) < 1e-3


def test_input_checks(small):
    """A density grid of another shape, a TF that is not (n, 4) or host
    tensors are refused before the device copy."""
    import torch
    print("\n[case] diff input checks", flush=True)
    vol = _vol(small)
    tf = torch.from_numpy(small.tf).cuda()
    den = torch.from_numpy(small.density).cuda()
    for bad_tf, bad_den in ((tf, den[:-1]), (tf[:, :3], den), (tf.cpu(), den)):
        with pytest.raises(ValueError):
            vol.forward_only(bad_tf, bad_den)
    rgba, ns = vol.forward_only(tf, den)
    torch.cuda.synchronize()
    assert int(ns.sum()) > 0


def test_adjoint_is_bit_reproducible():
    """The adjoint's sums are 64-bit fixed point per spatial bin (integer
    adds, in any order) and the bins are folded in a fixed order: two
    backward passes over the same inputs give bit-identical dTF and
    dDensity, at the full 256^3 x 8 views x 512^2 shape."""
    import torch
    print("\n[case] diff backward twice: bit-identical gradients", flush=True)
    wl = WL.diff256()
    vol = _vol(wl)
    grads = []
    w = torch.from_numpy(np.random.default_rng(7).standard_normal((len(wl.cams), wl.H, wl.W, 4))
                         .astype(np.float32)).cuda()
    for _ in range(2):
        tf = torch.from_numpy(wl.tf).cuda().requires_grad_(True)
        den = torch.from_numpy(wl.density).cuda().requires_grad_(True)
        (vol.render(tf, den) * w).sum().backward()
        torch.cuda.synchronize()
        grads.append((tf.grad.clone(), den.grad.clone()))
    assert float(grads[0][1].abs().sum()) > 0 and float(grads[0][0].abs().sum()) > 0
    assert torch.equal(grads[0][0], grads[1][0])
    assert torch.equal(grads[0][1], grads[1][1])
