"""Command-line tools over the C ABI: aloha_chopsuey (the volume mode of
chopSuey, chopSuey.cpp:361-440) on the CPU; aloha_headless (the
maui::Renderer C++ API, waikikiRenderer.h:28-118) on the GPU against the
oracle."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import aloha_amd as A
import oracle as O

BIN = Path(A.__file__).resolve().parent / "bin"


@pytest.mark.parametrize("dtype,name", [(np.uint8, "uint8"), (np.uint16, "uint16"),
                                        (np.float32, "float")])
def test_chopsuey_writes_reference_vol(tmp_path, dtype, name):
    rng = np.random.default_rng(0)
    raw = (rng.random((12, 10, 14)) * (255 if dtype is np.uint8 else 60000)).astype(dtype)
    src = tmp_path / "v.raw"
    raw.tofile(src)
    out = tmp_path / "v.vol"
    r = subprocess.run([str(BIN / "aloha_chopsuey"), str(src), "-dims", "14", "10", "12",
                        "-type", name, "-n", "4", "-o", str(out)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    cr, vr, sr, bricks, vox = A.vol_read(str(out))
    assert tuple(cr.upper) == (14, 10, 12) and tuple(vr.upper) == (15, 11, 13)
    ref = O.split_volume((14, 10, 12), 4)
    assert len(bricks) == len(ref) == 4
    bpc = np.dtype(dtype).itemsize
    for b, v, d in zip(bricks, vox, ref):
        assert b.as_dict()["cell"] == d["cell"] and b.as_dict()["voxel"] == d["voxel"]
        want, (lo, hi) = O.volume_get(raw, bpc, d["voxel"][0], d["voxel"][1])
        assert np.array_equal(v, want)
        assert (b.value_range.lower, b.value_range.upper) == (lo, hi)


def test_chopsuey_usage_errors(tmp_path):
    r = subprocess.run([str(BIN / "aloha_chopsuey"), "-n", "2"], capture_output=True, text=True)
    assert r.returncode != 0 and "no filename" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("P,islands,rank,mode", [(2, 1, 0, "ubyte"), (4, 2, 0, "ubyte"),
                                                 (2, 1, 1, "ubyte"), (2, 1, 0, "float"),
                                                 (4, 2, 0, "float")])
def test_headless_renderer_api_matches_oracle(tmp_path, P, islands, rank, mode):
    """maui::Renderer through aloha_headless against the oracle.  "ubyte" is
    the reference's frame (its defaults): rank frames packed to RGBA8 with the
    crosshair, blended in 8 bits; "float" blends the float rank images and
    packs once, without the crosshair."""
    print(f"\n[case] aloha_headless P={P} islands={islands} rank={rank} {mode}", flush=True)
    from aloha_amd import workloads as WL
    wl = WL.blobs128(num_ranks=P // islands, W=96, H=80, size=48)
    vols = [A.volume_extract(wl.raw, 1, b.voxel_range)[0] for b in wl.bricks]
    hdr = (A.Box3i((0, 0, 0), wl.dims), A.Box3i((0, 0, 0), tuple(d + 1 for d in wl.dims)),
           wl.bounds)
    path = tmp_path / "w.vol"
    A.vol_write(str(path), *hdr, wl.bricks, vols)
    wl.tf.tofile(tmp_path / "tf.f32")
    cam = [str(float(x)) for k in ("org", "dir_00", "dir_du", "dir_dv") for x in wl.cam[k]]
    out = tmp_path / "frame"
    extra = [] if mode == "ubyte" else ["-float-composite", "-no-crosshair"]
    r = subprocess.run([str(BIN / "aloha_headless"), str(path), "-ranks", str(P), "-islands",
                        str(islands), "-rank", str(rank), "-size", str(wl.W), str(wl.H),
                        "-camera", *cam, "-tf", str(tmp_path / "tf.f32"), "-dt", "0.5",
                        "-frames", "2", *extra, "-o", str(out)], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    got8 = np.fromfile(str(out) + ".rgba8", np.uint32).reshape(wl.H, wl.W)
    got = np.fromfile(str(out) + ".rgba", np.float32).reshape(wl.H, wl.W, 4)
    # the second frame runs with accumID = 1: average of frames 0 and 1
    isz = P // islands
    imgs = []
    for i in range(islands):
        for r_ in range(isz):
            bl = [(wl.bricks[b], vols[b]) for b in wl.rank_bricks[r_]]
            cross = mode == "ubyte"
            f0 = O.render_rank(wl.cam, wl.tf, bl, wl.W, wl.H, dt=0.5, max_depth=wl.max_depth,
                               island_id=i, num_islands=islands, crosshair=cross)
            f1 = O.render_rank(wl.cam, wl.tf, bl, wl.W, wl.H, dt=0.5, max_depth=wl.max_depth,
                               island_id=i, num_islands=islands, accum_id=1, accum=f0["accum"],
                               crosshair=cross)
            imgs.append(f1)
    order = [i * isz + r_ for i in range(islands) for r_ in wl.order]
    ref, ref8 = O.composite_over(np.stack([o["rgba"] for o in imgs]), order)
    assert np.abs(got - ref).max() <= 1e-4
    if rank != 0:
        assert np.array_equal(got8, imgs[rank]["rgba8"])
    elif mode == "ubyte":
        # IceT's 8-bit blend of the ranks' RGBA8 frames
        ref8u = O.composite_over_u8(np.stack([o["rgba8"] for o in imgs]), order)
        assert np.abs(got8.view(np.uint8).astype(int) - ref8u.view(np.uint8).astype(int)).max() <= 1
        assert (got8 == ref8u).mean() > 0.999
    else:
        assert np.abs(got8.view(np.uint8).astype(int) - ref8.view(np.uint8).astype(int)).max() <= 1
