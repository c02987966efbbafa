"""The frame fixtures against the reference itself: whole frames of the
reference's rend_render (rend.c:3603-3708) and of its single-ray shading
path, recomputed live from oracle/_ref/libref.so where it has been built
from the reference's sources, must equal the committed fixtures bit for
bit; and the fixtures' own invariants (one ray per pixel per frame, opaque
ARGB, packets and single rays agreeing on all but a few pixels)."""
import numpy as np
import pytest

from frames_ref import NAMES, load, scene_of
from nodebleed_amd import Scene
from oracle_lib import RefScene, reference


@pytest.mark.parametrize("name", NAMES)
def test_frame_fixture_invariants(name):
    f = load(name)
    w, h, n = int(f["width"]), int(f["height"]), int(f["frames"])
    assert int(f["rays_single"]) == int(f["rays_packets"]) == w * h * n
    for k in ("buf_single", "buf_packets"):
        assert (f[k] >> 24 == 0xFF).all()
    # normal tint (n * 0.5 + 0.5) * col: finite and non-negative (emitters'
    # col is their radiance, so values above 1 wrap in the 8-bit pack)
    assert np.isfinite(f["acc_single"]).all() and (f["acc_single"] >= 0).all()
    assert len(f["packet_diff_px"]) <= w * h // 1000 + 4


@pytest.mark.skipif(reference() is None, reason="reference not built here (oracle/_ref)")
@pytest.mark.parametrize("name", NAMES)
def test_frame_fixture_matches_live_reference(name):
    f = load(name)
    s = scene_of(f)
    r = RefScene(s)
    try:
        blk = tuple(int(b) for b in f["block"])
        buf, acc, rays = r.frames(s, int(f["frames"]), block=blk, threads=3)
        assert rays == int(f["rays_packets"])
        assert np.array_equal(buf, f["buf_packets"])
        assert np.array_equal(acc.view(np.uint32), f["acc_packets"].view(np.uint32))
        buf, acc, rays = r.frames(s, int(f["frames"]), block=blk, single=True)
        assert np.array_equal(buf, f["buf_single"])
        assert np.array_equal(acc.view(np.uint32), f["acc_single"].view(np.uint32))
    finally:
        r.close()
        s.close()
