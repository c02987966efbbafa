"""The drop-in frame on the GPU against the reference's own frames
(tests/golden/frames_*.npz, written from the reference's rend_render on the
build host, whose RCPSS/RSQRTPS tables the test selects).

rend_render is called from several host threads per frame with main.c's
frame protocol (blknum, rays, samples; main.c:407-426); rd->buf (the
application's ARGB buffer) and rd->acc (the accumulator) after three frames
must equal, bit for bit, the reference's block walk traced with its
single-ray trace_rays.  The reference's rend_render traces coherent packets
with its packet traversal, whose slab tests use an approximate reciprocal;
the pixels where that frame differs from the single-ray one are reported
and must be exactly the pixels where librend's frame differs from it."""
import ctypes as C
import sys
import threading

import numpy as np
import pytest

from frames_ref import NAMES, load, scene_of
from nodebleed_amd import (REND_MODE_NORMALS, REND_RCP_HOST_TABLE, REND_RCP_XEON_TABLE, Scene,
                           check, lib)

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.mark.parametrize("name", NAMES)
def test_rend_render_matches_the_reference_frames(name):
    f = load(name)
    L = lib()
    w, h, frames = int(f["width"]), int(f["height"]), int(f["frames"])
    s = scene_of(f)
    # the fixtures were written on the build host (an Intel Xeon): its
    # RCPSS/RSQRTPS tables reproduce them on any host
    check(L.rend_hip_set_rcp(REND_RCP_XEON_TABLE), "rcp")
    try:
        L.rend_staticinit(*(int(b) for b in f["block"]))
        check(L.rend_hip_set_mode(REND_MODE_NORMALS), "mode")
        L.rend_clraccum(s.rd)
        buf = np.zeros(w * h, np.uint32)
        samples, rays = 0, 0
        for fr in range(frames):
            L.rs_rdata_setframe(s.rd, 0, 0, samples, buf.ctypes.data)
            say(f"{name}: frame {fr}, 4 threads call rend_render")
            th = [threading.Thread(target=L.rend_render, args=(C.c_void_p(s.rd),))
                  for _ in range(4)]
            for t in th:
                t.start()
            for t in th:
                t.join()
            v = s.view()
            # rays are added only by a frame that rendered without error
            assert v.rays == w * h, (v.rays, L.rend_hip_errstr())
            rays += v.rays
            samples += v.spp
        assert rays == int(f["rays_single"])
        v = s.view()
        acc = np.frombuffer((C.c_float * (w * h * 3)).from_address(v.acc), np.float32)
        acc = acc.reshape(h, w, 3).copy()
        img = buf.reshape(h, w)
        bad = np.flatnonzero((acc != f["acc_single"]).any(axis=2) | (img != f["buf_single"]))
        assert len(bad) == 0, f"{len(bad)} pixels differ from the reference's single-ray frame"
        pk = set(np.flatnonzero((acc != f["acc_packets"]).any(axis=2) |
                                (img != f["buf_packets"])).tolist())
        say(f"{name}: {len(pk)} of {w * h} pixels differ from rend_render's packet frame "
            f"(its packet traversal's own hits)")
        assert pk == set(f["packet_diff_px"].tolist())
    finally:
        L.rend_hip_set_rcp(REND_RCP_HOST_TABLE)
        s.close()
