"""GPU reference-exact primary rays (rend_hip_gen_camrays_ref) against the
oracle's restatement of rend_render's camera rays, which the CPU suite pins
bit for bit to the reference's own make_camray_pckt; then the whole primary
pass (those rays through the GPU traversal) against the oracle."""
import sys

import numpy as np
import pytest

from nodebleed_amd import REND_RCP_HOST_TABLE, REND_RCP_XEON_TABLE, RAY, Scene, check, lib
from oracle_lib import or_camrays, or_trace, ref_camrays, reference
from raygen import mismatches, set_lens_camera

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.mark.parametrize("bs,samples", [((64, 32), 0), ((32, 8), 123), ((16, 16), 7)])
def test_reference_camrays_bit_exact(bs, samples):
    L = lib()
    bx, by = bs
    s = Scene(config=2, detail=0.25, width=1920, height=1080)
    L.rend_staticinit(bx, by)
    n = L.rend_hip_camrays_ref_count(s.rd)
    nblk = (1920 // bx) * (1080 // by)
    assert n == nblk * bx * by
    d = L.rend_hip_malloc(n * RAY.itemsize)
    try:
        say(f"reference camera rays {bx}x{by} blocks, samples {samples}: {n} rays")
        check(L.rend_hip_gen_camrays_ref(s.rd, d, samples, None))
        check(L.rend_hip_sync())
        got = np.zeros(n, RAY)
        check(L.rend_hip_memcpy_dtoh(got.ctypes.data, d, n * RAY.itemsize))
        want = or_camrays(s, bx, by, samples, range(nblk)).ravel()
        bad = np.flatnonzero(got.view(np.uint32).reshape(n, 8) != want.view(np.uint32).reshape(n, 8))
        assert bad.size == 0, f"{bad.size} words differ, first ray {bad[0] // 8}"
        # the primary pass: these rays through the GPU traversal, bit-exact
        # with the oracle on a sample
        rng = np.random.default_rng(samples)
        idx = np.sort(rng.choice(n, 30000, replace=False))
        say("trace a sample of them")
        hits = s.trace(got[idx])
        h_want, _ = or_trace(s, got[idx])
        assert mismatches(hits, h_want) == {"id": 0, "t": 0, "u": 0, "v": 0}
    finally:
        L.rend_hip_free(d)
        L.rend_staticinit(64, 32)
        s.close()


@pytest.mark.parametrize("bs,samples", [((64, 32), 0), ((32, 8), 9)])
def test_reference_thin_lens_camrays_bit_exact(bs, samples):
    """Thin-lens camera (rend.c:3320-3376): two more draws per packet, the
    lens angle's cos/sin from this host's libmvec through the table, origin
    on the lens, direction towards the focal plane; rays equal the oracle's
    (which calls libmvec itself), and on this host the reference's own."""
    L = lib()
    bx, by = bs
    s = Scene(config=2, detail=0.25, width=1920, height=1080)
    set_lens_camera(s, focangle=4.0, focdist=2.5)
    L.rend_staticinit(bx, by)
    n = L.rend_hip_camrays_ref_count(s.rd)
    nblk = (1920 // bx) * (1080 // by)
    d = L.rend_hip_malloc(n * RAY.itemsize)
    try:
        say(f"thin-lens reference camera rays {bx}x{by} blocks, samples {samples}: {n} rays")
        check(L.rend_hip_gen_camrays_ref(s.rd, d, samples, None))
        check(L.rend_hip_sync())
        got = np.zeros(n, RAY)
        check(L.rend_hip_memcpy_dtoh(got.ctypes.data, d, n * RAY.itemsize))
        want = or_camrays(s, bx, by, samples, range(nblk)).ravel()
        bad = np.flatnonzero(got.view(np.uint32).reshape(n, 8) != want.view(np.uint32).reshape(n, 8))
        assert bad.size == 0, f"{bad.size} words differ, first ray {bad[0] // 8}"
        assert (got["ox"] != got["ox"][0]).mean() > 0.99  # origins spread over the lens
        if reference() is not None:
            blocks = [0, nblk // 3, nblk - 1]
            ref = ref_camrays(s, bx, by, samples, blocks)
            npk8 = bx * by
            for i, b in enumerate(blocks):
                assert np.array_equal(got[b * npk8:(b + 1) * npk8].view(np.uint32),
                                      ref[i].view(np.uint32))
        rng = np.random.default_rng(samples + 1)
        idx = np.sort(rng.choice(n, 30000, replace=False))
        say("trace a sample of them")
        hits = s.trace(got[idx])
        h_want, _ = or_trace(s, got[idx])
        assert mismatches(hits, h_want) == {"id": 0, "t": 0, "u": 0, "v": 0}
    finally:
        L.rend_hip_free(d)
        L.rend_staticinit(64, 32)
        s.close()


def test_reference_camrays_reject_partial_packets():
    """Blocks that are not whole 4x2 packets have no reference walk."""
    L = lib()
    s = Scene(config=2, detail=0.05, width=256, height=128)
    d = L.rend_hip_malloc(256 * 128 * RAY.itemsize)
    try:
        L.rend_staticinit(30, 8)
        assert L.rend_hip_camrays_ref_count(s.rd) == 0
        assert L.rend_hip_gen_camrays_ref(s.rd, d, 0, None) != 0
    finally:
        L.rend_hip_free(d)
        L.rend_staticinit(64, 32)
        s.close()


@pytest.mark.parametrize("mode", ["host", "xeon"])
def test_camrays_follow_the_selected_host(mode):
    """This host's RSQRTPS (default), or the build host's tables: the GPU's
    rays equal the oracle's in the same mode and, in the default mode, the
    reference's own make_camray_pckt running on this host."""
    L = lib()
    s = Scene(config=1, detail=0.25, width=256, height=128)
    L.rend_staticinit(64, 32)
    n = L.rend_hip_camrays_ref_count(s.rd)
    d = L.rend_hip_malloc(n * RAY.itemsize)
    check(L.rend_hip_set_rcp(REND_RCP_XEON_TABLE if mode == "xeon" else REND_RCP_HOST_TABLE))
    try:
        say(f"camera rays, {mode} RSQRTPS: {n} rays")
        check(L.rend_hip_gen_camrays_ref(s.rd, d, 5, None))
        check(L.rend_hip_sync())
        got = np.zeros(n, RAY)
        check(L.rend_hip_memcpy_dtoh(got.ctypes.data, d, n * RAY.itemsize))
        nblk = (256 // 64) * (128 // 32)
        want = or_camrays(s, 64, 32, 5, range(nblk), rsqrt=mode).ravel()
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
        if mode == "host" and reference() is not None:
            ref = ref_camrays(s, 64, 32, 5, range(nblk)).ravel()
            assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    finally:
        L.rend_hip_set_rcp(REND_RCP_HOST_TABLE)
        L.rend_hip_free(d)
        s.close()
