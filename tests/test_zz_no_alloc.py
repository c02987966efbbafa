"""The allocation-free closest-hit entry point (rend_hip_trace_noalloc)
against rend_hip_trace through the C-ABI: identical hit records, bit for
bit, on two calls in a row, the second with fewer rays; and the documented
refusals (scene not on the device, too many rays) as clear errors."""
import ctypes as C
import sys

import numpy as np
import pytest

from nodebleed_amd import HIT, RAY, REND_RCP_HOST_TABLE, REND_RCP_IEEE, Scene, check, lib
from raygen import aimed_rays, random_rays

pytestmark = pytest.mark.gpu

REND_ESTATE, REND_EINVAL = -4, -3


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def _trace_both(L, s, rays):
    n = len(rays)
    st = L.rend_hip_stream()
    d = [L.rend_hip_malloc(max(1, n) * RAY.itemsize), L.rend_hip_malloc(max(1, n) * HIT.itemsize),
         L.rend_hip_malloc(max(1, n) * HIT.itemsize)]
    try:
        assert all(d)
        check(L.rend_hip_memcpy_htod(d[0], rays.ctypes.data, n * RAY.itemsize), "upload")
        check(L.rend_hip_trace(s.rd, d[0], d[1], n, st, None), "trace")
        check(L.rend_hip_trace_noalloc(s.rd, d[0], d[2], n, st), "trace_noalloc")
        check(L.rend_hip_sync(), "sync")
        a, b = np.zeros(n, HIT), np.zeros(n, HIT)
        check(L.rend_hip_memcpy_dtoh(a.ctypes.data, d[1], n * HIT.itemsize), "download")
        check(L.rend_hip_memcpy_dtoh(b.ctypes.data, d[2], n * HIT.itemsize), "download")
        return a, b
    finally:
        for p in d:
            L.rend_hip_free(p)


@pytest.mark.parametrize("config,rcp", [(2, REND_RCP_HOST_TABLE), (3, REND_RCP_HOST_TABLE),
                                        (3, REND_RCP_IEEE)])
def test_noalloc_trace_matches_trace(config, rcp):
    L = lib()
    s = Scene(config=config, detail=0.1 if config == 2 else 0.03)
    check(L.rend_hip_set_rcp(rcp), "rcp")
    try:
        check(L.rend_hip_sync_scene(s.rd), "scene upload")
        rays = np.concatenate([random_rays(s, 150_000, seed=21), aimed_rays(s, 50_000, seed=22)])
        for k, sub in enumerate((rays, rays[:7_777])):
            say(f"noalloc config {config} rcp {rcp}: call {k}, {len(sub)} rays")
            a, b = _trace_both(L, s, np.ascontiguousarray(sub))
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), \
                f"call {k}: {int((a.view(np.uint32) != b.view(np.uint32)).any(axis=1).sum())} differ"
            assert (a["id"] != 0xFFFFFFFF).sum() > 500
    finally:
        L.rend_hip_set_rcp(REND_RCP_HOST_TABLE)
        s.close()


def test_noalloc_trace_refusals():
    L = lib()
    s = Scene(config=2, detail=0.05)
    try:
        check(L.rend_hip_sync_scene(s.rd), "scene upload")
        # a pending rebuild must be uploaded first: the entry point would
        # otherwise have to allocate
        L.rend_prepdynamic(C.c_void_p(s.rd))
        assert L.rend_hip_trace_noalloc(s.rd, None, None, 0, None) == REND_ESTATE
        check(L.rend_hip_sync_scene(s.rd), "scene upload")
        assert L.rend_hip_trace_noalloc(s.rd, None, None, 0, None) == 0
        assert L.rend_hip_trace_noalloc(s.rd, None, None, 0x80000000, None) == REND_EINVAL
        assert b"at most" in L.rend_hip_errstr()
    finally:
        s.close()
