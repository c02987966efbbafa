"""Edge cases of the traversal boundary on the GPU: empty and ragged ray
counts (not multiples of a wave, a workgroup or a claim chunk), empty ray
intervals, and a ray count well above the benchmark's, each against the
oracle (bit-exact hits, equal any-hit flags); and the iteration-limit
error path, forced from the host with a small limit."""
import sys

import numpy as np
import pytest

from nodebleed_amd import HIT, Scene, check, lib
from oracle_lib import or_occluded, or_trace
from raygen import mismatches, random_rays

pytestmark = pytest.mark.gpu

MISS = 0xFFFFFFFF


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def scene():
    s = Scene(config=3, detail=0.03)
    yield s
    s.close()


def test_zero_rays(scene):
    L = lib()
    say("edges: zero rays")
    check(L.rend_hip_trace(scene.rd, None, None, 0, None, None), "trace 0")
    check(L.rend_hip_occluded(scene.rd, None, None, 0, None), "occluded 0")
    check(L.rend_hip_sync(), "sync")
    assert len(scene.trace(random_rays(scene, 1, seed=1)[:0])) == 0


@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 255, 257, 4097, 65537])
def test_ragged_ray_counts(scene, n):
    rays = random_rays(scene, n, seed=100 + n)
    say(f"edges: {n} rays")
    want, _ = or_trace(scene, rays)
    got = scene.trace(rays)
    assert mismatches(got, want) == {"id": 0, "t": 0, "u": 0, "v": 0}, n
    assert np.array_equal(scene.occluded(rays), or_occluded(scene, rays)), n


def test_empty_intervals(scene):
    """tmax of 0, negative or the smallest normal: no hit can satisfy
    0 < t < tmax (rend.c:846-905), so every ray misses with t = tmax."""
    rays = random_rays(scene, 30_000, seed=7)
    rays["tmax"] = np.tile(np.array([0.0, -1.0, np.finfo(np.float32).tiny], np.float32),
                           10_000)
    say("edges: empty intervals")
    want, _ = or_trace(scene, rays)
    got = scene.trace(rays)
    assert mismatches(got, want) == {"id": 0, "t": 0, "u": 0, "v": 0}
    assert (got["id"] == MISS).all()
    assert np.array_equal(got["t"].view(np.uint32), rays["tmax"].view(np.uint32))
    assert not scene.occluded(rays).any()


def test_eight_million_rays(scene):
    """A launch four times the benchmark's ray count: a 40 K sample
    bit-exact, size-independent properties over all of it."""
    n = 8 << 20
    rays = random_rays(scene, n, seed=9)
    say(f"edges: {n} rays")
    got = scene.trace(rays)
    idx = np.sort(np.random.default_rng(3).choice(n, 40_000, replace=False))
    want, _ = or_trace(scene, rays[idx])
    assert mismatches(got[idx], want) == {"id": 0, "t": 0, "u": 0, "v": 0}
    hit = got["id"] != MISS
    assert hit.any() and np.isfinite(got["t"][hit]).all()
    assert (got["t"][hit] > 0).all() and (got["t"][hit] < rays["tmax"][hit]).all()
    assert np.array_equal(got["t"][~hit].view(np.uint32), rays["tmax"][~hit].view(np.uint32))
    assert ((got["id"][hit] & 0xFFF) < scene.view().instcnt).all()



def test_iteration_limit_is_reported_then_cleared(scene):
    """With a 64-step limit a ray that needs more steps (node, leaf and
    instance visits, the oracle's per-ray counts) ends early and the call
    reports REND_EOVERFLOW ("iteration limit") naming it, for closest hit
    and, with a ray that misses everything, any hit; the report clears the
    error, and with the default limit the rays trace to the oracle's
    results again."""
    L = lib()
    rays = random_rays(scene, 4096, seed=7)
    want, _, per = or_trace(scene, rays, per_ray=True)
    steps = per.astype(np.int64).sum(axis=1)
    i = int(np.argmax(steps))
    miss = np.flatnonzero(want["id"] == MISS)
    j = int(miss[np.argmax(steps[miss])]) if len(miss) else i
    say(f"edges: iteration limit 64; closest-hit ray {i} takes {steps[i]} steps, "
        f"missing ray {j} {steps[j]}")
    assert steps[i] > 64 and steps[j] > 64
    hits = np.zeros(1, HIT)
    occ = np.zeros(1, np.uint8)
    try:
        check(L.rend_hip_set_iter_limit(64), "limit")
        one = np.ascontiguousarray(rays[i:i + 1])
        rc = L.rend_hip_trace_host(scene.rd, one.ctypes.data, hits.ctypes.data, 1, None)
        msg = L.rend_hip_errstr().decode()
        assert rc == -5 and "iteration limit" in msg and "ray 0 " in msg, (rc, msg)
        one = np.ascontiguousarray(rays[j:j + 1])
        rc = L.rend_hip_occluded_host(scene.rd, one.ctypes.data, occ.ctypes.data, 1)
        msg = L.rend_hip_errstr().decode()
        assert rc == -5 and "iteration limit" in msg, (rc, msg)
    finally:
        check(L.rend_hip_set_iter_limit(4194304), "limit")
    assert mismatches(scene.trace(rays), want) == {"id": 0, "t": 0, "u": 0, "v": 0}
    assert np.array_equal(scene.occluded(rays), or_occluded(scene, rays))
