"""Device TLAS build (rend_hip_set_tlas_build(REND_TLAS_GPU)) against the
reference's host-built SAH TLAS (rend_prepdynamic, rend.c:3073-3090).

The trees differ, the hits may not.  Bar: id, t, u and v bit-exact with the
oracle's traversal of the host tree, except where two hits are tied to
within the triangle test's reciprocal error (at most 1e-3 of the rays,
distances within 7.4e-4 relative): the traversal culls against the running
closest distance, so which of two such hits is reported depends on visit
order.  Any-hit flags are equal.  The scene is created with
the host build, so the oracle reads the reference's tree from rd->bnodes;
the device build leaves that region alone.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from nodebleed_amd import REND_TLAS_GPU, REND_TLAS_HOST, Scene, check, lib
from oracle_lib import or_occluded, or_trace
from raygen import aimed_rays, edge_rays, random_rays

pytestmark = pytest.mark.gpu

MISS = 0xFFFFFFFF
F3 = C.c_float * 3
Q4 = C.c_float * 4


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def rebuild(s, mode):
    L = lib()
    check(L.rend_hip_set_tlas_build(mode))
    L.rend_prepdynamic(C.c_void_p(s.rd))


def device_hits(s, rays):
    """GPU hits with the TLAS built on the device; restores the host mode."""
    try:
        rebuild(s, REND_TLAS_GPU)
        hits = s.trace(rays)
        r = rays.copy()
        with np.errstate(over="ignore"):
            tm = np.where(hits["id"] != MISS, hits["t"] * 1.5, 1.0).astype(np.float32)
        r["tmax"] = np.minimum(r["tmax"], tm)
        occ = s.occluded(r)
    finally:
        lib().rend_hip_set_tlas_build(REND_TLAS_HOST)
    return hits, r, occ


def assert_same_hits(s, rays, tag, pairs=False):
    """pairs: instances 2i and 2i + 1 coincide, so a hit may name either."""
    rebuild(s, REND_TLAS_HOST)
    want, _ = or_trace(s, rays)
    got, shadow, occ = device_hits(s, rays)
    other = (got["id"] != want["id"]) | (got["t"].view(np.uint32) != want["t"].view(np.uint32))
    for i in np.flatnonzero(other)[:8]:
        say(f"    {tag} ray {i}: got id {got['id'][i]:#x} t {got['t'][i]!r}, "
            f"oracle id {want['id'][i]:#x} t {want['t'][i]!r}")
    # Another hit may be reported only where two hits are (nearly) tied.
    # A triangle's t comes from the 12-bit RCPSS reciprocal (relative error
    # up to 1.5 * 2^-12 = 3.7e-4) while boxes are culled against the running
    # closest t with accurate slab distances: of two hits within that error
    # of each other, the one visited first can cull the other's box, and the
    # trees visit in other orders.  Two such t differ by at most 7.4e-4.
    g, w = got[other], want[other]
    assert ((g["id"] != MISS) & (w["id"] != MISS)).all(), f"{tag}: hit against miss"
    near = np.abs(g["t"].astype(np.float64) - w["t"]) <= 7.4e-4 * np.maximum(g["t"], w["t"])
    assert near.all(), f"{tag}: {int((~near).sum())} hits differ by more than a near tie"
    if pairs:
        assert ((g["id"] >> 12) == (w["id"] >> 12)).all() and (
            ((g["id"] & 0xFFF) ^ 1) == (w["id"] & 0xFFF)).all(), tag
    else:
        assert other.mean() < 1e-3, f"{tag}: {other.sum()} hits differ"
    same = ~other
    assert np.array_equal(got["u"][same], want["u"][same]) and np.array_equal(
        got["v"][same], want["v"][same]), tag
    assert np.array_equal(occ, or_occluded(s, shadow)), f"{tag}: any-hit"
    say(f"  {tag}: {len(rays)} rays, {int((want['id'] != MISS).sum())} hits, "
        f"{int(other.sum())} near ties")


def test_instanced_scene_hits_match_host_tlas():
    s = Scene(config=3, detail=0.03)
    try:
        say("device TLAS: 4096-instance scene")
        for kind, rays in (("random", random_rays(s, 100_000, seed=31)),
                           ("aimed", aimed_rays(s, 50_000, seed=32)),
                           ("edge", edge_rays(s, seed=33))):
            assert_same_hits(s, rays, kind)
    finally:
        s.close()


def test_overlapping_instance_boxes():
    """Instance boxes inflated to several times their size: a tree of heavily
    overlapping boxes (both builds see the same boxes)."""
    s = Scene(config=3, detail=0.03)
    try:
        ab = s.aabbs()
        rng = np.random.default_rng(5)
        c = 0.5 * (ab["min"] + ab["max"])
        e = 0.5 * (ab["max"] - ab["min"]) * rng.uniform(1.0, 4.0, (len(ab), 1)).astype(np.float32)
        ab["min"], ab["max"] = c - e, c + e
        say("device TLAS: inflated boxes")
        assert_same_hits(s, random_rays(s, 100_000, seed=34), "inflated")
    finally:
        s.close()


def small_scene(k, seed, duplicates=False):
    """k instances of a displaced icosphere at random poses; with duplicates
    every second instance repeats the previous one (exact distance ties)."""
    rng = np.random.default_rng(seed)

    def fill(L, rs):
        m = L.rs_material(rs, 0.6, 0.6, 0.6, 0)
        assert L.rs_mesh_begin(rs) == 0
        assert L.rs_icosphere(rs, 2, 0.0, 0.0, 0.0, 1.0, 0.1, 2.0, 3, m) == 0
        mesh = L.rs_mesh_end(rs)
        prev = None
        for i in range(k):
            if duplicates and i % 2 == 1:
                t, q, sc = prev
            else:
                q = rng.normal(size=4)
                q /= np.linalg.norm(q)
                t, sc = rng.uniform(-5, 5, 3), rng.uniform(0.5, 2.0)
                prev = (t, q, sc)
            assert L.rs_instance(rs, mesh, F3(*t), Q4(*q), F3(sc, sc, sc), 0) == i
        assert L.rs_camera(rs, F3(0, 2, 14), F3(0, 0, 0), 50.0, 1.0, 0.0) == 0

    return Scene.from_builder(fill)


@pytest.mark.parametrize("k", [1, 2, 3, 8, 9, 64, 300])
def test_small_instance_counts(k):
    s = small_scene(k, seed=40 + k)
    try:
        say(f"device TLAS: {k} instances")
        assert_same_hits(s, random_rays(s, 40_000, seed=k), f"{k} instances")
        assert_same_hits(s, aimed_rays(s, 20_000, seed=k + 1), f"{k} instances aimed")
    finally:
        s.close()


def test_coincident_instances_tie():
    """Pairs of instances with identical transforms: every hit on them is an
    exact tie, and the reported distance must still be the oracle's."""
    s = small_scene(16, seed=7, duplicates=True)
    try:
        say("device TLAS: coincident pairs")
        assert_same_hits(s, aimed_rays(s, 30_000, seed=8), "coincident", pairs=True)
    finally:
        s.close()


DNODE = np.dtype([("minx", "<f4", 8), ("maxx", "<f4", 8), ("miny", "<f4", 8), ("maxy", "<f4", 8),
                  ("minz", "<f4", 8), ("maxz", "<f4", 8), ("child", "<u4", 8), ("inv", "<u4", 8)])
FMAX = np.float32(3.4028235e38)


def device_tree(s):
    L = lib()
    try:
        rebuild(s, REND_TLAS_GPU)
        nn = C.c_uint(0)
        check(L.rend_hip_tlas_download(s.rd, None, 0, C.byref(nn)), "tlas size")
        out = np.zeros(nn.value, DNODE)
        check(L.rend_hip_tlas_download(s.rd, out.ctypes.data, out.nbytes, C.byref(nn)),
              "tlas download")
    finally:
        L.rend_hip_set_tlas_build(REND_TLAS_HOST)
    return out


@pytest.mark.parametrize("case", ["instances", "inflated", "k9", "k300", "coincident"])
def test_device_tree_structure(case):
    """The downloaded device TLAS: every instance in exactly one leaf slot,
    every node but the root referenced once at a 256-byte offset, child
    boxes equal to the union of their subtree's instance boxes, per-octant
    ranks a permutation of 0..7, empty slots the reference's empty box with
    a zero word, at most n - 1 nodes."""
    if case in ("instances", "inflated"):
        s = Scene(config=3, detail=0.03)
    else:
        s = small_scene({"k9": 9, "k300": 300, "coincident": 16}[case], seed=3,
                        duplicates=case == "coincident")
    try:
        ab = s.aabbs()
        if case == "inflated":
            c = 0.5 * (ab["min"] + ab["max"])
            ab["min"], ab["max"] = c - 3 * (c - ab["min"]), c + 3 * (ab["max"] - c)
        lo, hi = np.array(ab["min"]), np.array(ab["max"])
        say(f"device tree structure: {case}, {len(lo)} instances")
        t = device_tree(s)
        n = len(lo)
        assert 1 <= len(t) <= max(n - 1, 1)
        seen, refs = np.zeros(n, int), np.zeros(len(t), int)

        def walk(i, depth):
            assert depth < 64
            nd = t[i]
            blo, bhi = np.full(3, FMAX), np.full(3, -FMAX)
            for o in range(8):
                ranks = [(int(nd["inv"][o]) >> (3 * c)) & 7 for c in range(8)]
                assert sorted(ranks) == list(range(8)), (i, o, ranks)
            for c in range(8):
                w = int(nd["child"][c])
                box_lo = np.array([nd["minx"][c], nd["miny"][c], nd["minz"][c]])
                box_hi = np.array([nd["maxx"][c], nd["maxy"][c], nd["maxz"][c]])
                if w == 0:
                    assert (box_lo == FMAX).all() and (box_hi == -FMAX).all(), (i, c)
                    continue
                if w & 0x80000000:
                    k = w & 0x7FFFFFFF
                    assert k < n
                    seen[k] += 1
                    sub_lo, sub_hi = lo[k], hi[k]
                else:
                    assert w % 256 == 0 and 0 < w // 256 < len(t), (i, c, w)
                    refs[w // 256] += 1
                    sub_lo, sub_hi = walk(w // 256, depth + 1)
                assert (box_lo == sub_lo).all() and (box_hi == sub_hi).all(), (i, c)
                blo, bhi = np.minimum(blo, sub_lo), np.maximum(bhi, sub_hi)
            return blo, bhi

        walk(0, 0)
        assert (seen == 1).all(), "every instance exactly once"
        assert refs[0] == 0 and (refs[1:] == 1).all(), "every node referenced once"
    finally:
        s.close()


def test_rebuild_after_moving_instances():
    """Instances moved between builds (new globinv and boxes, as upd_rinsts
    does each frame): the device tree follows."""
    s = Scene(config=3, detail=0.03)
    try:
        ins, ab = s.insts(), s.aabbs()
        rng = np.random.default_rng(9)
        shift = rng.uniform(-3, 3, (len(ins), 3)).astype(np.float32)
        # world position p' = p + shift: globinv maps p' to the same object point
        g = ins["globinv"].reshape(-1, 3, 4)
        g[:, :, 3] -= np.einsum("nij,nj->ni", g[:, :, :3], shift)
        ins["globinv"] = g.reshape(len(ins), -1)
        ab["min"] += shift
        ab["max"] += shift
        say("device TLAS: moved instances")
        assert_same_hits(s, random_rays(s, 100_000, seed=35), "moved")
    finally:
        s.close()


def test_path_tracing_through_device_tlas():
    """Path-traced frames of the instanced scene (graph-replayed) with the
    host tree, then with the device tree after a rebuild: the same samples
    give the same image except where a near tie picks another hit."""
    s = Scene(config=3, detail=0.03, width=96, height=64, spp=4)
    try:
        say("device TLAS: path tracing, host tree then device tree")
        rebuild(s, REND_TLAS_HOST)
        a, ra = s.render(frames=3, maxdepth=4)
        try:
            rebuild(s, REND_TLAS_GPU)
            b, rb = s.render(frames=3, maxdepth=4)
        finally:
            lib().rend_hip_set_tlas_build(REND_TLAS_HOST)
        same = (a == b).all(axis=2)
        rel = float(((a - b) ** 2).sum() / max((a ** 2).sum(), 1e-30))
        say(f"  {same.mean():.5f} of pixels identical, relMSE {rel:.2e}, rays {ra} vs {rb}")
        assert a.sum() > 0 and same.mean() > 0.999 and rel < 1e-4
        assert abs(ra - rb) <= 1e-3 * ra
    finally:
        s.close()
