"""Device BLAS build (rend_hip_set_blas_build(REND_BLAS_GPU)) against the
reference's host-built SAH BLASes (rend_prepstatic, rend.c:3033-3071).

The trees differ, the hits may not.  Bar: id, t, u and v bit-exact with the
oracle's traversal of the host trees, except where two hits are tied to
within the triangle test's reciprocal error (at most 1e-3 of the rays,
distances within 7.4e-4 relative; which of two such hits is reported
depends on visit order).  Any-hit flags are equal.  The downloaded device
trees hold every triangle exactly once, with the reference's leaf data, and
every child box is the union of its subtree's triangle boxes.
"""
import ctypes as C
import sys
import time

import numpy as np
import pytest

from nodebleed_amd import REND_BLAS_GPU, REND_BLAS_HOST, RTRI, Scene, check, lib
from oracle_lib import or_occluded, or_trace
from raygen import aimed_rays, edge_rays, random_rays

pytestmark = pytest.mark.gpu

MISS = 0xFFFFFFFF
F3 = C.c_float * 3
Q4 = C.c_float * 4
DNODE = np.dtype([("minx", "<f4", 8), ("maxx", "<f4", 8), ("miny", "<f4", 8), ("maxy", "<f4", 8),
                  ("minz", "<f4", 8), ("maxz", "<f4", 8), ("child", "<u4", 8), ("inv", "<u4", 8)])
DLEAF = np.dtype([("v0", "<f4", (3, 4)), ("e0", "<f4", (3, 4)), ("e1", "<f4", (3, 4)),
                  ("id", "<u4", 4), ("tricnt", "<u4"), ("pad", "<u4", 7)])
LEAF = 0x80000000
FMAX = np.float32(3.4028235e38)


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def built_on_device(make):
    """A scene whose BLASes are built on the device (the mode is restored)."""
    L = lib()
    check(L.rend_hip_set_blas_build(REND_BLAS_GPU))
    try:
        s = make()
        t0 = time.perf_counter()
        check(L.rend_hip_sync_scene(s.rd), "device BLAS build")
        s.build_s = time.perf_counter() - t0
    finally:
        L.rend_hip_set_blas_build(REND_BLAS_HOST)
    return s


def assert_same_hits(ref, dev, rays, tag):
    want, _ = or_trace(ref, rays)
    got = dev.trace(rays)
    other = (got["id"] != want["id"]) | (got["t"].view(np.uint32) != want["t"].view(np.uint32))
    for i in np.flatnonzero(other)[:8]:
        say(f"    {tag} ray {i}: got id {got['id'][i]:#x} t {got['t'][i]!r}, "
            f"oracle id {want['id'][i]:#x} t {want['t'][i]!r}")
    # A triangle's t comes from the 12-bit RCPSS reciprocal (relative error
    # up to 1.5 * 2^-12) while boxes are culled against the running closest
    # t with accurate slab distances: of two hits within that error of each
    # other, the one visited first can cull the other's box.
    g, w = got[other], want[other]
    assert ((g["id"] != MISS) & (w["id"] != MISS)).all(), f"{tag}: hit against miss"
    near = np.abs(g["t"].astype(np.float64) - w["t"]) <= 7.4e-4 * np.maximum(g["t"], w["t"])
    assert near.all(), f"{tag}: {int((~near).sum())} hits differ by more than a near tie"
    assert other.mean() < 1e-3, f"{tag}: {other.sum()} hits differ"
    same = ~other
    assert np.array_equal(got["u"][same], want["u"][same]) and np.array_equal(
        got["v"][same], want["v"][same]), tag
    shadow = rays.copy()
    with np.errstate(over="ignore"):
        tm = np.where(want["id"] != MISS, want["t"] * 1.5, 1.0).astype(np.float32)
    shadow["tmax"] = np.minimum(shadow["tmax"], tm)
    assert np.array_equal(dev.occluded(shadow), or_occluded(ref, shadow)), f"{tag}: any-hit"
    say(f"  {tag}: {len(rays)} rays, {int((want['id'] != MISS).sum())} hits, "
        f"{int(other.sum())} near ties")


def download(s, triofs):
    L = lib()
    n = C.c_ulonglong(0)
    check(L.rend_hip_blas_download(s.rd, triofs, None, 0, C.byref(n)), "blas size")
    buf = np.zeros(n.value, np.uint8)
    check(L.rend_hip_blas_download(s.rd, triofs, buf.ctypes.data, buf.nbytes, C.byref(n)),
          "blas download")
    return buf


def check_tree(buf, tris):
    """Walks a downloaded device BLAS: returns (nodes, leaves) visited."""
    n = len(tris)
    boxes = np.stack([np.minimum(np.minimum(tris["v0"], tris["v1"]), tris["v2"]),
                      np.maximum(np.maximum(tris["v0"], tris["v1"]), tris["v2"])], 1)
    seen = np.zeros(n, np.int64)
    nodes = leaves = 0

    def node_at(ofs):
        assert ofs % 256 == 0 and ofs + 256 <= len(buf), ofs
        return np.frombuffer(buf, DNODE, 1, ofs)[0]

    def walk(ofs):
        """Returns the (lo, hi) of the subtree at node offset ofs."""
        nonlocal nodes, leaves
        nodes += 1
        nd = node_at(ofs)
        lo_all, hi_all = np.full(3, FMAX), np.full(3, -FMAX)
        for o in range(8):
            ranks = sorted(((int(nd["inv"][o]) >> (3 * c)) & 7) for c in range(8))
            assert ranks == list(range(8)), "ranks are not a permutation"
        for c in range(8):
            w = int(nd["child"][c])
            lo = np.array([nd["minx"][c], nd["miny"][c], nd["minz"][c]])
            hi = np.array([nd["maxx"][c], nd["maxy"][c], nd["maxz"][c]])
            if w == 0:
                assert (lo == FMAX).all() and (hi == -FMAX).all(), "empty slot box"
                continue
            if w & LEAF:
                lo_ofs = w & ~LEAF
                assert lo_ofs % 32 == 0 and lo_ofs + 192 <= len(buf), lo_ofs
                lf = np.frombuffer(buf, DLEAF, 1, lo_ofs)[0]
                leaves += 1
                k = int(lf["tricnt"])
                assert 1 <= k <= 4
                ids = lf["id"][:k].astype(np.int64)
                assert len(set(ids.tolist())) == k and (ids < n).all()
                assert (lf["id"][k:] == lf["id"][k - 1]).all(), "unused lanes repeat the last"
                seen[ids] += 1
                t = tris[ids]
                np.testing.assert_array_equal(lf["v0"][:, :k].T, t["v0"])
                np.testing.assert_array_equal(lf["e0"][:, :k].T, t["v1"] - t["v0"])
                np.testing.assert_array_equal(lf["e1"][:, :k].T, t["v2"] - t["v0"])
                sub_lo, sub_hi = boxes[ids, 0].min(0), boxes[ids, 1].max(0)
            else:
                sub_lo, sub_hi = walk(w)
            np.testing.assert_array_equal(lo, sub_lo)
            np.testing.assert_array_equal(hi, sub_hi)
            lo_all, hi_all = np.minimum(lo_all, lo), np.maximum(hi_all, hi)
        return lo_all, hi_all

    walk(0)
    assert (seen == 1).all(), f"{int((seen != 1).sum())} triangles not in exactly one leaf"
    assert nodes <= n // 2 + 2
    return nodes, leaves


def mesh_scene(tris_xyz, seed=0):
    """One mesh of the given triangles (k x 3 x 3), one identity instance."""
    def fill(L, rs):
        m = L.rs_material(rs, 0.6, 0.6, 0.6, 0)
        assert L.rs_mesh_begin(rs) == 0
        v = np.ascontiguousarray(tris_xyz.reshape(-1, 3), np.float32)
        idx = np.arange(len(v), dtype=np.uint32)
        assert L.rs_mesh_data(rs, v.ctypes.data, len(v), None, idx.ctypes.data,
                              len(tris_xyz), None, m) == 0
        mesh = L.rs_mesh_end(rs)
        assert L.rs_instance(rs, mesh, F3(0, 0, 0), Q4(0, 0, 0, 1), F3(1, 1, 1), 0) == 0
        assert L.rs_camera(rs, F3(0.5, 0.5, 4), F3(0.5, 0.5, 0.5), 40.0, 1.0, 0.0) == 0

    return lambda: Scene.from_builder(fill)


def random_tris(k, seed, dup=False):
    rng = np.random.default_rng(seed)
    c = rng.uniform(0, 1, (k, 1, 3))
    t = (c + rng.uniform(-0.15, 0.15, (k, 3, 3))).astype(np.float32)
    if dup:
        t[1::2] = t[0::2][: len(t[1::2])]
    return t


@pytest.mark.parametrize("cfg,detail", [(2, 0.1), (1, 0.25), (3, 0.03)])
def test_config_scenes_match_host_blas(cfg, detail):
    ref = Scene(config=cfg, detail=detail)
    dev = built_on_device(lambda: Scene(config=cfg, detail=detail))
    try:
        say(f"device BLAS: config {cfg} detail {detail}, {ref.ntris} tris, "
            f"built in {dev.build_s * 1e3:.1f} ms (with upload)")
        for kind, rays in (("random", random_rays(ref, 60_000, seed=11)),
                           ("aimed", aimed_rays(ref, 60_000, seed=12)),
                           ("edge", edge_rays(ref, seed=13))):
            assert_same_hits(ref, dev, rays, f"cfg{cfg} {kind}")
    finally:
        ref.close()
        dev.close()


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5, 8, 9, 33, 1000])
def test_small_meshes(k):
    make = mesh_scene(random_tris(k, seed=k))
    ref, dev = make(), built_on_device(make)
    try:
        say(f"device BLAS: {k} random triangles")
        assert_same_hits(ref, dev, random_rays(ref, 20_000, seed=k), f"{k} tris")
        assert_same_hits(ref, dev, aimed_rays(ref, 20_000, seed=k + 1), f"{k} tris aimed")
        check_tree(download(dev, 0), dev.tris())
    finally:
        ref.close()
        dev.close()


def test_duplicate_triangles():
    """Pairs of identical triangles: equal Morton codes (the index breaks the
    tie) and exact distance ties on every hit."""
    make = mesh_scene(random_tris(200, seed=5, dup=True))
    ref, dev = make(), built_on_device(make)
    try:
        check_tree(download(dev, 0), dev.tris())
        want, _ = or_trace(ref, aimed_rays(ref, 20_000, seed=6))
        got = dev.trace(aimed_rays(ref, 20_000, seed=6))
        assert np.array_equal(got["t"].view(np.uint32), want["t"].view(np.uint32))
        assert ((got["id"] >> 13) == (want["id"] >> 13)).all()  # one of the pair
    finally:
        ref.close()
        dev.close()


@pytest.mark.parametrize("cfg,detail", [(2, 0.1), (3, 0.03)])
def test_device_tree_structure(cfg, detail):
    dev = built_on_device(lambda: Scene(config=cfg, detail=detail))
    try:
        ins = dev.insts()
        tris = dev.tris()
        for tofs in sorted(set(ins["triofs"].tolist()))[:4]:
            tc = int(ins["tricnt"][ins["triofs"] == tofs][0])
            nodes, leaves = check_tree(download(dev, tofs), tris[tofs:tofs + tc])
            say(f"  cfg{cfg} BLAS @{tofs}: {tc} tris, {nodes} nodes, {leaves} leaves")
    finally:
        dev.close()


def test_full_terrain_builds_and_matches():
    """The 1M-triangle terrain of the traversal benchmark."""
    ref = Scene(config=2, detail=1.0)
    dev = built_on_device(lambda: Scene(config=2, detail=1.0))
    try:
        say(f"device BLAS: 1M-tri terrain built in {dev.build_s * 1e3:.1f} ms (with upload)")
        assert_same_hits(ref, dev, random_rays(ref, 100_000, seed=21), "terrain random")
        assert_same_hits(ref, dev, aimed_rays(ref, 100_000, seed=22), "terrain aimed")
    finally:
        ref.close()
        dev.close()
