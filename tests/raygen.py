"""Seeded host-side ray sets for parity tests: random rays through a scene's
bounds and the edge cases the reference's slab and triangle tests treat
specially (zero and negative-zero direction components, axis-parallel rays,
origins on box planes, short and zero tmax)."""
from __future__ import annotations

import numpy as np

from nodebleed_amd import RAY, make_rays


def bounds(scene):
    a = scene.aabbs()
    lo = a["min"].min(axis=0)
    hi = a["max"].max(axis=0)
    return lo.astype(np.float32), hi.astype(np.float32)


def random_rays(scene, n, seed=1, inside=0.5):
    """Origins in the (slightly grown) scene box, isotropic directions."""
    rng = np.random.default_rng(seed)
    lo, hi = bounds(scene)
    ext = hi - lo
    lo2, hi2 = lo - inside * ext, hi + inside * ext
    o = (lo2 + rng.random((n, 3)) * (hi2 - lo2)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return make_rays(o, d.astype(np.float32))


def aimed_rays(scene, n, seed=2):
    """Rays from outside aimed at points on random triangles (mostly hit)."""
    rng = np.random.default_rng(seed)
    tris = scene.tris()
    insts = scene.insts()
    k = rng.integers(0, len(insts), n)
    lo, hi = bounds(scene)
    # aim at the centroid of a random triangle of the instance's mesh,
    # transformed to world space by inverting globinv
    tgt = np.zeros((n, 3), np.float32)
    for j in np.unique(k):
        sel = np.nonzero(k == j)[0]
        ins = insts[j]
        t = tris[ins["triofs"] + rng.integers(0, ins["tricnt"], len(sel))]
        w = rng.dirichlet([1, 1, 1], len(sel)).astype(np.float32)
        p = w[:, :1] * t["v0"] + w[:, 1:2] * t["v1"] + w[:, 2:] * t["v2"]
        m = np.vstack([ins["globinv"].reshape(3, 4), [0, 0, 0, 1]]).astype(np.float64)
        g = np.linalg.inv(m)
        tgt[sel] = (p @ g[:3, :3].T + g[:3, 3]).astype(np.float32)
    c = (lo + hi) / 2
    r = np.linalg.norm(hi - lo)
    o = (c + rng.normal(size=(n, 3)) * r).astype(np.float32)
    d = (tgt - o)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return make_rays(o, d.astype(np.float32))


def axis_rays(scene, zero, n=64, seed=3):
    """Rays parallel to each axis whose other components are `zero` (+0.0 or
    -0.0).  With -0.0 the reference's slab test (1/d = -inf, octant taken as
    positive, x86 NaN operand rules) lets empty child slots pass, whose zero
    child word re-enters the root until its 64-entry stack assert fires
    (rend.c:834, 2089): such rays are error cases, not parity cases."""
    rng = np.random.default_rng(seed)
    lo, hi = bounds(scene)
    c = (lo + hi) / 2
    out = []
    axes = np.eye(3, dtype=np.float32)
    for s in (1.0, -1.0):
        for a in range(3):
            d = axes[a] * s
            d = np.where(d == 0, np.float32(zero), d).astype(np.float32)
            o = (lo + rng.random((n, 3)) * (hi - lo)).astype(np.float32)
            o[:, a] = (c[a] - s * 2 * (hi[a] - lo[a]))
            out.append(make_rays(o, np.repeat(d[None], n, 0)))
    return np.concatenate(out).astype(RAY)


def edge_rays(scene, seed=3):
    """Axis-parallel (+0.0 components), plane-touching and short rays."""
    rng = np.random.default_rng(seed)
    lo, hi = bounds(scene)
    out = [axis_rays(scene, 0.0, 128, seed)]
    # origins exactly on the scene box planes, random directions
    o = (lo + rng.random((256, 3)) * (hi - lo)).astype(np.float32)
    for i in range(256):
        o[i, i % 3] = lo[i % 3] if (i // 3) % 2 else hi[i % 3]
    d = rng.normal(size=(256, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    out.append(make_rays(o, d.astype(np.float32)))
    # short / zero / negative tmax
    r = random_rays(scene, 256, seed + 1)
    r["tmax"] = np.float32(rng.random(256) * np.linalg.norm(hi - lo) * 0.3)
    r["tmax"][:16] = 0.0
    r["tmax"][16:32] = -1.0
    out.append(r)
    return np.concatenate(out).astype(RAY)


def mismatches(a: np.ndarray, b: np.ndarray) -> dict:
    """Counts of differing ids and of t/u/v differing in any bit."""
    out = {"id": int((a["id"] != b["id"]).sum())}
    for f in ("t", "u", "v"):
        out[f] = int((a[f].view(np.uint32) != b[f].view(np.uint32)).sum())
    return out


def ulp_diff(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """|a - b| in units in the last place (same-sign finite floats)."""
    ia = a.view(np.int32).astype(np.int64)
    ib = b.view(np.int32).astype(np.int64)
    ia = np.where(ia < 0, -(ia & 0x7FFFFFFF), ia)
    ib = np.where(ib < 0, -(ib & 0x7FFFFFFF), ib)
    return np.abs(ia - ib)


def set_lens_camera(scene, focangle: float, focdist: float, vfov: float = 45.0) -> None:
    """Gives the scene's rdata a thin-lens camera (focus angle and distance,
    as main.c's upd_rcam reads them) at the current eye and view direction."""
    import ctypes as C
    c = scene.camera()
    eye = np.ascontiguousarray(c[0:3], np.float32)
    target = np.ascontiguousarray(c[0:3] + c[9:12], np.float32)
    v = scene.view()
    rc = scene.L.rs_set_rcam(scene.rd, eye.ctypes.data_as(C.POINTER(C.c_float)),
                             target.ctypes.data_as(C.POINTER(C.c_float)), vfov, focdist,
                             focangle, v.width, v.height)
    assert rc == 0
