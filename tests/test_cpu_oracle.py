"""CPU tests: the oracle against the reference's golden vectors and closed
forms, the builder against the reference's bytes, and the C-ABI surface.

No GPU is touched here (scene construction and builds are host code; device
uploads are deferred to the first trace)."""
from __future__ import annotations

import ctypes as C
import hashlib
import re
import struct
from pathlib import Path

import numpy as np
import pytest

import nodebleed_amd as nb
from nodebleed_amd import HIT, RendError, Scene, make_rays
from oracle_lib import RefScene, oracle, or_bruteforce, or_occluded, or_trace, reference
from raygen import aimed_rays, axis_rays, edge_rays, mismatches, random_rays

ROOT = Path(__file__).resolve().parent.parent
GOLDEN = ROOT / "tests" / "golden"


def f32(u: int) -> float:
    return struct.unpack("<f", struct.pack("<I", u))[0]


def bits(f: float) -> int:
    return struct.unpack("<I", struct.pack("<f", f))[0]


# ---- C-ABI surface -----------------------------------------------------------

def header_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b([a-z_][a-z0-9_]*)\s*\(", text, re.M):
            name = m.group(1)
            if name not in ("if", "for", "while", "return", "sizeof"):
                names.add(name)
    return names


def test_library_exports_every_declared_function():
    L = nb.lib()
    names = header_functions()
    # vec3.h keeps the reference's prototypes for callers' own vec3.c.
    names = {n for n in names if not n.startswith("vec3_")}
    assert {"rend_render", "rend_prepstatic", "rend_hip_trace", "rs_commit"} <= names
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    assert set(nb.EXPORTS) <= names | {"rend_init"}


def test_no_gpu_fails_loudly():
    """Without a HIP device the traversal raises; nothing falls back to CPU."""
    s = Scene(config=2, detail=0.02)
    rays = random_rays(s, 16)
    ndev = C.c_int(0)
    try:
        import torch  # noqa: F401
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is present")
    with pytest.raises(RendError):
        s.trace(rays)
    s.close()
    del ndev


# ---- reciprocal and x86 operand rules ------------------------------------------

def test_rcp_rule_known_values():
    o = oracle()
    assert bits(o.or_rcp(1.0)) == 0x3F7FF000      # RCPSS(1.0) on the capture host
    assert bits(o.or_rcp(1.5)) == 0x3F2AA000
    assert bits(o.or_rcp(0.0)) == 0x7F800000
    assert bits(o.or_rcp(-0.0)) == 0xFF800000
    assert bits(o.or_rcp(f32(0x000AE398))) == 0x7F800000   # denormal -> inf
    assert bits(o.or_rcp(f32(0x7E7FFFFF))) == 0x00800800   # smallest normal result
    assert bits(o.or_rcp(f32(0x7E800000))) == 0x00000000   # tiny result flushed
    assert bits(o.or_rcp(float("inf"))) == 0
    assert bits(o.or_rcp(float("-inf"))) == 0x80000000
    assert np.isnan(o.or_rcp(float("nan")))


def test_rcp_relative_error_bound():
    """|rcp(x) * x - 1| <= 1.5 * 2^-12 (Intel SDM bound for RCPSS)."""
    o = oracle()
    rng = np.random.default_rng(0)
    for x in rng.uniform(1.0, 2.0, 4000).astype(np.float32):
        r = o.or_rcp(float(x))
        assert abs(np.float64(r) * np.float64(x) - 1.0) <= 1.5 * 2.0 ** -12


def test_rcp_table_matches_host_instruction():
    ref = reference()
    if ref is None:
        pytest.skip("reference harness not built")
    import platform
    from nodebleed_amd import lib
    if lib().rend_hip_rcp_cpu().decode() not in open("/proc/cpuinfo").read():
        pytest.skip("host CPU differs from the table's capture CPU")
    o = oracle()
    rng = np.random.default_rng(1)
    us = rng.integers(0, 2 ** 32, 20000, dtype=np.uint64).astype(np.uint32)
    for u in us:
        x = f32(int(u))
        a, b = o.or_rcp(x), ref.ref_rcpps(x)
        assert bits(a) == bits(b) or (np.isnan(a) and np.isnan(b)), hex(int(u))
    del platform


def test_x86_minmax_operand_rules():
    o = oracle()
    nan = float("nan")
    assert o.or_xmax(nan, 1.0) == 1.0
    assert np.isnan(o.or_xmax(1.0, nan))
    assert bits(o.or_xmax(0.0, -0.0)) == 0x80000000
    assert o.or_xmin(nan, 1.0) == 1.0
    assert np.isnan(o.or_xmin(1.0, nan))


# ---- closed forms --------------------------------------------------------------

def one_triangle_scene():
    def fill(L, rs):
        m = L.rs_material(rs, 0.5, 0.5, 0.5, 0)
        L.rs_mesh_begin(rs)
        p = [np.array(v, np.float32) for v in ((0, 0, 0), (1, 0, 0), (0, 1, 0), (-1, 1, 0))]
        L.rs_quad(rs, *[nb.fptr(v) for v in p], m)
        mesh = L.rs_mesh_end(rs)
        L.rs_instance(rs, mesh, nb.fptr([0, 0, 0]), nb.fptr([0, 0, 0, 1]),
                      nb.fptr([1, 1, 1]), 0)
    return Scene.from_builder(fill)


def test_moeller_trumbore_known_answer():
    s = one_triangle_scene()
    r = make_rays(np.array([[0.25, 0.25, 1.0]], np.float32), np.array([[0, 0, -1]], np.float32))
    h, _ = or_trace(s, r, rcp_ieee=True)
    assert h["id"][0] == 0 and h["t"][0] == 1.0 and h["u"][0] == 0.25 and h["v"][0] == 0.25
    h, _ = or_trace(s, r)  # host-table reciprocal: every result scaled by rcp(det=1)
    k = f32(0x3F7FF000)
    assert h["t"][0] == np.float32(k) and h["u"][0] == np.float32(0.25 * k)
    # back face is hit too (no det sign test), a ray past tmax is not
    r2 = make_rays(np.array([[0.25, 0.25, -1.0]], np.float32), np.array([[0, 0, 1]], np.float32))
    assert or_trace(s, r2, rcp_ieee=True)[0]["id"][0] == 0
    r2["tmax"] = 0.5
    assert or_trace(s, r2, rcp_ieee=True)[0]["id"][0] == 0xFFFFFFFF
    s.close()


def test_oracle_matches_bruteforce():
    """BVH traversal equals the brute-force minimum over all triangles."""
    def fill(L, rs):
        m = L.rs_material(rs, 0.5, 0.5, 0.5, 0)
        L.rs_mesh_begin(rs)
        L.rs_icosphere(rs, 3, 0, 0, 0, 1.0, 0.1, 2.0, 7, m)
        mesh = L.rs_mesh_end(rs)
        for i, t in enumerate(((0, 0, 0), (2.5, 0, 0), (0, 2.2, 0.5))):
            q = np.array([0.1 * i, 0.2, 0.0, 1.0], np.float32)
            q /= np.linalg.norm(q)
            L.rs_instance(rs, mesh, nb.fptr(t), nb.fptr(q), nb.fptr([1.0 + 0.3 * i] * 3), 0)
    s = Scene.from_builder(fill)
    rays = np.concatenate([aimed_rays(s, 1500, seed=4), random_rays(s, 1500, seed=5)])
    got, _ = or_trace(s, rays)
    want = or_bruteforce(s, rays)
    same_t = got["t"] == want["t"]
    assert same_t.mean() > 0.999
    assert (got["id"][same_t] == want["id"][same_t]).mean() > 0.999
    s.close()


# ---- builder -------------------------------------------------------------------

def walk_blas(arena: np.ndarray):
    """Yields (node bytes, child words) over a host BLAS arena."""
    out_tris, nodes = [], 0
    stack = [0]
    while stack:
        o = stack.pop()
        nodes += 1
        n = arena[o:o + 256]
        f = n[:192].view(np.float32).reshape(6, 8)
        ch = n[192:224].view(np.uint32)
        perm = n[224:256].view(np.uint32)
        for oct_ in range(8):
            ranks = sorted(((perm[i] >> (3 * oct_)) & 7) for i in range(8))
            assert ranks == list(range(8)), "perm8 must be a permutation per octant"
        for i in range(8):
            c = int(ch[i])
            if c == 0:
                assert f[0, i] == np.float32(3.4028235e38) and f[1, i] == -np.float32(3.4028235e38)
                continue
            lo, hi = f[[0, 2, 4], i], f[[1, 3, 5], i]
            if c & nb.NODE_LEAF:
                lo_ = c & ~nb.NODE_LEAF
                leaf = arena[lo_:lo_ + 192]
                v = leaf[:144].view(np.float32).reshape(9, 4)
                ids = leaf[144:160].view(np.uint32)
                cnt = int(leaf[160:164].view(np.uint32)[0])
                assert 1 <= cnt <= 4
                for k in range(cnt):
                    v0 = v[0:3, k]
                    pts = [v0, v0 + v[3:6, k], v0 + v[6:9, k]]
                    for p in pts:  # leaf triangles inside the child box (edges
                        # are rounded in f32, so allow one ulp of slack)
                        assert (p >= lo - 1e-6 * (1 + abs(lo))).all() and \
                               (p <= hi + 1e-6 * (1 + abs(hi))).all()
                    out_tris.append(int(ids[k]))
                for k in range(cnt, 4):  # replicated lanes
                    assert ids[k] == ids[cnt - 1]
            else:
                assert c % 32 == 0 and c > o
                stack.append(c)
    return nodes, out_tris


@pytest.mark.parametrize("cfg,detail", [(2, 0.05), (1, 0.25)])
def test_builder_invariants(cfg, detail):
    s = Scene(config=cfg, detail=detail)
    assert s.hazards() == 0
    arena = s.arena(0, s.ntris * 512)
    nodes, tris = walk_blas(arena)
    assert sorted(tris) == list(range(s.ntris)), "each triangle in exactly one leaf"
    assert nodes >= s.ntris / 32
    s.close()


def test_config_sizes():
    s = Scene(config=1, detail=1.0)
    assert s.ntris == 81920 + 10 + 2 and s.hazards() == 0
    s.close()
    s = Scene(config=2, detail=1.0)
    assert s.ntris == 1_000_000 and s.hazards() == 0
    assert len(s.insts()) == 1
    s.close()
    s = Scene(config=3, detail=0.03)
    assert len(s.insts()) == 4096 and s.hazards() == 0
    s.close()


def golden_files():
    return sorted(GOLDEN.glob("ref_*.npz"))


@pytest.mark.parametrize("path", golden_files(), ids=lambda p: p.stem)
def test_builder_bytes_match_reference_golden(path):
    from golden.make_golden import regions  # noqa: F401  (same region list)
    g = np.load(path, allow_pickle=False)
    s = Scene(config=int(g["config"]), detail=float(g["detail"]))
    import importlib.util
    spec = importlib.util.spec_from_file_location("mg", GOLDEN / "make_golden.py")
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    got = [(lab, hashlib.sha256(s.arena(ofs, size).tobytes()).hexdigest())
           for lab, ofs, size in mg.regions(s)]
    assert [x[0] for x in got] == list(g["labels"])
    assert [x[1] for x in got] == list(g["sha256"])
    s.close()


@pytest.mark.parametrize("path", golden_files(), ids=lambda p: p.stem)
def test_oracle_matches_reference_golden(path):
    g = np.load(path, allow_pickle=False)
    s = Scene(config=int(g["config"]), detail=float(g["detail"]))
    got, _ = or_trace(s, g["rays"])
    assert mismatches(got, g["hits"]) == {"id": 0, "t": 0, "u": 0, "v": 0}
    occ = or_occluded(s, g["shadow_rays"])
    assert np.array_equal(occ, g["occluded"])
    s.close()


@pytest.mark.skipif(reference() is None, reason="reference harness not built")
@pytest.mark.parametrize("cfg,detail", [(2, 0.1), (3, 0.03)])
def test_oracle_and_builder_match_live_reference(cfg, detail):
    s = Scene(config=cfg, detail=detail)
    r = RefScene(s)
    ins = s.insts()
    for tofs in sorted(set(ins["triofs"].tolist())):
        tc = int(ins["tricnt"][ins["triofs"] == tofs][0])
        assert np.array_equal(s.arena(tofs * 512, tc * 512), r.arena(tofs * 512, tc * 512))
    tl = int(s.view().tlasofs) * 256
    assert np.array_equal(s.arena(tl, 2 * len(ins) * 256), r.arena(tl, 2 * len(ins) * 256))
    rays = np.concatenate([random_rays(s, 20000, seed=31), aimed_rays(s, 20000, seed=32),
                           edge_rays(s, seed=33)])
    assert mismatches(or_trace(s, rays)[0], r.trace(rays)) == {"id": 0, "t": 0, "u": 0, "v": 0}
    r.close()
    s.close()


def test_negative_zero_axis_rays_overflow_like_the_reference():
    """Rays with -0.0 components make the reference's traversal re-enter the
    root through empty slots until its stack assert fires; the oracle
    reports the same rays as overflows."""
    s = Scene(config=2, detail=0.05)
    rays = axis_rays(s, -0.0)
    h = np.zeros(len(rays), HIT)
    cnt = (C.c_ulonglong * 3)()
    n_over = oracle().or_trace(C.c_void_p(s.rd), nb.ptr(rays), nb.ptr(h), len(rays), 0, 2,
                               cnt, None)
    assert n_over > 0
    ok = axis_rays(s, 0.0)
    assert or_trace(s, ok)[0] is not None  # +0.0 components do not overflow
    s.close()


def test_noise_is_deterministic_and_bounded():
    L = nb.lib()
    rng = np.random.default_rng(2)
    pts = rng.uniform(-10, 10, (2000, 3)).astype(np.float32)
    a = np.array([L.rs_noise(*p, 1) for p in pts])
    b = np.array([L.rs_noise(*p, 1) for p in pts])
    c = np.array([L.rs_noise(*p, 2) for p in pts])
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert (np.abs(a) <= 1).all() and a.std() > 0.1
