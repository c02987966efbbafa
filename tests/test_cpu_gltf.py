"""glTF import (rs_gltf) and scene-graph animation (rs_animate) on the CPU.

Parity: the rdata after each update at every time of the fixture sequence
(triangles, normals, materials, instances with their inverse transforms,
instance boxes, camera, BLAS and TLAS bytes) equals, byte for byte, what
the reference's importer and main.c update produce (digests in
tests/golden/gltf_ref.npz, written by tests/golden/make_gltf.py from the
reference's own gltf.c/gltfimp.c/scene.c/rend.c), and, where
oracle/_ref/libref.so is built, the live reference over a longer random
time sequence.

Closed forms that do not come from the reference: the triangles are the
fixture's vertex arrays indexed as written; materials are the JSON values
rounded to float32 with emission = emissiveFactor x emissiveStrength; every
instance's world matrix (the inverse of its globinv) is the product of its
nodes' T * R * S down the hierarchy with the glTF sampler values (lerp,
slerp, STEP, cubic Hermite with tangents scaled by the key interval), and
the camera sits at its node's world translation looking down its -z axis.
"""
import ctypes as C
import json
import sys
from pathlib import Path

import numpy as np
import pytest

from nodebleed_amd import RendError, Scene, lib
from oracle_lib import RefGltf, rdata_digests, rdata_parts, reference

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE / "golden"))
import make_gltf as mg  # noqa: E402

GOLD = np.load(HERE / "golden" / "gltf_ref.npz", allow_pickle=False)
MESH, CAM, DYNAMIC, EMISSIVE, REFRACTIVE = 0x1, 0x2, 0x200, 0x1000, 0x2000


def load(name, **kw) -> Scene:
    kw.setdefault("width", mg.WIDTH)
    kw.setdefault("height", mg.HEIGHT)
    kw.setdefault("spp", mg.SPP)
    return Scene.from_gltf(mg.OUT / f"{name}.gltf", bg=mg.BG, **kw)


def doc(name):
    return json.loads((mg.OUT / f"{name}.gltf").read_text())


def accessor(d, name, i):
    a = d["accessors"][i]
    v = d["bufferViews"][a["bufferView"]]
    raw = (mg.OUT / f"{name}.bin").read_bytes()
    comp = {5121: np.uint8, 5123: np.uint16, 5125: np.uint32, 5126: np.float32}[
        a["componentType"]]
    n = {"SCALAR": 1, "VEC2": 2, "VEC3": 3, "VEC4": 4}[a["type"]]
    stride = v.get("byteStride", 0) or n * np.dtype(comp).itemsize
    base = v.get("byteOffset", 0) + a.get("byteOffset", 0)
    out = [np.frombuffer(raw, comp, n, base + k * stride) for k in range(a["count"])]
    return np.array(out).reshape(a["count"], n)


# ---- parity with the reference -------------------------------------------------

@pytest.mark.parametrize("name", mg.FIXTURES)
def test_rdata_matches_reference_golden_at_every_update(name):
    labels = list(GOLD[f"{name}_labels"])
    s = load(name)
    for i, t in enumerate(GOLD["times"]):
        mask = s.animate(float(t))
        assert mask == int(GOLD[f"{name}_masks"][i]), (t, mask)
        got = rdata_digests(s.rd, s.ntris, s.nmtls)
        assert [k for k, _ in got] == labels
        bad = [k for (k, v), g in zip(got, GOLD[f"{name}_sha256"][i]) if v != g]
        assert not bad, f"t={t}: {bad} differ from the reference"
    s.close()


@pytest.mark.skipif(reference() is None, reason="reference harness not built")
@pytest.mark.parametrize("name", mg.FIXTURES)
def test_rdata_matches_live_reference_over_random_times(name):
    rng = np.random.default_rng(7)
    times = np.concatenate([np.sort(rng.uniform(-0.5, 3.0, 24)), rng.uniform(0, 2.5, 8)])
    times = times.astype(np.float32)
    s = load(name)
    r = RefGltf(mg.OUT / f"{name}.gltf", mg.WIDTH, mg.HEIGHT, mg.SPP, mg.BG)
    assert (s.ntris, s.nmtls) == (r.ntris, r.nmtls)
    for t in times:
        assert s.animate(float(t)) == r.update(float(t))
        a, b = rdata_parts(s.rd, s.ntris, s.nmtls), rdata_parts(r.rd, r.ntris, r.nmtls)
        assert [k for k, _ in a] == [k for k, _ in b]
        bad = [k for (k, x), (_, y) in zip(a, b) if x != y]
        assert not bad, f"t={t}: {bad} differ from the reference"
    r.close()
    s.close()


def test_digests_tell_update_states_apart():
    """The state after one update at 0.5 equals the golden state after the
    sequence up to 0.5, and differs from the state at 1.0 in the parts an
    animation moves: the per-part digests name what went wrong."""
    s = load("anim_hier")
    s.animate(0.5)
    got = dict(rdata_digests(s.rd, s.ntris, s.nmtls))
    i = list(GOLD["times"]).index(np.float32(0.5))
    want = dict(zip(GOLD["anim_hier_labels"], GOLD["anim_hier_sha256"][i]))
    assert got == want
    other = dict(zip(GOLD["anim_hier_labels"], GOLD["anim_hier_sha256"][i + 2]))
    assert {k for k in got if got[k] != other[k]} >= {"insts", "aabbs", "cam", "tlas"}
    s.close()


# ---- closed forms -----------------------------------------------------------------

def trs(t, q, sc):
    x, y, z, w = q
    r = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    m = np.eye(4)
    m[:3, :3] = r @ np.diag(sc)
    m[:3, 3] = t
    return m


def sample(d, name, smp, t):
    """glTF sampler value at time t (keys clamped), in float64.  Keys are
    read from the start of their buffer view, as the importer does."""
    a = d["accessors"][smp["input"]]
    view = d["bufferViews"][a["bufferView"]]
    raw = (mg.OUT / f"{name}.bin").read_bytes()
    keys = np.frombuffer(raw, np.float32, a["count"], view["byteOffset"]).astype(np.float64)
    vals = accessor(d, name, smp["output"]).astype(np.float64)
    interp = smp.get("interpolation", "LINEAR")
    k = int(np.clip(np.searchsorted(keys, t) - 1, 0, len(keys) - 2))
    t0, t1 = keys[k], keys[k + 1]
    u = (min(max(t, t0), t1) - t0) / (t1 - t0)
    if interp == "STEP":
        return vals[k]
    if interp == "CUBICSPLINE":
        a0, v0, b0 = vals[3 * k], vals[3 * k + 1], vals[3 * k + 2]
        a1, v1 = vals[3 * k + 3], vals[3 * k + 4]
        dt = t1 - t0
        p = ((2 * u ** 3 - 3 * u ** 2 + 1) * v0 + (u ** 3 - 2 * u ** 2 + u) * dt * b0 +
             (-2 * u ** 3 + 3 * u ** 2) * v1 + (u ** 3 - u ** 2) * dt * a1)
        return p / np.linalg.norm(p) if len(p) == 4 else p
    v0, v1 = vals[k], vals[k + 1]
    if len(v0) == 4:
        c = float(np.dot(v0, v1))
        if c < 0:
            return None
        ang = np.arccos(min(c, 1.0))
        if ang < 1e-9:
            return v0
        return (np.sin((1 - u) * ang) * v0 + np.sin(u * ang) * v1) / np.sin(ang)
    return (1 - u) * v0 + u * v1


def world_matrices(name, t):
    """World matrix of every glTF node reachable from the scene at time t
    (None below a rotation track between keys of negative dot product: there
    the reference's spher_lerp keeps acos(d) and flips only the second
    weight, which is not the textbook slerp; the digests pin that case)."""
    d = doc(name)
    nodes = d["nodes"]
    comp = {i: [np.array(n.get("translation", [0, 0, 0]), np.float64),
                np.array(n.get("rotation", [0, 0, 0, 1]), np.float64),
                np.array(n.get("scale", [1, 1, 1]), np.float64)] for i, n in enumerate(nodes)}
    for an in d.get("animations", []):
        for ch in an["channels"]:
            slot = {"translation": 0, "rotation": 1, "scale": 2}[ch["target"]["path"]]
            comp[ch["target"]["node"]][slot] = sample(d, name, an["samplers"][ch["sampler"]], t)
    out = {}

    def walk(i, parent):
        out[i] = None if parent is None or any(c is None for c in comp[i]) else \
            parent @ trs(*comp[i])
        for c in nodes[i].get("children", []):
            walk(c, out[i])

    for r in d["scenes"][0]["nodes"]:
        walk(r, np.eye(4))
    order = []

    def dfs(i):
        order.append(i)
        for c in nodes[i].get("children", []):
            dfs(c)

    for r in d["scenes"][0]["nodes"]:
        dfs(r)
    return out, order


def inst_world(s):
    ins = s.insts()
    out = []
    for g in ins["globinv"]:
        m = np.eye(4)
        m[:3, :] = g.reshape(3, 4)
        out.append(np.linalg.inv(m))
    return out


@pytest.mark.parametrize("t", [0.0, 0.3, 0.55, 1.0, 1.45, 2.0, 3.0])
def test_instances_and_camera_follow_the_animated_hierarchy(t):
    name = "anim_hier"
    s = load(name)
    s.animate(float(t))
    worlds, order = world_matrices(name, float(t))
    nodes = doc(name)["nodes"]
    mesh_nodes = [i for i in order if "mesh" in nodes[i] and "camera" not in nodes[i]]
    got = inst_world(s)
    assert len(got) == len(mesh_nodes) == 4
    checked = 0
    for g, i in zip(got, mesh_nodes):
        if worlds[i] is not None:
            np.testing.assert_allclose(g, worlds[i], atol=2e-5, err_msg=nodes[i]["name"])
            checked += 1
    assert checked >= 2
    cam = s.camera()
    cw = worlds[[i for i in order if "camera" in nodes[i]][0]]
    np.testing.assert_allclose(cam[0:3], cw[:3, 3], atol=1e-5)
    fwd = -cw[:3, 2] / np.linalg.norm(cw[:3, 2])
    np.testing.assert_allclose(cam[9:12], fwd, atol=1e-5)
    s.close()


def test_imported_geometry_materials_and_instances():
    name = "anim_hier"
    d = doc(name)
    s = load(name)
    s.animate(0.0)
    tris = s.tris()
    assert s.ntris == 12 + 4 + 80
    # mesh 0: two primitives over one interleaved vertex view
    bv, _, bi = mg.box()
    want = bv[bi.reshape(-1, 3)]
    np.testing.assert_array_equal(np.stack([tris["v0"][:12], tris["v1"][:12],
                                            tris["v2"][:12]], 1), want)
    nr = np.frombuffer((C.c_char * (s.ntris * 48)).from_address(s.view().nrms),
                       np.float32).reshape(-1, 12)
    mtlid = nr[:, 3].view(np.uint32)
    assert list(mtlid[:12]) == [0] * 8 + [1] * 4      # primitive materials
    assert list(mtlid[12:16]) == [0] * 4             # no material: 0
    assert set(mtlid[16:]) == {2}
    tv, tn, ti = mg.tetra()
    np.testing.assert_array_equal(nr[12:16, 0:3], tn[ti.reshape(-1, 3)[:, 0]])  # not unit
    m = np.frombuffer((C.c_char * (s.nmtls * 32)).from_address(s.view().mtls),
                      np.float32).reshape(-1, 8)
    f = m[:, 6].view(np.uint32)
    mats = d["materials"]
    np.testing.assert_array_equal(m[0, :6], np.float32([0.8, 0.2, 0.1, 0.3, 0.7, 1.5]))
    np.testing.assert_array_equal(m[1, :3], np.float32([5.0, 4.5, 4.0]))  # emission
    assert f[1] == EMISSIVE and f[0] == 0
    assert f[2] == REFRACTIVE and m[2, 5] == np.float32(1.5)  # ior is not imported
    np.testing.assert_array_equal(m[3, :6], np.float32([1, 1, 1, 0, 0.5, 1.5]))
    assert f[4] == 0  # emissiveFactor without a strength emits nothing
    assert m[4, 3] == np.float32(mats[4]["pbrMetallicRoughness"]["metallicFactor"])
    ins = s.insts()
    # DFS order: crate, ico (child of crate), tetra, crate2
    assert list(ins["flags"]) == [MESH | EMISSIVE | DYNAMIC, MESH | REFRACTIVE,
                                  MESH | DYNAMIC, MESH | EMISSIVE | DYNAMIC]
    assert list(ins["tricnt"]) == [12, 80, 4, 12]
    assert list(ins["triofs"]) == [0, 16, 12, 0]
    # instance boxes hold their transformed triangles
    for k, w in enumerate(inst_world(s)):
        lo, hi = s.aabbs()["min"][k], s.aabbs()["max"][k]
        o, n = int(ins["triofs"][k]), int(ins["tricnt"][k])
        p = np.concatenate([tris[f][o:o + n] for f in ("v0", "v1", "v2")]).astype(np.float64)
        q = p @ w[:3, :3].T + w[:3, 3]
        assert (q >= lo - 1e-4).all() and (q <= hi + 1e-4).all()
    s.close()


def test_defaults_without_materials_or_cameras():
    s = load("static_default")
    assert s.animate(0.0) == 3 and s.animate(1.0) == 0
    assert s.ntris == 12   # the VEC2 primitive is skipped
    m = np.frombuffer((C.c_char * 32).from_address(s.view().mtls), np.float32)
    np.testing.assert_array_equal(m[:6], np.float32([0.5, 0.0, 0.5, 0.0, 0.5, 1.5]))
    w = inst_world(s)
    np.testing.assert_allclose(w[0][:3, 3], [0, 0, -5], atol=1e-6)
    np.testing.assert_allclose(w[1][:3, 3], [0, 1.5, -5], atol=1e-6)  # "matrix" unread
    cam = s.camera()
    np.testing.assert_array_equal(cam[0:3], [0, 0, 0])
    np.testing.assert_allclose(cam[9:12], [0, 0, -1], atol=1e-7)
    # the default camera's field of view is 60 * 180 / pi "degrees"
    vfov = np.float32(60.0 * 180.0 / 3.14159265)
    fov = 2.0 * np.tan(np.float32(0.5 * float(vfov) * 3.14159265 / 180.0)) * 10.0
    np.testing.assert_allclose(cam[16], fov, rtol=1e-5)
    s.close()


def test_camera_and_group_animation_show_only_with_a_moving_mesh():
    s = load("cam_group")
    masks = [s.animate(t) for t in (0.0, 0.5, 1.5)]
    assert masks == [3, 2, 2]
    w0, order = world_matrices("cam_group", 0.0)
    for g, i in zip(inst_world(s), [1, 2]):
        np.testing.assert_allclose(g, w0[i], atol=1e-6)
    cam = s.camera()
    np.testing.assert_allclose(cam[0:3], w0[3][:3, 3], atol=1e-6)
    s.close()


# ---- errors -----------------------------------------------------------------------

def write(tmp_path, name, d, raw=b"\0" * 64):
    (tmp_path / f"{name}.gltf").write_text(json.dumps(d) if isinstance(d, dict) else d)
    (tmp_path / f"{name}.bin").write_bytes(raw)
    return tmp_path / f"{name}.gltf"


@pytest.mark.parametrize("case", ["missing", "nobin", "notjson", "twobuffers", "badmesh",
                                  "onekey", "weights", "noscene"])
def test_malformed_files_are_errors(tmp_path, case):
    d = doc("anim_hier")
    raw = (mg.OUT / "anim_hier.bin").read_bytes()
    if case == "missing":
        path = tmp_path / "none.gltf"
    elif case == "nobin":
        path = tmp_path / "nobin.gltf"
        path.write_text(json.dumps(d))
    elif case == "notjson":
        path = write(tmp_path, "x", '{"nodes": [', raw)
    elif case == "twobuffers":
        d["buffers"] = d["buffers"] * 2
        path = write(tmp_path, "x", d, raw)
    elif case == "badmesh":
        d["nodes"][1]["mesh"] = 7
        path = write(tmp_path, "x", d, raw)
    elif case == "onekey":
        a = d["animations"][0]["samplers"][2]["input"]
        d["accessors"][a]["count"] = 1
        path = write(tmp_path, "x", d, raw)
    elif case == "weights":
        d["animations"][0]["channels"][0]["target"]["path"] = "weights"
        path = write(tmp_path, "x", d, raw)
    else:
        d["scenes"] = []
        path = write(tmp_path, "x", d, raw)
    with pytest.raises(RendError, match="rs_gltf"):
        Scene.from_gltf(path)


def test_import_into_a_used_builder_is_an_error():
    L = lib()
    rs = L.rs_new()
    assert L.rs_material(rs, 1.0, 1.0, 1.0, 0) == 0
    assert L.rs_gltf(rs, str(mg.OUT / "anim_hier.gltf").encode()) != 0
    assert b"empty" in L.rend_hip_errstr()
    L.rs_free(rs)


def test_animate_without_a_scene_graph_is_an_error():
    s = Scene(config=1, detail=0.1)
    with pytest.raises(RendError):
        s.animate(0.0)
    s.close()
