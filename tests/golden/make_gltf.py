"""Writes the glTF 2.0 fixtures and the reference's results for them.

The fixtures are synthetic scenes of our own, built here with numpy, that
exercise every part of the glTF subset the reference's importer reads
(gltf.c, gltfimp.c): materials with baseColorFactor / metallicFactor /
roughnessFactor / emissiveFactor x KHR_materials_emissive_strength /
KHR_materials_transmission, a perspective camera on a node, interleaved
(byteStride) and plain vertex views, u8/u16/u32 index accessors with and
without byteOffset, multi-primitive meshes, a node hierarchy, a node no scene
references (not imported), and STEP / LINEAR (lerp and slerp) /
CUBICSPLINE animations over translation, rotation and scale in two
animations, including a track on a node with neither mesh nor camera and an
input accessor with a byteOffset into its buffer view (the reference reads
animation data from the start of the view).  A second file has no
materials and no cameras (the importer's defaults) and a primitive it skips;
a third animates only a camera and a group node, which the reference's
update never shows (world matrices are recomputed only when a mesh node's
track ran).

    python tests/golden/make_gltf.py          # fixtures, then reference digests

When oracle/_ref/libref.so has been built from the reference's sources
(make -C oracle ref), the reference's importer and scene-graph update
(main.c init()/update(), restated in oracle/refharness.c) run over each
fixture at TIMES, and the SHA-256 of every part of its rdata after each
update is written to tests/golden/gltf_ref.npz, with, for FRAME_FIXTURES at
FRAME_TIMES, a new accumulation of FRAMES frames of the reference's block
walk traced with its single-ray trace_rays (ARGB buffer and accumulator)
and the pixels where its rend_render packet frame differs from it.
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
OUT = HERE / "gltf"
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

#: animation times of the update sequence (the last one goes back in time:
#: key search continues from the last key found)
TIMES = (0.0, 0.25, 0.5, 0.8, 1.0, 1.3, 1.9, 2.5, 0.4)
WIDTH, HEIGHT, SPP = 128, 64, 2
BG = (0.4, 0.4, 0.4)


class Writer:
    """One .bin buffer with 4-byte aligned buffer views and accessors."""

    def __init__(self):
        self.bin = bytearray()
        self.views, self.accessors = [], []

    def view(self, data: bytes, stride: int | None = None) -> int:
        while len(self.bin) % 4:
            self.bin.append(0)
        v = {"buffer": 0, "byteOffset": len(self.bin), "byteLength": len(data)}
        if stride:
            v["byteStride"] = stride
        self.bin += data
        self.views.append(v)
        return len(self.views) - 1

    def accessor(self, view: int, comp: int, count: int, type_: str, offset: int = 0) -> int:
        a = {"bufferView": view, "componentType": comp, "count": count, "type": type_}
        if offset:
            a["byteOffset"] = offset
        self.accessors.append(a)
        return len(self.accessors) - 1

    def floats(self, a, type_: str) -> int:
        a = np.ascontiguousarray(a, np.float32)
        n = {"SCALAR": 1, "VEC3": 3, "VEC4": 4}[type_]
        return self.accessor(self.view(a.tobytes()), 5126, a.size // n, type_)

    def indices(self, idx, dtype, pad: int = 0) -> int:
        comp = {np.uint8: 5121, np.uint16: 5123, np.uint32: 5125}[dtype]
        data = bytes(pad) + np.ascontiguousarray(idx, dtype).tobytes()
        return self.accessor(self.view(data), comp, len(idx), "SCALAR", pad)

    def save(self, name: str, doc: dict):
        OUT.mkdir(exist_ok=True)
        doc["asset"] = {"version": "2.0", "generator": "make_gltf.py"}
        doc["buffers"] = [{"byteLength": len(self.bin), "uri": f"{name}.bin"}]
        doc["bufferViews"] = self.views
        doc["accessors"] = self.accessors
        (OUT / f"{name}.gltf").write_text(json.dumps(doc, indent=1))
        (OUT / f"{name}.bin").write_bytes(bytes(self.bin))


def box():
    """24 vertices (4 per face) with face normals; 12 triangles."""
    v, n, idx = [], [], []
    for axis in range(3):
        for sgn in (-1.0, 1.0):
            u, w = (axis + 1) % 3, (axis + 2) % 3
            base = len(v)
            for a, b in ((-1, -1), (1, -1), (1, 1), (-1, 1)):
                p = [0.0, 0.0, 0.0]
                p[axis], p[u], p[w] = sgn * 0.5, a * 0.5, b * 0.5
                v.append(p)
                nn = [0.0, 0.0, 0.0]
                nn[axis] = sgn
                n.append(nn)
            tri = [(0, 1, 2), (0, 2, 3)] if sgn > 0 else [(0, 2, 1), (0, 3, 2)]
            for t in tri:
                idx += [base + t[0], base + t[1], base + t[2]]
    return np.array(v, np.float32), np.array(n, np.float32), np.array(idx)


def tetra():
    v = np.array([[0, 0.8, 0], [-0.7, -0.4, 0.4], [0.7, -0.4, 0.4], [0, -0.4, -0.8]],
                 np.float32)
    idx = np.array([0, 1, 2, 0, 2, 3, 0, 3, 1, 1, 3, 2])
    n = 2.0 * v  # not unit: normals are taken as they are
    return v, n.astype(np.float32), idx


def icosphere():
    t = (1.0 + 5 ** 0.5) / 2
    v = [[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t],
         [0, -1, -t], [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
         (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
         (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    v = [list(np.array(p) / np.linalg.norm(p)) for p in v]
    mid = {}

    def m(a, b):
        k = (min(a, b), max(a, b))
        if k not in mid:
            p = (np.array(v[a]) + np.array(v[b])) / 2
            v.append(list(p / np.linalg.norm(p)))
            mid[k] = len(v) - 1
        return mid[k]

    f2 = []
    for a, b, c in f:
        ab, bc, ca = m(a, b), m(b, c), m(c, a)
        f2 += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
    v = np.array(v, np.float32)
    return v, v.copy(), np.array(f2).reshape(-1)


def quat(axis, ang):
    axis = np.asarray(axis, np.float64)
    axis = axis / np.linalg.norm(axis)
    return [*(np.sin(ang / 2) * axis), np.cos(ang / 2)]


def anim_hier():
    W = Writer()
    bv, bn, bi = box()
    inter = np.concatenate([bv, bn], axis=1).astype(np.float32)  # 24 bytes a vertex
    iv = W.view(inter.tobytes(), stride=24)
    pos = W.accessor(iv, 5126, len(bv), "VEC3", 0)
    nrm = W.accessor(iv, 5126, len(bv), "VEC3", 12)
    crate = {"name": "crate", "primitives": [
        {"attributes": {"POSITION": pos, "NORMAL": nrm},
         "indices": W.indices(bi[:24], np.uint16), "material": 0},
        {"attributes": {"NORMAL": nrm, "POSITION": pos},
         "indices": W.indices(bi[24:], np.uint8, pad=4), "material": 1}]}
    tv, tn, ti = tetra()
    tet = {"name": "tetra", "primitives": [
        {"attributes": {"POSITION": W.floats(tv, "VEC3"), "NORMAL": W.floats(tn, "VEC3")},
         "indices": W.indices(ti, np.uint32)}]}
    sv, sn, si = icosphere()
    ico = {"name": "ico", "primitives": [
        {"indices": W.indices(si, np.uint16, pad=8), "material": 2,
         "attributes": {"NORMAL": W.floats(sn, "VEC3"), "POSITION": W.floats(sv, "VEC3")}}]}

    materials = [
        {"name": "rust", "pbrMetallicRoughness": {
            "baseColorFactor": [0.8, 0.2, 0.1, 1.0], "metallicFactor": 0.3,
            "roughnessFactor": 0.7}},
        {"name": "lamp", "emissiveFactor": [1.0, 0.9, 0.8],
         "pbrMetallicRoughness": {"baseColorFactor": [0.1, 0.1, 0.1, 1.0]},
         "extensions": {"KHR_materials_emissive_strength": {"emissiveStrength": 5.0}}},
        {"name": "glass", "pbrMetallicRoughness": {
            "baseColorFactor": [0.9, 0.9, 1.0, 1.0], "roughnessFactor": 0.05},
         "extensions": {"KHR_materials_transmission": {"transmissionFactor": 1.0},
                        "KHR_materials_ior": {"ior": 1.33}}},
        {"name": "plain"},
        {"name": "dim", "emissiveFactor": [0.5, 0.5, 0.5],
         "pbrMetallicRoughness": {"metallicFactor": 1e-3, "roughnessFactor": 0.30000001}},
    ]
    nodes = [
        {"name": "group", "translation": [0.5, 0.0, -6.0], "rotation": quat([0, 1, 0], 0.2),
         "children": [1, 2, 5]},
        {"name": "crate", "mesh": 0, "translation": [-1.5, 0.0, 0.0], "children": [4]},
        {"name": "tetra", "mesh": 1, "translation": [1.5, 0.2, 0.0],
         "rotation": quat([1, 1, 0], 0.4), "scale": [1.2, 1.2, 1.2]},
        {"name": "eye", "camera": 0, "translation": [0.0, 0.6, -2.0],
         "rotation": quat([1, 0, 0], -0.1)},
        {"name": "ico", "mesh": 2, "translation": [0.0, 1.5, 0.0], "scale": [0.5, 0.5, 0.5]},
        {"name": "crate2", "mesh": 0, "translation": [0.0, -1.5, 0.25]},
        {"name": "orphan", "mesh": 1, "translation": [9.0, 9.0, 9.0]},
    ]

    keys3 = W.floats([0.0, 1.0, 2.0], "SCALAR")
    # node 1 translation, LINEAR
    t1 = W.floats([[-1.5, 0, 0], [-1.0, 0.5, 0], [-1.5, 1.0, -0.5]], "VEC3")
    # node 1 rotation, LINEAR (slerp; the last key on the far hemisphere)
    r1 = W.floats([quat([0, 1, 0], 0.0), quat([0, 1, 0], 1.0),
                   list(-np.array(quat([0, 1, 0], 2.0)))], "VEC4")
    # node 2 scale, STEP
    k2 = W.floats([0.0, 0.5, 1.5], "SCALAR")
    s2 = W.floats([[1.2, 1.2, 1.2], [0.6, 1.0, 1.4], [1.0, 1.0, 1.0]], "VEC3")
    # node 0 (no mesh, no camera) translation, LINEAR
    t0 = W.floats([[0.5, 0.0, -6.0], [0.5, 0.3, -6.5], [0.5, 0.0, -7.0]], "VEC3")
    # node 5 rotation and translation, CUBICSPLINE (in-tangent, value, out-tangent)
    kc = W.floats([0.0, 1.2], "SCALAR")
    rc = W.floats([[0, 0, 0, 0], quat([0, 0, 1], 0.0), [0, 0, 0.4, 0],
                   [0, 0, 0.3, 0], quat([0, 0, 1], 0.9), [0, 0, 0, 0]], "VEC4")
    tcub = W.floats([[0, 0, 0], [0, -1.5, 0.25], [1, 0, 0],
                     [0, 1, 0], [0.5, -1.0, 0.25], [0, 0, 0]], "VEC3")
    # node 3 (camera) translation, LINEAR; its keys live at byteOffset 8 of
    # their view, which the reference ignores (it reads 0.0, 1.0 from the
    # view's start)
    kview = W.view(np.array([0.0, 1.0, 0.5, 2.0], np.float32).tobytes())
    kcam = W.accessor(kview, 5126, 2, "SCALAR", 8)
    tcam = W.floats([[0.0, 0.6, -2.0], [0.4, 0.8, -2.5]], "VEC3")

    animations = [
        {"name": "move",
         "channels": [{"sampler": 0, "target": {"node": 1, "path": "translation"}},
                      {"sampler": 1, "target": {"node": 1, "path": "rotation"}},
                      {"sampler": 2, "target": {"node": 2, "path": "scale"}},
                      {"sampler": 3, "target": {"node": 0, "path": "translation"}}],
         "samplers": [{"input": keys3, "output": t1, "interpolation": "LINEAR"},
                      {"input": keys3, "output": r1, "interpolation": "LINEAR"},
                      {"input": k2, "output": s2, "interpolation": "STEP"},
                      {"input": keys3, "output": t0, "interpolation": "LINEAR"}]},
        {"name": "spin",
         "channels": [{"sampler": 1, "target": {"node": 5, "path": "rotation"}},
                      {"sampler": 0, "target": {"node": 5, "path": "translation"}},
                      {"sampler": 2, "target": {"node": 3, "path": "translation"}}],
         "samplers": [{"input": kc, "output": tcub, "interpolation": "CUBICSPLINE"},
                      {"input": kc, "output": rc, "interpolation": "CUBICSPLINE"},
                      {"input": kcam, "output": tcam, "interpolation": "LINEAR"}]},
    ]
    doc = {"scene": 0, "scenes": [{"nodes": [0, 3]}], "nodes": nodes,
           "meshes": [crate, tet, ico], "materials": materials,
           "cameras": [{"type": "perspective", "perspective": {
               "yfov": 0.66, "znear": 0.1, "aspectRatio": 1.5}}],
           "animations": animations}
    W.save("anim_hier", doc)


def static_default():
    W = Writer()
    bv, bn, bi = box()
    pos, nrm = W.floats(bv, "VEC3"), W.floats(bn, "VEC3")
    flat = W.accessor(W.view(np.zeros(8, np.float32).tobytes()), 5126, 4, "VEC2")
    mesh = {"primitives": [
        {"attributes": {"POSITION": pos, "NORMAL": nrm}, "indices": W.indices(bi, np.uint16)},
        # POSITION is VEC2: the importer skips this primitive
        {"attributes": {"POSITION": flat, "NORMAL": nrm},
         "indices": W.indices([0, 1, 2], np.uint16)}]}
    nodes = [{"mesh": 0, "translation": [0.0, 0.0, -5.0], "children": [1]},
             # "matrix" is not read: translation only
             {"mesh": 0, "matrix": [2, 0, 0, 0, 0, 2, 0, 0, 0, 0, 2, 0, 0, 1, 0, 1],
              "translation": [0.0, 1.5, 0.0]}]
    W.save("static_default", {"scenes": [{"nodes": [0]}], "nodes": nodes, "meshes": [mesh]})


def cam_group():
    """Animated camera and animated group node, static mesh nodes: the
    reference's update() recomputes world matrices only when a mesh node's
    track ran, so neither animation shows after the first update."""
    W = Writer()
    bv, bn, bi = box()
    tv, tn, ti = tetra()
    meshes = [{"primitives": [{"attributes": {"POSITION": W.floats(bv, "VEC3"),
                                              "NORMAL": W.floats(bn, "VEC3")},
                               "indices": W.indices(bi, np.uint16), "material": 1}]},
              {"primitives": [{"attributes": {"POSITION": W.floats(tv, "VEC3"),
                                              "NORMAL": W.floats(tn, "VEC3")},
                               "indices": W.indices(ti, np.uint8), "material": 0}]}]
    materials = [{"pbrMetallicRoughness": {"baseColorFactor": [0.3, 0.6, 0.9, 1.0]}},
                 {"emissiveFactor": [1.0, 1.0, 1.0], "extensions": {
                     "KHR_materials_emissive_strength": {"emissiveStrength": 2.5}}}]
    nodes = [{"name": "group", "children": [1], "translation": [0.0, 0.0, -4.0]},
             {"name": "box", "mesh": 0, "translation": [-0.8, 0.0, 0.0]},
             {"name": "tet", "mesh": 1, "translation": [1.0, 0.0, -4.0]},
             {"name": "eye", "camera": 0, "translation": [0.0, 0.5, 1.0]}]
    k = W.floats([0.0, 1.0, 2.0], "SCALAR")
    tg = W.floats([[0.0, 0.0, -4.0], [0.0, 1.0, -4.0], [0.0, 0.0, -6.0]], "VEC3")
    tc = W.floats([[0.0, 0.5, 1.0], [1.0, 0.5, 1.0], [0.0, 1.5, 2.0]], "VEC3")
    rc = W.floats([quat([0, 1, 0], 0.0), quat([0, 1, 0], 0.3), quat([1, 0, 0], -0.2)],
                  "VEC4")
    animations = [{"channels": [{"sampler": 0, "target": {"node": 0, "path": "translation"}},
                                {"sampler": 1, "target": {"node": 3, "path": "translation"}},
                                {"sampler": 2, "target": {"node": 3, "path": "rotation"}}],
                   "samplers": [{"input": k, "output": tg, "interpolation": "LINEAR"},
                                {"input": k, "output": tc, "interpolation": "LINEAR"},
                                {"input": k, "output": rc, "interpolation": "STEP"}]}]
    W.save("cam_group", {"scenes": [{"nodes": [0, 2, 3]}], "nodes": nodes, "meshes": meshes,
                         "materials": materials, "animations": animations,
                         "cameras": [{"type": "perspective",
                                      "perspective": {"yfov": 0.9, "znear": 0.05}}]})


FIXTURES = ("anim_hier", "static_default", "cam_group")
#: fixtures with reference frames, the times at which they are rendered,
#: frames per accumulation and the block size
FRAME_FIXTURES = ("anim_hier",)
FRAME_TIMES = (0.0, 0.8, 1.9)
FRAMES = 2
BLOCK = (64, 32)


def write_fixtures():
    anim_hier()
    static_default()
    cam_group()


def reference_digests():
    from oracle_lib import RefGltf, rdata_digests, reference
    if reference() is None:
        raise SystemExit("oracle/_ref/libref.so is not built (make -C oracle ref)")
    out = {"times": np.array(TIMES, np.float32), "frame_times": np.array(FRAME_TIMES, np.float32),
           "size": np.array([WIDTH, HEIGHT, SPP, FRAMES, *BLOCK], np.int64)}
    for name in FIXTURES:
        g = RefGltf(OUT / f"{name}.gltf", WIDTH, HEIGHT, SPP, BG)
        labels, shas, masks = None, [], []
        for t in TIMES:
            masks.append(g.update(t))
            d = rdata_digests(g.rd, g.ntris, g.nmtls)
            labels = [k for k, _ in d]
            shas.append([v for _, v in d])
            if name in FRAME_FIXTURES and t in FRAME_TIMES:
                key = f"{name}_t{FRAME_TIMES.index(t)}"
                buf, acc, n = g.frames(WIDTH, HEIGHT, FRAMES, BLOCK, single=True)
                out[f"{key}_buf_single"], out[f"{key}_acc_single"] = buf, acc
                out[f"{key}_rays_single"] = np.int64(n)
                pbuf, pacc, _ = g.frames(WIDTH, HEIGHT, FRAMES, BLOCK, single=False)
                diff = np.flatnonzero((pacc != acc).any(axis=2) | (pbuf != buf))
                out[f"{key}_packet_diff_px"] = diff.astype(np.int64)
                print(f"{key}: {n} rays, {len(diff)} pixels differ in the packet frame")
        g.close()
        out[f"{name}_labels"] = np.array(labels)
        out[f"{name}_sha256"] = np.array(shas)
        out[f"{name}_masks"] = np.array(masks, np.int32)
        print(name, "masks", masks)
    np.savez_compressed(HERE / "gltf_ref.npz", **out)


if __name__ == "__main__":
    write_fixtures()
    reference_digests()
