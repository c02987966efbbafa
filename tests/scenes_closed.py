"""Small scenes with closed-form radiance, shared by the oracle (CPU) and
GPU path-tracing tests."""
import ctypes as C

import numpy as np

from nodebleed_amd import Scene

F3 = C.c_float * 3
Q4 = C.c_float * 4


def cube_scene(albedo, bg, w=48, h=48, spp=16):
    """Flat-shaded cube [-1,1]^3 (separate vertices per face) under a uniform
    background, no emitters."""
    def fill(L, rs):
        m = L.rs_material(rs, *albedo, 0)
        assert L.rs_mesh_begin(rs) == 0
        P = [(x, y, z) for z in (-1, 1) for y in (-1, 1) for x in (-1, 1)]
        faces = [(0, 2, 3, 1), (4, 5, 7, 6), (0, 1, 5, 4), (2, 6, 7, 3), (0, 4, 6, 2),
                 (1, 3, 7, 5)]
        for q in faces:
            assert L.rs_quad(rs, *[F3(*P[i]) for i in q], m) == 0
        mesh = L.rs_mesh_end(rs)
        assert L.rs_instance(rs, mesh, F3(0, 0, 0), Q4(0, 0, 0, 1), F3(1, 1, 1), 0) == 0
        assert L.rs_camera(rs, F3(2.6, 2.1, 3.2), F3(0, 0, 0), 45.0, 1.0, 0.0) == 0
        assert L.rs_background(rs, *bg) == 0

    return Scene.from_builder(fill, width=w, height=h, spp=spp)


LIGHT_Y, LIGHT_HALF, LIGHT_LE, FLOOR_ALBEDO = 1.0, 0.25, 10.0, 0.5
TARGET = (0.5, 0.0, 0.0)


def lit_floor_scene(w=4, h=4, spp=1):
    """Diffuse floor y=0 under a square emitter at y=1 facing down; the
    camera looks at TARGET through a 1-degree field of view from below the
    light's height, so it sees only the floor."""
    def fill(L, rs):
        mf = L.rs_material(rs, FLOOR_ALBEDO, FLOOR_ALBEDO, FLOOR_ALBEDO, 0)
        ml = L.rs_material(rs, LIGHT_LE, LIGHT_LE, LIGHT_LE, 0x1000)
        assert L.rs_mesh_begin(rs) == 0
        assert L.rs_quad(rs, F3(-5, 0, -5), F3(5, 0, -5), F3(5, 0, 5), F3(-5, 0, 5), mf) == 0
        a, y = LIGHT_HALF, LIGHT_Y
        assert L.rs_quad(rs, F3(-a, y, -a), F3(a, y, -a), F3(a, y, a), F3(-a, y, a), ml) == 0
        mesh = L.rs_mesh_end(rs)
        assert L.rs_instance(rs, mesh, F3(0, 0, 0), Q4(0, 0, 0, 1), F3(1, 1, 1), 0) == 0
        assert L.rs_camera(rs, F3(2.0, 0.5, 0.0), F3(*TARGET), 1.0, 1.0, 0.0) == 0
        assert L.rs_background(rs, 0.0, 0.0, 0.0) == 0

    return Scene.from_builder(fill, width=w, height=h, spp=spp)


def lit_floor_two_instances(w=4, h=4, spp=1):
    """The lit floor with the floor and the emitter as two instances of two
    meshes (instance 1 = the emitter, centred at the origin of its mesh), so
    the emitter can be moved by its instance transform."""
    def fill(L, rs):
        mf = L.rs_material(rs, FLOOR_ALBEDO, FLOOR_ALBEDO, FLOOR_ALBEDO, 0)
        ml = L.rs_material(rs, LIGHT_LE, LIGHT_LE, LIGHT_LE, 0x1000)
        assert L.rs_mesh_begin(rs) == 0
        assert L.rs_quad(rs, F3(-5, 0, -5), F3(5, 0, -5), F3(5, 0, 5), F3(-5, 0, 5), mf) == 0
        floor = L.rs_mesh_end(rs)
        assert L.rs_mesh_begin(rs) == 0
        a, y = LIGHT_HALF, LIGHT_Y
        assert L.rs_quad(rs, F3(-a, y, -a), F3(a, y, -a), F3(a, y, a), F3(-a, y, a), ml) == 0
        light = L.rs_mesh_end(rs)
        assert L.rs_instance(rs, floor, F3(0, 0, 0), Q4(0, 0, 0, 1), F3(1, 1, 1), 0) == 0
        assert L.rs_instance(rs, light, F3(0, 0, 0), Q4(0, 0, 0, 1), F3(1, 1, 1), 0) == 1
        assert L.rs_camera(rs, F3(2.0, 0.5, 0.0), F3(*TARGET), 1.0, 1.0, 0.0) == 0
        assert L.rs_background(rs, 0.0, 0.0, 0.0) == 0

    return Scene.from_builder(fill, width=w, height=h, spp=spp)


def move_instance(scene, i, shift):
    """Translates instance i of a scene whose instances have no rotation or
    scale, as the application's upd_rinsts does (main.c:117-157): new
    globinv and world box in the host rdata, then rend_prepdynamic."""
    ins, ab = scene.insts(), scene.aabbs()
    g = ins["globinv"][i].reshape(3, 4)
    g[:, 3] -= np.asarray(shift, np.float32)
    for k, key in enumerate(("min", "max")):
        ab[key][i] += np.asarray(shift, np.float32)
    ab["min"][i] -= 1e-5
    ab["max"][i] += 1e-5
    scene.L.rend_prepdynamic(scene.rd)


def lit_floor_radiance_at(light_centre):
    a = LIGHT_HALF
    cx, cy, cz = light_centre
    poly = [(cx - a, cy, cz - a), (cx + a, cy, cz - a), (cx + a, cy, cz + a),
            (cx - a, cy, cz + a)]
    return FLOOR_ALBEDO / np.pi * polygon_irradiance(TARGET, (0, 1, 0), poly, LIGHT_LE)


def polygon_irradiance(x, n, poly, Le):
    """Irradiance at x (normal n) from a Lambertian polygon of radiance Le
    (Lambert's formula: E = Le/2 * |sum_i theta_i n . g_i|, theta_i the angle
    edge i subtends at x, g_i the unit normal of the plane through x and
    edge i)."""
    x = np.asarray(x, np.float64)
    n = np.asarray(n, np.float64)
    u = [np.asarray(p, np.float64) - x for p in poly]
    u = [v / np.linalg.norm(v) for v in u]
    s = 0.0
    for i in range(len(u)):
        a, b = u[i], u[(i + 1) % len(u)]
        g = np.cross(a, b)
        s += np.arccos(np.clip(a @ b, -1, 1)) * (n @ (g / np.linalg.norm(g)))
    return 0.5 * Le * abs(s)


def lit_floor_radiance():
    a, y = LIGHT_HALF, LIGHT_Y
    poly = [(-a, y, -a), (a, y, -a), (a, y, a), (-a, y, a)]
    E = polygon_irradiance(TARGET, (0, 1, 0), poly, LIGHT_LE)
    return FLOOR_ALBEDO / np.pi * E


def cube_coverage(img, albedo):
    """Per-pixel coverage fraction per channel from v = 1 - f (1 - a), with
    the albedo as stored (float32)."""
    a = np.asarray(albedo, np.float32).astype(np.float64)
    return (1.0 - img) / (1.0 - a)


def icosphere(subdiv: int):
    """Unit icosphere: vertices (n x 3) on the sphere and triangles (m x 3),
    each face split in four per level, new vertices projected outwards."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t),
         (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
         (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
         (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    v = [np.array(p, np.float64) / np.linalg.norm(p) for p in v]
    for _ in range(subdiv):
        mid, nf = {}, []

        def m(a, b):
            k = (min(a, b), max(a, b))
            if k not in mid:
                p = v[a] + v[b]
                v.append(p / np.linalg.norm(p))
                mid[k] = len(v) - 1
            return mid[k]
        for a, b, c in f:
            ab, bc, ca = m(a, b), m(b, c), m(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = nf
    return np.array(v), np.array(f, np.int64)


def integrating_sphere(albedo, le, cap_y, subdiv=5, w=16, h=16, spp=1):
    """Closed sphere of radius 1 seen from its centre: triangles whose
    centroid lies above y = cap_y emit le (and reflect nothing), the rest
    are Lambertian with the given albedo.  Returns (scene, emitter area
    fraction of the mesh).  The camera looks down, away from the cap."""
    v, f = icosphere(subdiv)
    f = f[:, [0, 2, 1]]  # wound to face the inside (emitters emit on their front)
    cen = v[f].mean(axis=1)
    emit = cen[:, 1] > cap_y
    e0, e1 = v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]]
    area = 0.5 * np.linalg.norm(np.cross(e0, e1), axis=1)
    frac = float(area[emit].sum() / area.sum())
    vv = np.ascontiguousarray(v, np.float32)
    nn = np.ascontiguousarray(-v / np.linalg.norm(v, axis=1, keepdims=True), np.float32)
    ii = np.ascontiguousarray(f, np.uint32)

    def fill(L, rs):
        mw = L.rs_material(rs, albedo, albedo, albedo, 0)
        ml = L.rs_material(rs, le, le, le, 0x1000)
        mt = np.ascontiguousarray(np.where(emit, ml, mw), np.uint32)
        assert L.rs_mesh_begin(rs) == 0
        assert L.rs_mesh_data(rs, vv.ctypes.data, len(vv), nn.ctypes.data, ii.ctypes.data,
                              len(ii), mt.ctypes.data, 0) == 0
        mesh = L.rs_mesh_end(rs)
        assert L.rs_instance(rs, mesh, F3(0, 0, 0), Q4(0, 0, 0, 1), F3(1, 1, 1), 0) == 0
        assert L.rs_camera(rs, F3(0, 0, 0), F3(0.01, -1, 0.02), 60.0, 1.0, 0.0) == 0
        assert L.rs_background(rs, 0.0, 0.0, 0.0) == 0

    s = Scene.from_builder(fill, width=w, height=h, spp=spp)
    s.emitters = int(emit.sum())
    return s, frac


def integrating_sphere_radiance(albedo, le, frac, depth):
    """Radiance of the sphere's wall estimated with `depth` path vertices,
    each with next-event estimation (emitters seen by BSDF sampling add
    nothing): on a sphere every surface element sees any region in
    proportion to its area, so the wall's radiance is uniform and
    L_D = a Le f sum_{k<D} (a (1 - f))^k."""
    q = albedo * (1.0 - frac)
    return albedo * le * frac * (1.0 - q ** depth) / (1.0 - q)
