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
