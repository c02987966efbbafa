"""nodebleed_amd: Python bindings of librend, the MI355X path-tracing core.

librend (``nodebleed_amd/librend.so``, built by ``python -m
nodebleed_amd.build``) exports the reference renderer's C API (``rend.h``)
plus ray-buffer, scene-construction and timing entry points
(``rend_hip.h``, ``rend_scene.h``).  This module binds them with ctypes for
tests and the benchmark; it adds no compute of its own.  Loading fails
loudly when the library has not been built: there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIBPATH = PKG / "librend.so"

RAY = np.dtype([("ox", "<f4"), ("oy", "<f4"), ("oz", "<f4"), ("tmax", "<f4"),
                ("dx", "<f4"), ("dy", "<f4"), ("dz", "<f4"), ("pad", "<f4")])
HIT = np.dtype([("t", "<f4"), ("u", "<f4"), ("v", "<f4"), ("id", "<u4")])
RTRI = np.dtype([("v0", "<f4", 3), ("p0", "<f4"), ("v1", "<f4", 3), ("p1", "<f4"),
                 ("v2", "<f4", 3), ("p2", "<f4")])
RINST = np.dtype([("globinv", "<f4", 12), ("flags", "<u4"), ("triofs", "<u4"),
                  ("tricnt", "<u4"), ("pad0", "<u4")])
AABB = np.dtype([("min", "<f4", 3), ("p0", "<f4"), ("max", "<f4", 3), ("p1", "<f4")])
#: struct rend_view (rend_hip.h): one camera of a batched render
VIEW = np.dtype([("eye", "<f4", 3), ("target", "<f4", 3), ("vfov", "<f4"),
                 ("focdist", "<f4"), ("focangle", "<f4"), ("w", "<u4"), ("h", "<u4"),
                 ("spp", "<u4")])

REND_MODE_NORMALS = 0
REND_MODE_PATHTRACE = 1
REND_RCP_HOST_TABLE = 0
REND_RCP_IEEE = 1
REND_RCP_XEON_TABLE = 2
REND_TLAS_HOST = 0
REND_TLAS_GPU = 1
REND_BLAS_HOST = 0
REND_BLAS_GPU = 1
INST_ID_BITS = 12
NODE_LEAF = 0x80000000


class RsView(C.Structure):
    _fields_ = [("mtls", C.c_void_p), ("tris", C.c_void_p), ("nrms", C.c_void_p),
                ("insts", C.c_void_p), ("aabbs", C.c_void_p), ("bnodes", C.c_void_p),
                ("acc", C.c_void_p), ("buf", C.c_void_p), ("instcnt", C.c_uint),
                ("tlasofs", C.c_uint), ("width", C.c_uint), ("height", C.c_uint),
                ("spp", C.c_uint), ("samples", C.c_uint), ("rays", C.c_uint),
                ("blknum", C.c_int), ("bgcol", C.c_float * 3), ("cam", C.c_void_p),
                ("camsize", C.c_uint)]


class TraceStats(C.Structure):
    _fields_ = [("nodes", C.c_ulonglong), ("leaves", C.c_ulonglong),
                ("insts", C.c_ulonglong), ("rays", C.c_ulonglong)]


class RendError(RuntimeError):
    pass


_lib = None

_SIGS = {
    "rend_staticinit": (None, [C.c_uint, C.c_uint]),
    "rend_render": (C.c_int, [C.c_void_p]),
    "rend_clraccum": (None, [C.c_void_p]),
    "rend_resaccum": (None, [C.c_void_p, C.c_uint, C.c_uint]),
    "rend_release": (None, [C.c_void_p]),
    "rend_prepstatic": (None, [C.c_void_p]),
    "rend_prepdynamic": (None, [C.c_void_p]),
    "rend_hip_lasterror": (C.c_int, []),
    "rend_hip_errstr": (C.c_char_p, []),
    "rend_hip_rcp_cpu": (C.c_char_p, []),
    "rend_hip_devinfo": (C.c_int, [C.POINTER(C.c_int), C.c_char_p, C.c_uint]),
    "rend_hip_set_rcp": (C.c_int, [C.c_int]),
    "rend_hip_set_tlas_build": (C.c_int, [C.c_int]),
    "rend_hip_set_blas_build": (C.c_int, [C.c_int]),
    "rend_hip_blas_download": (C.c_int, [C.c_void_p, C.c_uint, C.c_void_p, C.c_ulonglong,
                                         C.POINTER(C.c_ulonglong)]),
    "rend_hip_sync_scene": (C.c_int, [C.c_void_p]),
    "rend_hip_tlas_download": (C.c_int, [C.c_void_p, C.c_void_p, C.c_ulonglong,
                                         C.POINTER(C.c_uint)]),
    "rend_hip_set_tuning": (C.c_int, [C.c_uint, C.c_uint]),
    "rend_hip_set_trace_variant": (C.c_int, [C.c_int]),
    "rend_hip_set_step_batch": (C.c_int, [C.c_uint]),
    "rend_hip_trace_lanes": (C.c_int, [C.c_void_p]),
    "rend_hip_set_iter_limit": (C.c_int, [C.c_uint]),
    "rend_hip_get_tuning": (C.c_int, [C.c_void_p]),
    "rend_hip_set_mode": (C.c_int, [C.c_int]),
    "rend_hip_pt_config": (C.c_int, [C.c_uint, C.c_uint, C.c_uint]),
    "rend_hip_upload": (C.c_int, [C.c_void_p]),
    "rend_hip_hazards": (C.c_uint, [C.c_void_p]),
    "rend_hip_arena_bytes": (C.c_longlong, [C.c_void_p]),
    "rend_hip_trace": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint,
                                 C.c_void_p, C.POINTER(TraceStats)]),
    "rend_hip_trace_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint,
                                         C.c_void_p]),
    "rend_hip_occluded": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint,
                                    C.c_void_p]),
    "rend_hip_trace_host": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint,
                                      C.POINTER(TraceStats)]),
    "rend_hip_occluded_host": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_uint]),
    "rend_hip_render_frame": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint)]),
    "rend_hip_download_acc": (C.c_int, [C.c_void_p]),
    "rend_hip_camrays_ref_count": (C.c_ulonglong, [C.c_void_p]),
    "rend_hip_gen_camrays_ref": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint, C.c_void_p]),
    "rend_hip_gen_camrays": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint, C.c_uint,
                                       C.c_uint, C.c_uint, C.c_void_p]),
    "rend_hip_gen_diffuse": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_uint, C.c_uint, C.c_void_p]),
    "rend_hip_stream": (C.c_void_p, []),
    "rend_hip_malloc": (C.c_void_p, [C.c_ulonglong]),
    "rend_hip_free": (C.c_int, [C.c_void_p]),
    "rend_hip_memcpy_htod": (C.c_int, [C.c_void_p, C.c_void_p, C.c_ulonglong]),
    "rend_hip_memcpy_dtoh": (C.c_int, [C.c_void_p, C.c_void_p, C.c_ulonglong]),
    "rend_hip_sync": (C.c_int, []),
    "rend_hip_copy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_ulonglong, C.c_void_p]),
    "rend_hip_event_create": (C.c_void_p, []),
    "rend_hip_event_record": (C.c_int, [C.c_void_p, C.c_void_p]),
    "rend_hip_event_elapsed": (C.c_float, [C.c_void_p, C.c_void_p]),
    "rend_hip_event_destroy": (C.c_int, [C.c_void_p]),
    "rend_hip_pt_graphs": (C.c_int, [C.c_int]),
    "rend_hip_render_views": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint, C.c_uint,
                                        C.c_void_p, C.c_void_p, C.POINTER(C.c_ulonglong)]),
    "rend_hip_autotune": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint, C.c_void_p,
                                    C.POINTER(C.c_float)]),
    "rend_hip_guide_enable": (C.c_int, [C.c_void_p, C.c_float, C.c_uint, C.c_uint,
                                        C.c_float, C.c_uint]),
    "rend_hip_guide_disable": (C.c_int, [C.c_void_p]),
    "rend_hip_guide_set": (C.c_int, [C.c_void_p, C.c_float, C.c_uint]),
    "rend_hip_guide_training": (C.c_int, [C.c_void_p, C.c_uint, C.c_float, C.c_float, C.c_int]),
    "rend_hip_guide_sampling": (C.c_int, [C.c_void_p, C.c_uint, C.c_int]),
    "rend_hip_guide_target": (C.c_int, [C.c_void_p, C.c_int]),
    "rend_hip_guide_objective": (C.c_int, [C.c_void_p, C.c_int]),
    "rend_hip_guide_warmup": (C.c_int, [C.c_void_p, C.c_uint, C.c_uint]),
    "rend_hip_guide_cdf_host": (C.c_int, [C.c_void_p, C.c_uint, C.c_float, C.c_void_p]),
    "rend_hip_guide_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "rend_hip_guide_params": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
    "rend_hip_guide_forward_host": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint,
                                              C.c_void_p]),
    "rend_hip_guide_grad_host": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint, C.c_void_p,
                                           C.POINTER(C.c_float)]),
    "rend_hip_guide_mix_host": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint, C.c_void_p]),
    "rend_hip_guide_train_host": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint, C.c_uint,
                                            C.POINTER(C.c_float)]),
    "rend_hip_guide_sample_host": (C.c_int, [C.c_void_p, C.c_float, C.c_float, C.c_float,
                                             C.c_void_p]),
    "rend_hip_mfma_check": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "rs_new": (C.c_void_p, []),
    "rs_free": (None, [C.c_void_p]),
    "rs_config": (C.c_int, [C.c_void_p, C.c_int, C.c_uint, C.c_float]),
    "rs_commit": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint, C.c_uint, C.c_uint]),
    "rs_material": (C.c_int, [C.c_void_p, C.c_float, C.c_float, C.c_float, C.c_uint]),
    "rs_mesh_begin": (C.c_int, [C.c_void_p]),
    "rs_mesh_end": (C.c_int, [C.c_void_p]),
    "rs_icosphere": (C.c_int, [C.c_void_p, C.c_uint, C.c_float, C.c_float, C.c_float,
                               C.c_float, C.c_float, C.c_float, C.c_uint, C.c_uint]),
    "rs_torus": (C.c_int, [C.c_void_p, C.c_uint, C.c_uint, C.c_float, C.c_float,
                           C.c_float, C.c_float, C.c_uint, C.c_uint]),
    "rs_terrain": (C.c_int, [C.c_void_p, C.c_uint, C.c_uint, C.c_float, C.c_float,
                             C.c_float, C.c_float, C.c_float, C.c_float, C.c_uint,
                             C.c_uint]),
    "rs_quad": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float),
                          C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_uint]),
    "rs_instance": (C.c_int, [C.c_void_p, C.c_uint, C.POINTER(C.c_float),
                              C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_uint]),
    "rs_camera": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float),
                            C.c_float, C.c_float, C.c_float]),
    "rs_set_rcam": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float),
                              C.c_float, C.c_float, C.c_float, C.c_uint, C.c_uint]),
    "rs_background": (C.c_int, [C.c_void_p, C.c_float, C.c_float, C.c_float]),
    "rs_tricount": (C.c_uint, [C.c_void_p]),
    "rs_mtlcount": (C.c_uint, [C.c_void_p]),
    "rs_mesh_data": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint, C.c_void_p, C.c_void_p,
                               C.c_uint, C.c_void_p, C.c_uint]),
    "rs_gltf": (C.c_int, [C.c_void_p, C.c_char_p]),
    "rs_animate": (C.c_int, [C.c_void_p, C.c_void_p, C.c_float]),
    "rs_noise": (C.c_float, [C.c_float, C.c_float, C.c_float, C.c_uint]),
    "rs_instance_glob": (C.c_int, [C.c_void_p, C.c_uint, C.POINTER(C.c_float)]),
    "rs_rdata_view": (None, [C.c_void_p, C.POINTER(RsView)]),
    "rs_rdata_setframe": (None, [C.c_void_p, C.c_int, C.c_uint, C.c_uint, C.c_void_p]),
    "rs_rdata_camera": (None, [C.c_void_p, C.POINTER(C.c_float)]),
    "rs_rdata_new": (C.c_void_p, []),
    "rs_rdata_free": (None, [C.c_void_p]),
    "rs_sizeof_rdata": (C.c_uint, []),
}

#: Every C-ABI symbol the public headers declare (checked by the CPU tests).
EXPORTS = tuple(_SIGS) + ("rend_init", "rend_staticrelease", "mulpos", "muldir")


def lib() -> C.CDLL:
    """Loads librend.so; raises when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIBPATH.exists():
        raise RendError(f"{LIBPATH} is missing: run `python -m nodebleed_amd.build` "
                        "(there is no non-HIP fallback)")
    L = C.CDLL(str(LIBPATH))
    for name, (res, args) in _SIGS.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    L.rend_init.restype = None
    L.rend_init.argtypes = [C.c_void_p, C.c_uint, C.c_uint, C.c_uint, C.c_uint]
    _lib = L
    return L


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        L = lib()
        raise RendError(f"{what}: error {rc}: {L.rend_hip_errstr().decode()}")


def fptr(a) -> C.POINTER(C.c_float):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a.ctypes.data_as(C.POINTER(C.c_float))


def ptr(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)


class Scene:
    """An rdata built from a procedural configuration or an rs builder.

    ``Scene(config=2)`` builds the 1M-triangle terrain, ``Scene(config=1)``
    the sphere in an open box, ``Scene(config=3)`` 4096 instances of 16
    meshes; ``detail`` < 1 shrinks triangle counts for tests.
    """

    def __init__(self, config: int | None = None, seed: int = 1, detail: float = 1.0,
                 width: int = 64, height: int = 64, spp: int = 1, builder=None):
        L = lib()
        self.L = L
        L.rend_staticinit(64, 32)
        self.rs = builder if builder is not None else L.rs_new()
        if config is not None:
            check(L.rs_config(self.rs, config, seed, detail), f"rs_config({config})")
        self.rd = L.rs_rdata_new()
        check(L.rs_commit(self.rs, self.rd, width, height, spp), "rs_commit")
        self.ntris = L.rs_tricount(self.rs)
        self.nmtls = L.rs_mtlcount(self.rs)

    @classmethod
    def from_builder(cls, fill, width=64, height=64, spp=1):
        L = lib()
        rs = L.rs_new()
        fill(L, rs)
        return cls(None, builder=rs, width=width, height=height, spp=spp)

    @classmethod
    def from_gltf(cls, path, width=64, height=64, spp=1, bg=(0.4, 0.4, 0.4)):
        """A glTF 2.0 scene (path.gltf beside path.bin), flattened at time 0."""
        L = lib()
        rs = L.rs_new()
        rc = L.rs_gltf(rs, str(path).encode())
        if rc != 0:
            msg = L.rend_hip_errstr().decode()
            L.rs_free(rs)
            raise RendError(f"rs_gltf({path}): error {rc}: {msg}")
        L.rs_background(rs, *bg)
        return cls(None, builder=rs, width=width, height=height, spp=spp)

    def animate(self, time: float) -> int:
        """Advances the scene graph to `time`; returns the mask of what
        changed (1 instances and TLAS, 2 camera)."""
        rc = self.L.rs_animate(self.rs, self.rd, time)
        if rc < 0:
            check(rc, "rs_animate")
        return rc

    def view(self) -> RsView:
        v = RsView()
        self.L.rs_rdata_view(self.rd, C.byref(v))
        return v

    def _arr(self, p, dtype, n):
        if not p or n == 0:
            return np.zeros(0, dtype)
        buf = (C.c_char * (n * np.dtype(dtype).itemsize)).from_address(p)
        return np.frombuffer(buf, dtype=dtype, count=n)

    def tris(self) -> np.ndarray:
        return self._arr(self.view().tris, RTRI, self.ntris)

    def insts(self) -> np.ndarray:
        v = self.view()
        return self._arr(v.insts, RINST, v.instcnt)

    def aabbs(self) -> np.ndarray:
        v = self.view()
        return self._arr(v.aabbs, AABB, v.instcnt)

    def globs(self) -> np.ndarray:
        n = self.view().instcnt
        out = np.zeros((n, 16), np.float32)
        for i in range(n):
            self.L.rs_instance_glob(self.rs, i, out[i].ctypes.data_as(C.POINTER(C.c_float)))
        return out

    def arena(self, byte_ofs: int, nbytes: int) -> np.ndarray:
        """Raw bytes of the host BVH arena (rdata.bnodes)."""
        return self._arr(self.view().bnodes + byte_ofs, np.uint8, nbytes).copy()

    def hazards(self) -> int:
        return self.L.rend_hip_hazards(self.rd)

    def trace(self, rays: np.ndarray, stats: bool = False):
        rays = np.ascontiguousarray(rays, dtype=RAY)
        hits = np.zeros(len(rays), HIT)
        st = TraceStats()
        check(self.L.rend_hip_trace_host(self.rd, ptr(rays), ptr(hits), len(rays),
                                         C.byref(st) if stats else None), "trace")
        return (hits, st) if stats else hits

    def occluded(self, rays: np.ndarray) -> np.ndarray:
        rays = np.ascontiguousarray(rays, dtype=RAY)
        occ = np.zeros(len(rays), np.uint8)
        check(self.L.rend_hip_occluded_host(self.rd, ptr(rays), ptr(occ), len(rays)),
              "occluded")
        return occ

    def camera(self) -> np.ndarray:
        """eye, ri, up, fwd, 1/w, 1/h, aspect, focdist, fovfdist, focrad."""
        out = np.zeros(18, np.float32)
        self.L.rs_rdata_camera(self.rd, out.ctypes.data_as(C.POINTER(C.c_float)))
        return out

    def render(self, frames: int = 1, mode: int = REND_MODE_PATHTRACE, maxdepth: int = 4,
               rr_depth: int = 3, batch: int = 0, block=(1, 1)) -> tuple[np.ndarray, int]:
        """Renders `frames` frames through the frame protocol of rend_render
        and returns (mean radiance image h x w x 3, rays traced)."""
        L = self.L
        v = self.view()
        w, h = v.width, v.height
        L.rend_staticinit(*block)
        check(L.rend_hip_set_mode(mode), "mode")
        check(L.rend_hip_pt_config(maxdepth, rr_depth, batch), "pt config")
        L.rend_clraccum(self.rd)
        buf = np.zeros(w * h, np.uint32)
        rays, samples = 0, 0
        for _ in range(frames):
            L.rs_rdata_setframe(self.rd, 0, 0, samples, buf.ctypes.data)
            n = C.c_uint(0)
            check(L.rend_hip_render_frame(self.rd, C.byref(n)), "render")
            rays += n.value
            samples += v.spp
        check(L.rend_hip_download_acc(self.rd), "download")
        acc = self._arr(self.view().acc, np.float32, w * h * 3).reshape(h, w, 3)
        self.argb = buf.reshape(h, w)
        return acc / max(samples, 1), rays

    def render_views(self, views: np.ndarray, sample_base: int = 0, maxdepth: int = 4,
                     rr_depth: int = 3, batch: int = 0, calls: int = 1):
        """Path traces the VIEW records in one batched launch sequence per
        call (sample_base advancing by each view's spp between calls) and
        returns ([mean radiance h x w x 3 per view], rays traced)."""
        L = self.L
        check(L.rend_hip_pt_config(maxdepth, rr_depth, batch), "pt config")
        views = np.ascontiguousarray(views, VIEW)
        sizes = [int(v["w"]) * int(v["h"]) * 3 for v in views]
        total = sum(sizes)
        d_acc = L.rend_hip_malloc(total * 4)
        if not d_acc:
            raise RendError(L.rend_hip_lasterror(), "accumulator allocation")
        try:
            zero = np.zeros(total, np.float32)
            check(L.rend_hip_memcpy_htod(d_acc, zero.ctypes.data, total * 4), "clear")
            rays = 0
            base = sample_base
            for _ in range(calls):
                n = C.c_ulonglong(0)
                check(L.rend_hip_render_views(self.rd, views.ctypes.data, len(views), base,
                                              d_acc, None, C.byref(n)), "render_views")
                rays += n.value
                base += int(views["spp"].max())
            check(L.rend_hip_sync(), "sync")
            acc = np.zeros(total, np.float32)
            check(L.rend_hip_memcpy_dtoh(acc.ctypes.data, d_acc, total * 4), "download")
        finally:
            L.rend_hip_free(d_acc)
        out, o = [], 0
        for v, sz in zip(views, sizes):
            img = acc[o:o + sz].reshape(int(v["h"]), int(v["w"]), 3)
            out.append(img / float(int(v["spp"]) * calls))
            o += sz
        return out, rays

    def close(self):
        if self.rd:
            self.L.rend_release(self.rd)
            self.L.rs_rdata_free(self.rd)
            self.rd = None
        if self.rs:
            self.L.rs_free(self.rs)
            self.rs = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def uneven_views(seed: int = 1, n: int = 32) -> np.ndarray:
    """The batched uneven workload: n cameras over the terrain, resolutions
    drawn from {320x240, 640x480, 1280x720, 1920x1080, 3840x2160}, spp
    uniform in 4..64, eyes on a ring above the terrain (VIEW records)."""
    rng = np.random.default_rng(seed)
    res = [(320, 240), (640, 480), (1280, 720), (1920, 1080), (3840, 2160)]
    v = np.zeros(n, VIEW)
    for i in range(n):
        w, h = res[rng.integers(len(res))]
        ang = rng.uniform(0, 2 * np.pi)
        r = rng.uniform(0.3, 0.9)
        eye = (r * np.cos(ang), rng.uniform(0.3, 0.9), 0.5 * r * np.sin(ang))
        tgt = (rng.uniform(-0.3, 0.3), 0.0, rng.uniform(-0.15, 0.15))
        v[i] = (eye, tgt, rng.uniform(30, 60), 1.0, 0.0, w, h, int(rng.integers(4, 65)))
    return v


def make_rays(o: np.ndarray, d: np.ndarray, tmax: float | np.ndarray = 3.4028235e38) -> np.ndarray:
    r = np.zeros(len(o), RAY)
    r["ox"], r["oy"], r["oz"] = o[:, 0], o[:, 1], o[:, 2]
    r["dx"], r["dy"], r["dz"] = d[:, 0], d[:, 1], d[:, 2]
    r["tmax"] = tmax
    return r
