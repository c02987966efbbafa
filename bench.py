"""Closest-hit traversal benchmark on one MI355X.

Workload (the headline configuration): a single BLAS of 1,000,000 triangles
(seeded displaced terrain, 1000 x 500 cells, height amplitude 0.05), one
instance, 1920x1080 jittered primary rays and one cosine-hemisphere diffuse
ray per primary hit.  Rays are generated on the GPU, so inputs are resident
in HBM before timing starts.  A step is one closest-hit launch over the
2,073,600 incoherent (diffuse) rays; the primary-ray launch is timed beside
it and reported as an extra field.  The same run then measures the
path-guiding target on the occluded-lights scene (1920x1080): the wall time
guided path tracing needs to reach plain path tracing's relMSE after 1,024
frames (16,384 spp), over plain's time, added to the line as
"guided_vs_plain" (about four minutes; --no-guide leaves it out).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-seconds S]

Prints one JSON line.  Only one GPU is supported: --gpus above 1 exits with
a message.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from nodebleed_amd import HIT, RAY, Scene, TraceStats, check, lib, uneven_views  # noqa: E402

W, H = 1920, 1080
PEAK_HBM_GBS = 8000.0
# Bytes one lane reads per visit in the device layout (nodebleed_amd/csrc/dlayout.h):
# node = 6x8 bounds + 8 child words + 1 rank word, leaf = 3 vertices/edges x 4
# triangles + 4 ids, instance = 12-float inverse + 8-byte BLAS offset; each ray
# reads 32 B and writes 16 B.
S_NODE, S_LEAF, S_INST, S_RAY = 228, 160, 56, 48


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# MI355X: 256 CUs x 4 SIMDs; a wave64 VALU instruction occupies its SIMD for
# 4 cycles; engine clock 2.4 GHz (peak).
SIMDS, VALU_CYCLES, CLOCK_HZ = 1024, 4, 2.4e9
HOST_SHARE = 16
PMC_PROFILE = "profiles/trace_terrain_diffuse_pmc.md"


def visit_bytes(st: TraceStats, nrays: int) -> int:
    """Bytes the traversal's visits read plus ray and hit records: what the
    caches deliver to the lanes (most of it from L1/L2/MALL)."""
    return nrays * S_RAY + st.nodes * S_NODE + st.leaves * S_LEAF + st.insts * S_INST


def compulsory_bytes(scene: Scene, nrays: int) -> int:
    """The least a launch must move across HBM when nothing stays cached
    between launches: every ray record read and every hit written once, and
    the scene's device tree (BLAS + TLAS arena) and instance records read
    once."""
    L = lib()
    arena = int(L.rend_hip_arena_bytes(scene.rd))
    return nrays * S_RAY + max(arena, 0) + 64 * int(scene.view().instcnt)


def profile_counters():
    """Per-launch figures of the committed rocprofv3 --pmc summary of this
    workload: HBM bytes (FETCH_SIZE + WRITE_SIZE scaled by a copy kernel's
    reading) and SQ_INSTS_VALU; None where absent."""
    out = {"hbm_bytes": None, "valu_insts": None}
    try:
        with open(os.path.join(ROOT, PMC_PROFILE)) as f:
            for line in f:
                t = line.strip()
                if t.startswith("hbm_bytes_per_launch:"):
                    out["hbm_bytes"] = int(t.split(":", 1)[1].strip().split()[0])
                elif t.startswith("| SQ_INSTS_VALU |"):
                    out["valu_insts"] = int(t.split("|")[2].strip().replace(",", ""))
    except (OSError, ValueError, IndexError):
        pass
    return out


def timed_launches(L, scene, d_rays, d_hits, n, k):
    stream = L.rend_hip_stream()
    e0, e1 = L.rend_hip_event_create(), L.rend_hip_event_create()
    check(L.rend_hip_sync(), "sync")
    check(L.rend_hip_event_record(e0, stream), "event")
    for _ in range(k):
        check(L.rend_hip_trace(scene.rd, d_rays, d_hits, n, stream, None), "trace")
    check(L.rend_hip_event_record(e1, stream), "event")
    ms = L.rend_hip_event_elapsed(e0, e1)
    check(L.rend_hip_sync(), "sync")
    L.rend_hip_event_destroy(e0)
    L.rend_hip_event_destroy(e1)
    return ms


def cpu_baseline(scene: Scene, rays: np.ndarray, seconds: float) -> dict:
    """The reference's single-ray traversal (or the oracle port) on host
    cores, over a bounded prefix of the same diffuse rays."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    online = os.sysconf("SC_NPROCESSORS_ONLN") if hasattr(os, "sysconf") else os.cpu_count()
    allowed = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else online
    # the reference runs one thread per online core (main.c:386); a one-GPU
    # share of the GPU host is 16 threads, whatever the host reports
    threads = max(1, min(int(allowed), HOST_SHARE))
    ref = O.reference(fast=True)
    if ref is not None:
        rs = O.RefScene(scene, fast=True)
        run = lambda r: rs.trace(r, threads=threads)  # noqa: E731
        kind, what = "reference", "reference intersect_tlas (speed build)"
    else:
        rs = None
        run = lambda r: O.or_trace(scene, r, threads=threads)  # noqa: E731
        kind, what = "port", "C restatement of intersect_tlas"
    probe = rays[:20000]
    t = time.perf_counter()
    run(probe)
    dt = max(time.perf_counter() - t, 1e-6)
    n = int(min(len(rays), max(20000, len(probe) / dt * seconds)))
    # whole passes over the (prefix of the) buffer until `seconds` of CPU
    # work have elapsed (the probe's per-ray time is only an estimate)
    passes = 0
    t = time.perf_counter()
    while passes == 0 or time.perf_counter() - t < seconds:
        run(rays[:n])
        passes += 1
    dt = time.perf_counter() - t
    n *= passes
    if rs is not None:
        rs.close()
    cpu = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": round(n / dt / 1e6, 4), "unit": "Mrays/s", "cores": threads,
            "kind": kind, "online_cores": int(online), "allowed_cores": int(allowed),
            "thread_policy": f"one per allowed core, capped at the {HOST_SHARE}-thread "
                             "host share of one GPU",
            "sample": f"{what}, {n} rays ({passes} pass(es) over the first "
                      f"{n // passes} of {len(rays)} diffuse rays), {threads} threads "
                      f"on {cpu}, {dt:.1f} s"}


def emit(metric, value, unit, steps, warmup, ms, config, higher=True, dtype="f32",
         extra=None):
    out = {"metric": metric, "value": round(value, 3), "unit": unit, "n_gpus": 1,
           "steps": steps, "warmup": warmup, "ms_per_step": round(ms, 4),
           "higher_is_better": higher, "scaling": "weak", "vs_baseline": None,
           "dtype": dtype, "data": "synthetic (seeded procedural scenes)", "config": config}
    if extra:
        out.update(extra)
    print(json.dumps(out), flush=True)


def frames(L, scene, k):
    """k path-traced frames through rend_hip_render_frame; (seconds, rays)."""
    rays = 0
    samples = scene.view().samples
    spp = scene.view().spp
    t = time.perf_counter()
    for _ in range(k):
        L.rs_rdata_setframe(scene.rd, 0, 0, samples, None)
        n = C.c_uint(0)
        check(L.rend_hip_render_frame(scene.rd, C.byref(n)), "render")
        rays += n.value
        samples += spp
    return time.perf_counter() - t, rays


def bench_pt(a):
    """Instanced scene: 16 meshes of 50k-500k triangles, 4,096 instances
    (4,095 meshes + light), 1920x1080, 64 spp per frame, up to 8 segments,
    Russian roulette from depth 3.  Step = one frame."""
    from nodebleed_amd import REND_MODE_PATHTRACE
    L = lib()
    t0 = time.perf_counter()
    scene = Scene(config=3, seed=a.seed, detail=1.0, width=W, height=H, spp=64)
    log(f"scene: {scene.ntris} tris, {scene.view().instcnt} instances, "
        f"build {time.perf_counter() - t0:.1f} s")
    L.rend_staticinit(32, 8)
    check(L.rend_hip_set_mode(REND_MODE_PATHTRACE), "mode")
    check(L.rend_hip_pt_config(8, 3, 0), "pt config")
    L.rend_clraccum(scene.rd)
    frames(L, scene, a.warmup)
    log("warmup done")
    sec, rays = frames(L, scene, a.steps)
    ms = sec * 1e3 / a.steps
    log(f"pt: {ms:.1f} ms/frame, {rays / a.steps / 1e6:.1f} M rays/frame")
    emit("path-traced Mrays/s (extension + shadow rays)", rays / sec / 1e6, "Mrays/s",
         a.steps, a.warmup, ms,
         {"workload": "path tracing, 16 meshes 50k-500k tris x 4,096 instances, 1920x1080, "
                      "64 spp/frame, max depth 8, RR after depth 3",
          "tris_unique": int(scene.ntris), "rays_per_frame": rays // a.steps,
          "samples_per_s": round(W * H * 64 / (ms * 1e-3) / 1e6, 2)})
    scene.close()
    return 0


def bench_tlas(a):
    """Per-frame TLAS rebuild of the instanced scene (4,096 instances over 16
    meshes): rend_prepdynamic + upload with the reference's host SAH build
    (build_bvh8, repack, upload) against the device LBVH build, and the
    closest-hit throughput of 1920x1080 diffuse rays through either tree.
    Step = one rebuild."""
    from nodebleed_amd import REND_TLAS_GPU, REND_TLAS_HOST
    L = lib()
    scene = Scene(config=3, seed=a.seed, detail=1.0, width=W, height=H, spp=1)
    ninst = scene.view().instcnt
    log(f"scene: {scene.ntris} tris, {ninst} instances")
    n = W * H
    bufs = [L.rend_hip_malloc(n * RAY.itemsize), L.rend_hip_malloc(n * HIT.itemsize),
            L.rend_hip_malloc(n * RAY.itemsize), L.rend_hip_malloc(n * HIT.itemsize)]
    d_prim, d_phit, d_diff, d_dhit = bufs
    stream = L.rend_hip_stream()
    check(L.rend_hip_gen_camrays(scene.rd, d_prim, W, H, 1, a.seed, stream), "camrays")
    check(L.rend_hip_trace(scene.rd, d_prim, d_phit, n, stream, None), "primary trace")
    check(L.rend_hip_gen_diffuse(scene.rd, d_prim, d_phit, d_diff, n, a.seed, stream),
          "diffuse rays")
    check(L.rend_hip_sync(), "sync")
    res = {}
    try:
        for name, mode in (("host", REND_TLAS_HOST), ("device", REND_TLAS_GPU)):
            check(L.rend_hip_set_tlas_build(mode), "tlas mode")
            for _ in range(max(1, a.warmup)):
                L.rend_prepdynamic(C.c_void_p(scene.rd))
                check(L.rend_hip_sync_scene(scene.rd), "rebuild")
            t = time.perf_counter()
            for _ in range(a.steps):
                L.rend_prepdynamic(C.c_void_p(scene.rd))
                check(L.rend_hip_sync_scene(scene.rd), "rebuild")
            ms = (time.perf_counter() - t) * 1e3 / a.steps
            st = TraceStats()
            check(L.rend_hip_trace(scene.rd, d_diff, d_dhit, n, stream, C.byref(st)), "count")
            timed_launches(L, scene, d_diff, d_dhit, n, 2)
            tr = timed_launches(L, scene, d_diff, d_dhit, n, a.steps) / a.steps
            res[name] = {"rebuild_ms": round(ms, 4), "diffuse_trace_ms": round(tr, 4),
                         "diffuse_mrays_s": round(n / tr / 1e3, 1),
                         "visits_per_ray": {"nodes": round(st.nodes / n, 3),
                                            "leaves": round(st.leaves / n, 3),
                                            "instances": round(st.insts / n, 3)}}
            log(f"{name} TLAS: rebuild {ms:.3f} ms, diffuse trace {tr:.3f} ms")
    finally:
        L.rend_hip_set_tlas_build(REND_TLAS_HOST)
        for p in bufs:
            L.rend_hip_free(p)
        scene.close()
    dev = res["device"]["rebuild_ms"]
    emit("TLAS rebuild ms (rend_prepdynamic + upload), device build", dev, "ms",
         a.steps, a.warmup, dev,
         {"workload": f"TLAS over {ninst} instances of 16 meshes (50k-500k tris), rebuilt "
                      "per frame; diffuse rays 1920x1080 traced through each tree",
          "host": res["host"], "device": res["device"],
          "speedup": round(res["host"]["rebuild_ms"] / dev, 1)}, higher=False)
    return 0


def bench_blas(a):
    """Scene load of the 1M-triangle terrain (rs_commit: rend_prepstatic,
    rend_prepdynamic; then the upload) with the reference's host SAH BLAS
    build (build_bvh8, repack, upload) against the device LBVH build, and
    the closest-hit throughput of 1920x1080 diffuse rays through either
    tree.  Step = one traversal launch (the build is timed once per tree)."""
    from nodebleed_amd import REND_BLAS_GPU, REND_BLAS_HOST, RsView
    L = lib()
    n = W * H
    res = {}
    try:
        for name, mode in (("host", REND_BLAS_HOST), ("device", REND_BLAS_GPU)):
            check(L.rend_hip_set_blas_build(mode), "blas mode")
            L.rend_staticinit(64, 32)
            rs = L.rs_new()
            check(L.rs_config(rs, 2, a.seed, 1.0), "scene")
            scene = Scene(None, builder=rs, width=W, height=H, spp=1)
            # time a fresh commit of the same builder: rend_prepstatic +
            # rend_prepdynamic, then the upload (and, for device, the build)
            rd = L.rs_rdata_new()
            check(L.rs_commit(rs, rd, W, H, 1), "warm commit")
            check(L.rend_hip_sync_scene(rd), "warm upload")
            L.rend_release(rd)
            L.rs_rdata_free(rd)
            rd = L.rs_rdata_new()
            t = time.perf_counter()
            check(L.rs_commit(rs, rd, W, H, 1), "commit")
            tc = time.perf_counter() - t
            check(L.rend_hip_sync_scene(rd), "upload")
            tl = time.perf_counter() - t
            L.rend_release(rd)
            L.rs_rdata_free(rd)
            bufs = [L.rend_hip_malloc(n * RAY.itemsize), L.rend_hip_malloc(n * HIT.itemsize),
                    L.rend_hip_malloc(n * RAY.itemsize), L.rend_hip_malloc(n * HIT.itemsize)]
            d_prim, d_phit, d_diff, d_dhit = bufs
            stream = L.rend_hip_stream()
            check(L.rend_hip_gen_camrays(scene.rd, d_prim, W, H, 1, a.seed, stream), "camrays")
            check(L.rend_hip_trace(scene.rd, d_prim, d_phit, n, stream, None), "primary")
            check(L.rend_hip_gen_diffuse(scene.rd, d_prim, d_phit, d_diff, n, a.seed, stream),
                  "diffuse rays")
            st = TraceStats()
            check(L.rend_hip_trace(scene.rd, d_diff, d_dhit, n, stream, C.byref(st)), "count")
            timed_launches(L, scene, d_diff, d_dhit, n, max(1, a.warmup))
            tr = timed_launches(L, scene, d_diff, d_dhit, n, a.steps) / a.steps
            res[name] = {"commit_s": round(tc, 4), "load_s": round(tl, 4),
                         "diffuse_trace_ms": round(tr, 4),
                         "diffuse_mrays_s": round(n / tr / 1e3, 1),
                         "visits_per_ray": {"nodes": round(st.nodes / n, 3),
                                            "leaves": round(st.leaves / n, 3)}}
            log(f"{name} BLAS: commit {tc * 1e3:.1f} ms, commit + upload {tl * 1e3:.1f} ms, "
                f"diffuse trace {tr:.3f} ms, nodes/ray {st.nodes / n:.2f}")
            for p in bufs:
                L.rend_hip_free(p)
            scene.close()
    finally:
        L.rend_hip_set_blas_build(REND_BLAS_HOST)
    dev = res["device"]["load_s"] * 1e3
    emit("scene load ms (rs_commit + upload), device BLAS build", dev, "ms", a.steps, a.warmup,
         res["device"]["diffuse_trace_ms"],
         {"workload": "1M-triangle displaced terrain: BLAS build + TLAS + upload, then "
                      "1920x1080 diffuse rays traced through the tree",
          "host": res["host"], "device": res["device"],
          "speedup": round(res["host"]["load_s"] * 1e3 / dev, 1)}, higher=False)
    return 0


def bench_views(a):
    """32 uneven cameras over the 1M-triangle terrain in one batched,
    graph-captured launch sequence; launch shape from the autotuner (run on
    the same views at 1 spp).  Step = one call rendering every view once."""
    L = lib()
    scene = Scene(config=2, seed=a.seed, detail=1.0, width=64, height=64)
    views = uneven_views(a.seed)
    paths = int((views["w"].astype(np.int64) * views["h"] * views["spp"]).sum())
    px = int((views["w"].astype(np.int64) * views["h"]).sum())
    log(f"views: {len(views)} cameras, {px / 1e6:.1f} M pixels, {paths / 1e6:.0f} M paths")
    check(L.rend_hip_pt_config(4, 3, 0), "pt config")
    d_acc = L.rend_hip_malloc(px * 12)
    if not d_acc:
        raise SystemExit("accumulator allocation failed")
    proxy = views.copy()
    proxy["spp"] = 1
    out = (C.c_float * 5)()
    t = time.perf_counter()
    check(L.rend_hip_autotune(scene.rd, proxy.ctypes.data, len(proxy), d_acc, out), "autotune")
    log(f"autotune {time.perf_counter() - t:.1f} s: batch {int(out[0])} refill {int(out[1])} "
        f"variant {int(out[4])}: {out[2]:.1f} ms vs default {out[3]:.1f} ms (1-spp proxy)")
    n = C.c_ulonglong(0)
    for i in range(a.warmup):
        check(L.rend_hip_render_views(scene.rd, views.ctypes.data, len(views), 1000 * i, d_acc,
                                      None, C.byref(n)), "views")
    check(L.rend_hip_sync(), "sync")
    log("warmup done")
    stream = L.rend_hip_stream()
    e0, e1 = L.rend_hip_event_create(), L.rend_hip_event_create()
    rays = 0
    check(L.rend_hip_event_record(e0, stream), "event")
    for i in range(a.steps):
        check(L.rend_hip_render_views(scene.rd, views.ctypes.data, len(views), 100000 + 100 * i,
                                      d_acc, None, C.byref(n)), "views")
        rays += n.value
    check(L.rend_hip_event_record(e1, stream), "event")
    check(L.rend_hip_sync(), "sync")
    ms = L.rend_hip_event_elapsed(e0, e1) / a.steps
    L.rend_hip_free(d_acc)
    log(f"views: {ms:.1f} ms per call, {rays / a.steps / 1e6:.0f} M rays per call")
    emit("batched path-traced Mrays/s (32 uneven views)", rays / a.steps / (ms * 1e-3) / 1e6,
         "Mrays/s", a.steps, a.warmup, ms,
         {"workload": "32 cameras over the 1M-tri terrain, resolutions from "
                      "{320x240..3840x2160}, spp U{4..64}, max depth 4, RR after 3, one "
                      "graph-captured batched launch sequence per step",
          "pixels": px, "paths_per_step": paths, "rays_per_step": rays // a.steps,
          "autotune": {"batch": int(out[0]), "refill": int(out[1]), "variant": int(out[4]),
                       "proxy_ms_best": round(out[2], 2), "proxy_ms_default": round(out[3], 2)}})
    scene.close()
    return 0


def rel_mse(x, r):
    """Relative MSE of an image against a reference: sum (x - r)^2 / sum r^2
    over pixels and channels."""
    x = np.asarray(x, np.float64)
    r = np.asarray(r, np.float64)
    return float(((x - r) ** 2).sum() / (r ** 2).sum())


def guide_equal_error(a):
    """Equal-error wall time on the occluded-lights scene: plain path tracing
    runs `guide_frames` frames (time T_p, relMSE E_p against a high-spp plain
    GPU reference); guided path tracing (trained online from its own paths,
    starting from an untrained network) runs until its relMSE first reaches
    E_p (time T_g, checked every few frames with the check itself excluded
    from the clock).  Returns the result record; ratio = T_g / T_p."""
    from nodebleed_amd import REND_MODE_PATHTRACE
    L = lib()
    # Guided frames launch their kernels directly (the network changes
    # between frames); plain frames would replay a captured graph, which at
    # this frame size is ~3 % slower than direct launches.  Both are timed
    # with direct launches, so plain runs at its faster setting.
    check(L.rend_hip_pt_graphs(0), "graphs")
    w, h = (W, H) if not a.guide_small else (480, 270)
    spf = a.guide_spf
    scene = Scene(config=5, seed=a.seed, detail=1.0, width=w, height=h, spp=spf)
    L.rend_staticinit(32, 8)
    check(L.rend_hip_set_mode(REND_MODE_PATHTRACE), "mode")
    maxdepth, rr = 6, 3
    check(L.rend_hip_pt_config(maxdepth, rr, 0), "pt config")
    cfg = (a.guide_mix, a.guide_stride, a.guide_steps, a.guide_lr)

    def download(nspp):
        check(L.rend_hip_download_acc(scene.rd), "download")
        v = scene.view()
        acc = scene._arr(v.acc, np.float32, w * h * 3).reshape(h, w, 3)
        return acc.astype(np.float64) / nspp

    def run(nframes, guided, sample0=0, ref=None, check_every=0, stop_at=None, tmax=None,
            tag=""):
        L.rend_clraccum(scene.rd)
        if guided:
            # a fresh network; the first frames train harder
            check(L.rend_hip_guide_enable(scene.rd, *cfg, 7), "guide")
            check(L.rend_hip_guide_training(scene.rd, a.guide_chunk, cfg[3], a.guide_lr_mlp, 1),
                  "schedule")
            check(L.rend_hip_guide_warmup(scene.rd, a.guide_warm_frames, a.guide_warm_steps),
                  "warm-up")
            check(L.rend_hip_guide_target(scene.rd, a.guide_target), "target")
        else:
            L.rend_hip_guide_disable(scene.rd)
        curve, el, samples, k = [], 0.0, sample0, 0
        while k < nframes:
            t = time.perf_counter()
            L.rs_rdata_setframe(scene.rd, 0, 0, samples, None)
            nr = C.c_uint(0)
            check(L.rend_hip_render_frame(scene.rd, C.byref(nr)), "render")
            el += time.perf_counter() - t
            samples += spf
            k += 1
            # near the target the error is checked every frame, so the time
            # reported is that of the first frame that reaches it
            near = stop_at is not None and curve and curve[-1][2] <= 1.1 * stop_at
            if check_every and (k % check_every == 0 or near):
                err = rel_mse(download(k * spf), ref)
                curve.append((round(el, 3), k * spf, err))
                log(f"  {tag} {k * spf} spp {el:.2f} s relMSE {err:.4f}")
                if stop_at is not None and err <= stop_at:
                    break
                if tmax is not None and el > tmax:
                    break
            elif k % 256 == 0:
                log(f"  {tag} frame {k}/{nframes} {el:.1f} s")
        return el, download(k * spf), curve

    if a.guide_profile_frames:
        # guided (or, with --guide-profile-plain, plain) frames only, for a
        # profile of their kernels; no error protocol
        guided = not a.guide_profile_plain
        tag = "guided" if guided else "plain"
        el, _, _ = run(a.guide_profile_frames, guided, tag=tag)
        log(f"{tag}: {a.guide_profile_frames} frames in {el:.2f} s "
            f"({el * 1e3 / a.guide_profile_frames:.1f} ms per frame)")
        L.rend_hip_guide_disable(scene.rd)
        scene.close()
        return {"profile_frames": a.guide_profile_frames, "seconds": round(el, 3)}
    t = time.perf_counter()
    _, ref, _ = run(a.guide_ref_frames, False, sample0=1 << 24, tag="reference")
    log(f"reference: {a.guide_ref_frames * spf} spp plain in {time.perf_counter() - t:.1f} s")
    every = max(1, a.guide_frames // 8)
    tp, plain, cp = run(a.guide_frames, False, ref=ref, check_every=every, tag="plain")
    err_p = rel_mse(plain, ref)
    log(f"plain: {a.guide_frames * spf} spp in {tp:.2f} s, relMSE {err_p:.4f}")
    tg, guided, cg = run(1 << 30, True, ref=ref, check_every=max(1, min(8, every // 8)),
                         stop_at=err_p,
                         tmax=3.0 * tp, tag="guided")
    err_g = rel_mse(guided, ref)
    info = (C.c_float * 4)()
    L.rend_hip_guide_info(scene.rd, info)
    reached = err_g <= err_p
    ratio = tg / tp if reached else float("nan")
    log(f"guided: relMSE {err_g:.4f} after {tg:.2f} s ({'reached' if reached else 'not reached'}"
        f" plain's {err_p:.4f}); ratio {ratio:.3f}")
    L.rend_hip_guide_disable(scene.rd)
    scene.close()
    check(L.rend_hip_pt_graphs(0 if a.no_graphs else 1), "graphs")
    return {"ratio": round(ratio, 4) if reached else None,
            "launch": "direct kernel launches for plain and guided frames",
            "workload": f"occluded-lights terrain (1M tris, 8 upward emitters capping chimneys "
                        f"light ceiling spots, terrain lit indirectly), {w}x{h}, {spf} spp/frame, "
                        f"max depth {maxdepth}, RR after {rr}; reference "
                        f"{a.guide_ref_frames * spf} spp plain (GPU)",
            "plain": {"seconds": round(tp, 3), "frames": a.guide_frames,
                      "spp": a.guide_frames * spf, "relmse": err_p, "curve": cp},
            "guided": {"seconds": round(tg, 3), "frames": cg[-1][1] // spf if cg else None,
                       "spp": cg[-1][1] if cg else None, "relmse": err_g, "reached": reached,
                       "curve": cg, "adam_steps": int(info[2])},
            "guide": {"mix": cfg[0], "train_stride": cfg[1], "steps_per_frame": cfg[2],
                      "target": "next-vertex direct light" if a.guide_target else
                                "path radiance after the vertex",
                      "warm_frames": a.guide_warm_frames,
                      "warm_steps_per_frame": a.guide_warm_steps,
                      "records_per_step": a.guide_chunk,
                      "lr_grid": cfg[3], "lr_mlp": a.guide_lr_mlp}}


def bench_guide(a):
    res = guide_equal_error(a)
    if a.guide_profile_frames:
        return 0
    emit("guided/plain wall time to equal relMSE", res["ratio"] if res["ratio"] is not None
         else float("nan"), "ratio", a.guide_frames, 0,
         res["plain"]["seconds"] * 1e3 / a.guide_frames,
         {k: v for k, v in res.items() if k != "ratio"}, higher=False)
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # defaults: the headline (trace) line times 100 launches (~36 ms of GPU
    # time) after 10 warm-up ones, since with 20 after 3 its mean still
    # includes launches at lower clocks (~3 % slower); the other modes, whose
    # steps are whole frames or calls, time 20 after 3
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--sweep", action="store_true",
                    help="time traversal launch shapes (blocks/CU x refill) on stderr")
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MIB",
                    help="after timing, run one plain copy kernel of MIB MiB "
                         "(reference for hardware byte counters)")
    ap.add_argument("--mode", default="trace",
                    choices=("trace", "pt", "views", "guide", "tlas", "blas"),
                    help="trace: closest-hit microbench (the headline line); pt: "
                         "instanced path tracing; views: 32 uneven cameras in one batched "
                         "launch; guide: equal-time guided vs plain path tracing; tlas: "
                         "per-frame TLAS rebuild, host SAH vs device LBVH")
    ap.add_argument("--guide-frames", type=int, default=1024,
                    help="plain frames (16 spp each) whose error guided must reach")
    ap.add_argument("--guide-ref-frames", type=int, default=4096)
    ap.add_argument("--guide-spf", type=int, default=16, help="samples per pixel per frame")
    ap.add_argument("--guide-mix", type=float, default=0.5)
    ap.add_argument("--guide-stride", type=int, default=32,
                    help="record a training sample at every n-th path")
    ap.add_argument("--guide-steps", type=int, default=8)
    ap.add_argument("--guide-lr", type=float, default=1e-2, help="hash-grid learning rate")
    ap.add_argument("--guide-lr-mlp", type=float, default=1e-3, help="MLP learning rate")
    ap.add_argument("--guide-warm-frames", type=int, default=16)
    ap.add_argument("--guide-chunk", type=int, default=4096,
                    help="training records per Adam step")
    ap.add_argument("--guide-warm-steps", type=int, default=64)
    ap.add_argument("--guide-target", type=int, default=1,
                    help="training target: 1 next vertex's direct light, 0 path radiance")
    ap.add_argument("--guide-small", action="store_true", help="480x270 instead of 1920x1080")
    ap.add_argument("--guide-profile-frames", type=int, default=0,
                    help="--mode guide: only render this many guided frames (for profiling)")
    ap.add_argument("--guide-profile-plain", action="store_true",
                    help="with --guide-profile-frames: plain frames instead of guided ones")
    ap.add_argument("--no-guide", action="store_true",
                    help="trace mode: leave out the guided/plain equal-error run")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch path-tracing frames directly instead of replaying HIP graphs")
    a = ap.parse_args()
    if a.steps is None:
        a.steps = 100 if a.mode == "trace" else 20
    if a.warmup is None:
        a.warmup = 10 if a.mode == "trace" else 3
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {a.gpus}); "
              "multi-GPU sharding is a design note, not an implemented path", file=sys.stderr)
        sys.exit(2)
    L = lib()
    if a.no_graphs:
        check(L.rend_hip_pt_graphs(0), "graphs")
    if a.mode != "trace":
        sys.exit({"pt": bench_pt, "views": bench_views, "guide": bench_guide,
                  "tlas": bench_tlas, "blas": bench_blas}[a.mode](a))

    t0 = time.perf_counter()
    scene = Scene(config=2, seed=a.seed, detail=1.0, width=W, height=H, spp=1)
    log(f"scene: {scene.ntris} tris, build {time.perf_counter() - t0:.2f} s")
    n = W * H
    d_prim = L.rend_hip_malloc(n * RAY.itemsize)
    d_phit = L.rend_hip_malloc(n * HIT.itemsize)
    d_diff = L.rend_hip_malloc(n * RAY.itemsize)
    d_dhit = L.rend_hip_malloc(n * HIT.itemsize)
    if not all((d_prim, d_phit, d_diff, d_dhit)):
        raise SystemExit(f"device allocation failed: {L.rend_hip_errstr().decode()}")
    stream = L.rend_hip_stream()
    check(L.rend_hip_gen_camrays(scene.rd, d_prim, W, H, 1, a.seed, stream), "camrays")
    check(L.rend_hip_trace(scene.rd, d_prim, d_phit, n, stream, None), "primary trace")
    check(L.rend_hip_gen_diffuse(scene.rd, d_prim, d_phit, d_diff, n, a.seed, stream),
          "diffuse rays")
    check(L.rend_hip_sync(), "sync")

    st_d, st_p = TraceStats(), TraceStats()
    check(L.rend_hip_trace(scene.rd, d_diff, d_dhit, n, stream, C.byref(st_d)), "count")
    lanes = (C.c_ulonglong * 4)()
    check(L.rend_hip_trace_lanes(lanes), "lane counters")
    wave = {"iterations": int(lanes[0]),
            "lanes_per_iteration": round(lanes[1] / max(lanes[0], 1), 2),
            "leaf_steps": int(lanes[2]), "setup_steps": int(lanes[3])}
    log(f"diffuse launch wave loop: {wave}")
    check(L.rend_hip_trace(scene.rd, d_prim, d_phit, n, stream, C.byref(st_p)), "count")
    phits = np.zeros(n, HIT)
    check(L.rend_hip_memcpy_dtoh(phits.ctypes.data, d_phit, n * HIT.itemsize), "dtoh")
    prim_hit_frac = float((phits["id"] != 0xFFFFFFFF).mean())
    log(f"primary hit fraction {prim_hit_frac:.4f}; diffuse visits/ray: "
        f"nodes {st_d.nodes / n:.2f} leaves {st_d.leaves / n:.2f}")

    for _ in range(a.warmup):
        check(L.rend_hip_trace(scene.rd, d_diff, d_dhit, n, stream, None), "warmup")
        check(L.rend_hip_trace(scene.rd, d_prim, d_phit, n, stream, None), "warmup")
    check(L.rend_hip_sync(), "sync")
    log("warmup done")

    if a.sweep:
        tun0 = (C.c_uint * 3)()
        check(L.rend_hip_get_tuning(tun0), "tuning")
        for variant in range(4):
            check(L.rend_hip_set_trace_variant(variant), "variant")
            for mult in (0,):
                for refill in (32, 48, 56):
                    check(L.rend_hip_set_tuning(mult, refill), "tuning")
                    ms = timed_launches(L, scene, d_diff, d_dhit, n, 5) / 5
                    msp = timed_launches(L, scene, d_prim, d_phit, n, 5) / 5
                    log(f"sweep variant {variant} blocks/CU {mult or 'occ'} refill {refill}: "
                        f"diffuse {ms:.3f} ms ({n / ms / 1e3:.0f} Mrays/s), primary "
                        f"{msp:.3f} ms")
        check(L.rend_hip_set_trace_variant(int(tun0[2])), "variant")
        check(L.rend_hip_set_tuning(tun0[0], tun0[1]), "tuning")

    ms_d = timed_launches(L, scene, d_diff, d_dhit, n, a.steps)
    log(f"diffuse: {ms_d / a.steps:.3f} ms/launch")
    ms_p = timed_launches(L, scene, d_prim, d_phit, n, a.steps)
    log(f"primary: {ms_p / a.steps:.3f} ms/launch")

    if a.calib_copy:
        nb_ = a.calib_copy << 20
        src, dst = L.rend_hip_malloc(nb_), L.rend_hip_malloc(nb_)
        check(L.rend_hip_copy(src, dst, nb_, stream), "copy")
        check(L.rend_hip_sync(), "sync")
        log(f"calibration copy: {nb_} bytes read, {nb_} bytes written")
        L.rend_hip_free(src)
        L.rend_hip_free(dst)

    ms_step = ms_d / a.steps
    mrays = n / (ms_step * 1e-3) / 1e6
    mrays_p = n / (ms_p / a.steps * 1e-3) / 1e6
    bytes_alg = compulsory_bytes(scene, n)
    achieved = bytes_alg / (ms_step * 1e-3) / 1e9
    bytes_vis = visit_bytes(st_d, n)
    prof = profile_counters()

    cpu = None
    if not a.no_cpu:
        rays = np.zeros(n, RAY)
        check(L.rend_hip_memcpy_dtoh(rays.ctypes.data, d_diff, n * RAY.itemsize), "dtoh")
        cpu = cpu_baseline(scene, rays, a.cpu_seconds)
        log(f"cpu baseline: {cpu}")

    for p in (d_prim, d_phit, d_diff, d_dhit):
        L.rend_hip_free(p)
    name = C.create_string_buffer(256)
    cus = C.c_int(0)
    L.rend_hip_devinfo(C.byref(cus), name, 256)
    out = {
        "metric": "incoherent Mrays/s (closest hit)",
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural terrain and GPU-generated rays)",
        "config": {
            "workload": "closest-hit traversal microbench: 1M-tri displaced terrain "
                        "(1000x500 cells, amp 0.05), 1920x1080 cosine-hemisphere "
                        "diffuse rays from primary hits; primary rays timed beside it",
            "tris": int(scene.ntris), "rays_per_step": n, "seed": a.seed,
            "primary_mrays_s": round(mrays_p, 2),
            "primary_hit_fraction": round(prim_hit_frac, 4),
            "visits_per_ray": {"nodes": round(st_d.nodes / n, 3),
                               "leaves": round(st_d.leaves / n, 3),
                               "insts": round(st_d.insts / n, 3)},
            "device": f"{name.value.decode()} ({cus.value} CUs)",
            "wave_loop": wave,
        },
        "roofline": {
            "bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
            "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
            "traffic": prof["hbm_bytes"],
            "algorithmic_bytes_per_launch": int(bytes_alg),
            "algorithmic_bytes": "rays x 48 B (32 B ray in, 16 B hit out) + device "
                                 "arena (BLAS + TLAS) + 64 B per instance, once per launch",
            "traffic_source": f"{PMC_PROFILE} (FETCH_SIZE + WRITE_SIZE per launch, "
                              "copy-kernel scaled)",
            "hbm_frac_measured": (round(prof["hbm_bytes"] / (ms_step * 1e-3) / 1e9
                                        / PEAK_HBM_GBS, 4) if prof["hbm_bytes"] else None),
            "visit_bytes_per_launch": int(bytes_vis),
            "visit_bytes_rate_GBs": round(bytes_vis / (ms_step * 1e-3) / 1e9, 1),
            "visit_bytes": "node 228 B, leaf 160 B, instance 56 B per visit of the "
                           "reference algorithm + ray records: served by L1/L2/MALL",
            "valu_issue_frac": (round(prof["valu_insts"] * VALU_CYCLES / SIMDS / CLOCK_HZ
                                      / (ms_step * 1e-3), 4) if prof["valu_insts"] else None),
            "valu_source": f"{PMC_PROFILE} (SQ_INSTS_VALU per launch x 4 cycles / "
                           "(1024 SIMDs x 2.4 GHz x launch time))",
            "limiter": "VALU issue and memory latency of divergent traversal; HBM "
                       "bandwidth is not the binding ceiling"},
        "cpu_baseline": cpu,
    }
    scene.close()
    if not a.no_guide:
        # the path-guiding target on the same GPU in the same run: guided
        # path tracing's wall time to plain's converged relMSE, over plain's
        t = time.perf_counter()
        try:
            out["guided_vs_plain"] = guide_equal_error(a)
        except Exception as e:  # the headline line is still printed
            out["guided_vs_plain"] = {"error": f"{type(e).__name__}: {e}"}
        log(f"guided vs plain: {time.perf_counter() - t:.0f} s, "
            f"ratio {out['guided_vs_plain'].get('ratio')}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
