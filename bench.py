"""Closest-hit traversal benchmark on one MI355X.

Workload (the headline configuration): a single BLAS of 1,000,000 triangles
(seeded displaced terrain, 1000 x 500 cells, height amplitude 0.05), one
instance, 1920x1080 jittered primary rays and one cosine-hemisphere diffuse
ray per primary hit.  Rays are generated on the GPU, so inputs are resident
in HBM before timing starts.  A step is one closest-hit launch over the
2,073,600 incoherent (diffuse) rays; the primary-ray launch is timed beside
it and reported as an extra field.

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

from nodebleed_amd import HIT, RAY, Scene, TraceStats, check, lib  # noqa: E402

W, H = 1920, 1080
PEAK_HBM_GBS = 8000.0
# Bytes one lane reads per visit in the device layout (nodebleed_amd/csrc/dlayout.h):
# node = 6x8 bounds + 8 child words + 1 rank word, leaf = 3 vertices/edges x 4
# triangles + 4 ids, instance = 12-float inverse + 8-byte BLAS offset; each ray
# reads 32 B and writes 16 B.
S_NODE, S_LEAF, S_INST, S_RAY = 228, 160, 56, 48


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(st: TraceStats, nrays: int) -> int:
    return nrays * S_RAY + st.nodes * S_NODE + st.leaves * S_LEAF + st.insts * S_INST


def measured_traffic():
    """HBM bytes per diffuse-ray launch from the committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE per dispatch, scaled by the ratio the
    same counters show for a plain copy kernel of known size), or None."""
    path = os.path.join(ROOT, "profiles", "trace_terrain_diffuse_pmc.md")
    try:
        with open(path) as f:
            for line in f:
                if line.strip().startswith("hbm_bytes_per_launch:"):
                    return int(line.split(":", 1)[1].strip().split()[0])
    except (OSError, ValueError):
        pass
    return None


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
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    threads = max(1, min(int(threads), 16))
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
    # whole passes over the buffer until about `seconds` of CPU work
    passes = max(1, int(seconds / max(dt / len(probe) * n, 1e-9)))
    t = time.perf_counter()
    for _ in range(passes):
        run(rays[:n])
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
            "kind": kind,
            "sample": f"{what}, {n} rays ({passes} pass(es) over the first "
                      f"{n // passes} of {len(rays)} diffuse rays), {threads} threads "
                      f"on {cpu}, {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--sweep", action="store_true",
                    help="time traversal launch shapes (blocks/CU x refill) on stderr")
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MIB",
                    help="after timing, run one plain copy kernel of MIB MiB "
                         "(reference for hardware byte counters)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {a.gpus}); "
              "multi-GPU sharding is a design note, not an implemented path", file=sys.stderr)
        sys.exit(2)

    L = lib()
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
        for mult in (0, 2, 3, 4, 6, 8):
            for refill in (4, 8, 16, 32):
                check(L.rend_hip_set_tuning(mult, refill), "tuning")
                ms = timed_launches(L, scene, d_diff, d_dhit, n, 5) / 5
                msp = timed_launches(L, scene, d_prim, d_phit, n, 5) / 5
                log(f"sweep blocks/CU {mult or 'occ'} refill {refill}: diffuse {ms:.3f} ms "
                    f"({n / ms / 1e3:.0f} Mrays/s), primary {msp:.3f} ms")
        check(L.rend_hip_set_tuning(0, 16), "tuning")

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
    bytes_d = algorithmic_bytes(st_d, n)
    achieved = bytes_d / (ms_step * 1e-3) / 1e9

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
        },
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
                     "traffic": measured_traffic(),
                     "algorithmic_bytes_per_launch": int(bytes_d)},
        "cpu_baseline": cpu,
    }
    scene.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
