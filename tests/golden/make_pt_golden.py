"""Writes FP64 reference images for the GPU path-tracing parity tests.

The oracle (oracle/pt64.c: the path tracer's estimator in double precision,
independent traversal and random numbers) renders each benchmark scene at a
reduced resolution but at the scene's full geometric size; the block means
of the image (what the GPU tests compare) go to tests/golden/pt64_<name>.npz
with the parameters that made them.  Data only; takes about ten minutes on
8 cores.

    python tests/golden/make_pt_golden.py [name ...]
"""
from __future__ import annotations

import subprocess
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

from nodebleed_amd import Scene  # noqa: E402
from oracle_lib import or_pt64  # noqa: E402

#: name -> config, detail, width, height, max depth, roulette depth, spp, block
CASES = {
    # sphere in the open box, 81,920 + 12 triangles, 256 x 256, depth 4,
    # roulette from depth 3 (the path-tracing benchmark's CPU-sized case)
    "box_full": (1, 1.0, 256, 256, 4, 3, 1024, 2),
    # 16 instanced meshes (2,479,120 triangles) x 4,095 instances + light,
    # depth 8, roulette from depth 3
    "instances_full": (3, 1.0, 64, 36, 8, 3, 4096, 4),
    # the occluded-lights terrain (1,000,000 terrain triangles, 16 emitter
    # triangles), depth 6, roulette from depth 3, 65,536 spp
    "occluded_full": (5, 1.0, 96, 54, 6, 3, 65536, 6),
}
SEED = 7


def blocks(img: np.ndarray, k: int) -> np.ndarray:
    h, w, _ = img.shape
    return img.reshape(h // k, k, w // k, k, 3).mean(axis=(1, 3))


def main(names):
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], check=True)
    for name in names:
        cfg, det, w, h, depth, rr, spp, k = CASES[name]
        s = Scene(config=cfg, detail=det, width=w, height=h)
        t = time.time()
        img = or_pt64(s, spp, maxdepth=depth, rr_depth=rr, seed=SEED)
        np.savez_compressed(HERE / f"pt64_{name}.npz", config=cfg, detail=det, width=w,
                            height=h, maxdepth=depth, rr_depth=rr, spp=spp, block=k,
                            seed=SEED, blocks=blocks(img, k).astype(np.float32))
        print(f"{name}: {w}x{h} {spp} spp depth {depth} rr {rr}: {time.time() - t:.0f} s, "
              f"mean {img.mean():.5g}", flush=True)
        s.close()


if __name__ == "__main__":
    main(sys.argv[1:] or list(CASES))
