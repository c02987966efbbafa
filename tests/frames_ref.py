"""Loads the frame fixtures of tests/golden/frames_<name>.npz: three
frames of the reference's rend_render (normal-tint shading, 64x32 blocks,
its own primary rays; box_lens through a thin lens) and of the same block
walk traced by its single-ray trace_rays (tests/golden/make_golden.py)."""
from __future__ import annotations

from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden"
NAMES = ("box", "terrain", "box_lens")


def load(name: str) -> dict:
    z = np.load(GOLDEN / f"frames_{name}.npz")
    f = {k: z[k] for k in z.files}
    w, h = int(f["width"]), int(f["height"])
    buf_p = f["buf_single"].copy().reshape(-1)
    acc_p = f["acc_single"].copy().reshape(-1, 3)
    buf_p[f["packet_diff_px"]] = f["buf_packets_diff"]
    acc_p[f["packet_diff_px"]] = f["acc_packets_diff"]
    f["buf_packets"] = buf_p.reshape(h, w)
    f["acc_packets"] = acc_p.reshape(h, w, 3)
    return f


def scene_of(f: dict):
    """The fixture's scene and camera (a thin lens where the fixture has one)."""
    from nodebleed_amd import Scene
    from raygen import set_lens_camera
    s = Scene(config=int(f["config"]), detail=float(f["detail"]), width=int(f["width"]),
              height=int(f["height"]))
    lens = f.get("lens")
    if lens is not None and float(lens[0]) > 0:
        set_lens_camera(s, float(lens[0]), float(lens[1]))
    return s
