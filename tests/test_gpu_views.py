"""The batched uneven workload at its stated size: 32 cameras over the
1M-triangle terrain, resolutions from 320x240 to 3840x2160, spp 4..64, all
views in one batched, graph-captured launch sequence per call.

* Split and co-batching invariance at full size: with one sample per pixel
  every accumulator receives exactly one addition, so a view's image must
  not depend on the other views of the batch.
* Against the FP64 oracle: a full-resolution view's image, box-filtered
  down by an odd factor k, estimates what the oracle's camera renders at the
  coarse resolution.  Pixel x samples [x - 1/2, x + 1/2) (the reference's
  jitter, rend.c:3302-3308), so coarse pixel i covers fine pixels
  k i - (k - 1)/2 .. k i + (k - 1)/2 exactly; jittered samples cover each
  footprint uniformly, and block means agree within the image parity bar
  (coarse row and column 0 reach past the fine image and are left out)."""
import ctypes as C
import sys

import numpy as np
import pytest

from nodebleed_amd import Scene, check, lib, uneven_views
from oracle_lib import or_pt64, rel_mse

pytestmark = pytest.mark.gpu

MAXDEPTH, RR = 4, 3


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def terrain():
    s = Scene(config=2, detail=1.0, width=64, height=64)
    yield s
    s.close()


def coarse(img, k):
    """Coarse pixels 1.. of an odd box factor k (see the module note)."""
    h, w, _ = img.shape
    r = (k - 1) // 2
    c = img[r + 1:h - r, r + 1:w - r]
    return c.reshape(h // k - 1, k, w // k - 1, k, 3).mean(axis=(1, 3))


def blocks2(img):
    h, w, _ = img.shape
    h, w = h - h % 2, w - w % 2
    return img[:h, :w].reshape(h // 2, 2, w // 2, 2, 3).mean(axis=(1, 3))


def test_full_size_views_are_independent_of_the_batch(terrain):
    v = uneven_views(1)
    one = v.copy()
    one["spp"] = 1
    say(f"32 views at 1 spp, {int((one['w'].astype(np.int64) * one['h']).sum())} paths, batched")
    together, rays = terrain.render_views(one, maxdepth=MAXDEPTH, rr_depth=RR)
    assert rays > int((one["w"].astype(np.int64) * one["h"]).sum())
    for i in (2, 12):
        say(f"view {i} ({int(one[i]['w'])}x{int(one[i]['h'])}) alone")
        alone, _ = terrain.render_views(one[i:i + 1], maxdepth=MAXDEPTH, rr_depth=RR)
        assert np.array_equal(alone[0], together[i]), i
    for img in together:
        assert np.isfinite(img).all() and (img >= 0).all()


#: view index -> (odd box-filter factor, oracle spp): a 3840x2160, a
#: 1920x1080 and a 320x240 view of the set, checked on 256x144, 128x72 and
#: 64x48 grids
CHECKED = {10: (15, 1024), 4: (15, 2048), 2: (5, 2048)}


def test_full_size_views_match_fp64_oracle(terrain):
    v = uneven_views(1)
    say(f"32 views at their spp: {int((v['w'].astype(np.int64) * v['h'] * v['spp']).sum())} "
        "paths in one batched call")
    imgs, rays = terrain.render_views(v, maxdepth=MAXDEPTH, rr_depth=RR)
    say(f"{rays} rays")
    L = lib()
    for i, (k, spp) in CHECKED.items():
        x = v[i]
        w, h = int(x["w"]) // k, int(x["h"]) // k
        s = Scene(config=2, detail=1.0, width=w, height=h)
        try:
            f3 = C.c_float * 3
            check(L.rs_set_rcam(s.rd, f3(*x["eye"]), f3(*x["target"]), float(x["vfov"]),
                                float(x["focdist"]), float(x["focangle"]), w, h), "camera")
            say(f"view {i}: oracle {w}x{h} at {spp} spp")
            ref = or_pt64(s, spp, maxdepth=MAXDEPTH, rr_depth=RR, seed=17 + i, threads=16)
        finally:
            s.close()
        got, want = coarse(imgs[i], k), ref[1:, 1:]
        err = rel_mse(blocks2(got), blocks2(want))
        say(f"view {i} ({int(x['w'])}x{int(x['h'])}, {int(x['spp'])} spp): relMSE of 2x2 "
            f"blocks of the {w}x{h} grid {err:.2e}, mean {got.mean():.5f} vs {want.mean():.5f}")
        assert err <= 1e-3, (i, err)
        assert abs(got.mean() / want.mean() - 1) < 2e-3, (i, got.mean(), want.mean())
