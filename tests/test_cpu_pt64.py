"""Closed-form checks of the FP64 path-tracing oracle (oracle/pt64.c), so
that GPU-versus-oracle image comparisons rest on an oracle that is itself
pinned:
  * Lambertian indirect weight and background: a convex diffuse object
    under uniform background radiance B with no emitters returns exactly
    a*B per sample (every bounce escapes), B on a miss;
  * next-event estimation: direct light on a floor from a square emitter
    equals albedo/pi times Lambert's polygon irradiance (which a dropped
    cosine, a wrong pdf conversion or a missing 1/pi would break);
  * Monte Carlo convergence of the oracle's images themselves."""
import numpy as np

import scenes_closed as S
from oracle_lib import or_pt64


def test_oracle_convex_cube_is_exact():
    a = (0.2, 0.5, 0.8)
    s = S.cube_scene(a, (1.0, 1.0, 1.0), w=24, h=24)
    img = or_pt64(s, 16, maxdepth=4, rr_depth=0, seed=2)
    s.close()
    f = S.cube_coverage(img, a)
    assert np.abs(f[..., 0] - f[..., 1]).max() < 1e-12
    assert np.abs(f[..., 1] - f[..., 2]).max() < 1e-12
    assert f.min() > -1e-12 and f.max() < 1 + 1e-12
    inside = np.abs(f[..., 0] - 1) < 1e-12
    assert inside.sum() > 50 and (np.abs(f[..., 0]) < 1e-12).sum() > 50
    assert np.allclose(img[inside], np.float32(a).astype(np.float64), rtol=1e-12)


def test_oracle_direct_light_matches_polygon_irradiance():
    s = S.lit_floor_scene(w=4, h=4)
    img = or_pt64(s, 65536, maxdepth=1, rr_depth=0, seed=4)
    s.close()
    want = S.lit_floor_radiance()
    # the view centre is the centre of pixel (w/2, h/2): the reference's
    # sx = ((x + u - 0.5) / w - 0.5) puts pixel x's centre at x / w - 0.5
    got = img[2, 1:4].mean()
    # 3 pixels x 65536 samples: relative standard error ~0.1 %; the
    # footprint's curvature of E is below that
    assert abs(got / want - 1) < 0.005, (got, want)
    # the floor is lit more strongly towards the light (image top)
    assert (np.diff(img[:, 2, 0]) < 0).all()
    assert np.allclose(img[..., 0], img[..., 1]) and np.allclose(img[..., 1], img[..., 2])


def test_polygon_irradiance_formula():
    """The closed form itself: a disc-like limit (a large square far above)
    and the unit-hemisphere limit (an infinite emitter gives E = pi Le)."""
    big = [(-1e4, 1, -1e4), (1e4, 1, -1e4), (1e4, 1, 1e4), (-1e4, 1, 1e4)]
    assert abs(S.polygon_irradiance((0, 0, 0), (0, 1, 0), big, 1.0) / np.pi - 1) < 1e-3
    # small square of side d at height h: E ~ Le * d^2 / h^2 (cosines = 1)
    d, h = 1e-3, 1.0
    sq = [(-d / 2, h, -d / 2), (d / 2, h, -d / 2), (d / 2, h, d / 2), (-d / 2, h, d / 2)]
    assert abs(S.polygon_irradiance((0, 0, 0), (0, 1, 0), sq, 1.0) / (d * d / h / h) - 1) < 1e-5
