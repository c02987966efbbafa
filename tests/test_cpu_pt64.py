"""Closed-form checks of the FP64 path-tracing oracle (oracle/pt64.c), so
that GPU-versus-oracle image comparisons rest on an oracle that is itself
pinned:
  * Lambertian indirect weight and background: a convex diffuse object
    under uniform background radiance B with no emitters returns exactly
    a*B per sample (every bounce escapes), B on a miss;
  * next-event estimation: direct light on a floor from a square emitter
    equals albedo/pi times Lambert's polygon irradiance (which a dropped
    cosine, a wrong pdf conversion or a missing 1/pi would break);
  * multi-bounce throughput, path depth and Russian roulette: inside a
    closed diffuse sphere with an emissive cap the wall radiance has the
    closed form a Le f sum_{k<D} (a (1 - f))^k for D path vertices
    (integrating-sphere property); a dropped albedo factor, a missing 1/q
    after a roulette survival or a depth off by one each miss it by far more
    than the tolerance;
  * Russian roulette is unbiased on a scene without a closed form (the box
    of the path-tracing benchmark): roulette from the first bounce and no
    roulette agree within the Monte Carlo error;
  * Monte Carlo convergence of the oracle's images themselves."""
import numpy as np
import pytest

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


SPHERE = dict(albedo=0.5, le=1.0, cap_y=0.8)


@pytest.mark.parametrize("depth,rr", [(1, 0), (2, 0), (4, 0), (4, 1), (8, 3), (8, 1)])
def test_oracle_integrating_sphere_closed_form(depth, rr):
    s, frac = S.integrating_sphere(SPHERE["albedo"], SPHERE["le"], SPHERE["cap_y"], subdiv=5)
    img = or_pt64(s, 2048, maxdepth=depth, rr_depth=rr, seed=3)
    s.close()
    want = S.integrating_sphere_radiance(SPHERE["albedo"], SPHERE["le"], frac, depth)
    # 256 pixels x 2048 paths: the mean's standard error is < 0.1 % (with
    # roulette from the first bounce); the 20,480-facet mesh is within
    # 3e-4 of the smooth sphere's closed form
    got = img.mean()
    assert abs(got / want - 1) < 3e-3, (depth, rr, got, want)
    # the same tolerance separates every depth from the next one
    nxt = S.integrating_sphere_radiance(SPHERE["albedo"], SPHERE["le"], frac, depth + 1)
    assert nxt / want - 1 > 3 * 3e-3 or depth >= 8
    # and the wall is uniform: every pixel is within its own noise of it
    assert np.abs(img / want - 1).max() < 0.2


def test_closed_form_detects_planted_mistakes():
    """What the closed form would read for an estimator that drops the
    albedo from the throughput, or forgets the 1/q after a roulette
    survival (q = the surviving throughput's maximum, below 0.95), at
    depth 4: both far outside the tolerance above."""
    a, le, f = SPHERE["albedo"], SPHERE["le"], 0.1
    good = S.integrating_sphere_radiance(a, le, f, 4)
    no_albedo = a * le * f * sum((1 - f) ** k for k in range(4))
    # roulette from the first bounce without 1/q: vertex k survives with
    # prod_{j<=k} min(a^j, 0.95) and is not re-weighted
    no_inv_q = sum(a * le * f * (a * (1 - f)) ** k * np.prod([min(a ** j, 0.95)
                   for j in range(1, k + 1)]) for k in range(4))
    assert abs(no_albedo / good - 1) > 0.5
    assert abs(no_inv_q / good - 1) > 0.1


def test_oracle_russian_roulette_is_unbiased_on_the_box():
    from nodebleed_amd import Scene
    s = Scene(config=1, detail=0.25, width=16, height=16)
    off = or_pt64(s, 1024, maxdepth=6, rr_depth=0, seed=11)
    on = or_pt64(s, 1024, maxdepth=6, rr_depth=1, seed=12)
    s.close()
    # whole-image means: 262,144 paths each, relative standard error of the
    # roulette image ~0.3 %
    assert abs(on.mean() / off.mean() - 1) < 0.015, (on.mean(), off.mean())
    # and per 4x4 block
    b_on = on.reshape(4, 4, 4, 4, 3).mean(axis=(1, 3))
    b_off = off.reshape(4, 4, 4, 4, 3).mean(axis=(1, 3))
    assert np.abs(b_on - b_off).max() < 0.1 * b_off.max()


def test_oracle_tlas_pruning_changes_nothing():
    """The oracle prunes instances with the host TLAS when that tree is
    sound; images are identical to testing every instance."""
    from nodebleed_amd import Scene
    from oracle_lib import oracle
    s = Scene(config=3, detail=0.03, width=16, height=12)
    try:
        a = or_pt64(s, 16, maxdepth=8, rr_depth=3, seed=5)
        oracle().or_pt64_tlas(0)
        b = or_pt64(s, 16, maxdepth=8, rr_depth=3, seed=5)
    finally:
        oracle().or_pt64_tlas(1)
        s.close()
    assert np.array_equal(a, b) and a.mean() > 0


PT_FIXTURES = {"box_full": 16, "instances_full": 64, "occluded_full": 256}


@pytest.mark.parametrize("name", sorted(PT_FIXTURES))
def test_fp64_fixtures_follow_the_oracle(name):
    """Each committed FP64 reference image (tests/golden/make_pt_golden.py)
    is what the oracle renders now: a low-sample render of the same case
    agrees with it within that render's own noise (block relMSE and image
    mean), so a change of the estimator shows up here first."""
    from pathlib import Path

    from nodebleed_amd import Scene
    from oracle_lib import rel_mse
    f = np.load(Path(__file__).resolve().parent / "golden" / f"pt64_{name}.npz")
    w, h, k = int(f["width"]), int(f["height"]), int(f["block"])
    s = Scene(config=int(f["config"]), detail=float(f["detail"]), width=w, height=h)
    try:
        spp = PT_FIXTURES[name]
        img = or_pt64(s, spp, maxdepth=int(f["maxdepth"]), rr_depth=int(f["rr_depth"]),
                      seed=101)
    finally:
        s.close()
    b = img.reshape(h // k, k, w // k, k, 3).mean(axis=(1, 3))
    ref = f["blocks"].astype(np.float64)
    # noise of the low-sample render alone: per-block sample variance from
    # the two half-images would need the pixels; bound it by the measured
    # relMSE levels of these cases (box 4e-4, instances 3.4e-3, occluded
    # 3.4e-2 at these sample counts) with a factor 4 of margin
    bound = {"box_full": 2e-3, "instances_full": 1.5e-2, "occluded_full": 0.15}[name]
    err = rel_mse(b, ref)
    assert err < bound, (name, err)
    assert abs(b.mean() / ref.mean() - 1) < {"box_full": 0.01, "instances_full": 0.03,
                                              "occluded_full": 0.1}[name]
