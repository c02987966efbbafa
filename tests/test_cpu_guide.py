"""CPU checks of the path-guiding mathematics (numpy restatement in
guide_ref.py): vMF mixture density normalisation, the sampling routine's
consistency with the density, the one-sample MIS weight, and the analytic
loss gradient against finite differences."""
import numpy as np
import pytest

import guide_ref as G


def random_outputs(n, seed):
    rng = np.random.default_rng(seed)
    o = rng.normal(size=(n, G.GOUT))
    o[:, 4:5 * G.LOBES:5] = rng.uniform(-1.0, 4.0, (n, G.LOBES))
    return o


def uniform_sphere(n, rng):
    d = rng.normal(size=(n, 3))
    return d / np.linalg.norm(d, axis=1, keepdims=True)


def sample_vmf_mixture(o, u):
    """numpy restatement of gd_sample (Jakob 2012) for checks."""
    pi, mu, k = G.mixture(o)
    cdf = np.cumsum(pi, axis=1)
    j = np.minimum((u[:, :1] >= cdf).sum(axis=1), G.LOBES - 1)
    n = np.arange(len(o))
    m, kk = mu[n, j], k[n, j]
    v = 1 - u[:, 1]
    w = 1 + np.log(v + (1 - v) * np.exp(-2 * kk)) / kk
    sw = np.sqrt(np.maximum(0, 1 - w * w))
    phi = 2 * np.pi * u[:, 2]
    sgn = np.where(m[:, 2] >= 0, 1.0, -1.0)
    a = -1 / (sgn + m[:, 2])
    b = m[:, 0] * m[:, 1] * a
    t1 = np.stack([1 + sgn * m[:, 0] ** 2 * a, sgn * b, -sgn * m[:, 0]], 1)
    t2 = np.stack([b, sgn + m[:, 1] ** 2 * a, -m[:, 1]], 1)
    return w[:, None] * m + sw[:, None] * (np.cos(phi)[:, None] * t1 + np.sin(phi)[:, None] * t2)


def test_mixture_density_integrates_to_one():
    rng = np.random.default_rng(0)
    o = random_outputs(4, 1)
    d = uniform_sphere(400_000, rng)
    for i in range(len(o)):
        est = 4 * np.pi * G.pdf(np.repeat(o[i:i + 1], len(d), 0), d).mean()
        assert abs(est - 1.0) < 0.03, est


def test_sampling_matches_density():
    """E_p[f] from samples equals the quadrature of f * p (f = smooth test
    function), and the samples are unit vectors."""
    rng = np.random.default_rng(2)
    o = np.repeat(random_outputs(1, 3), 200_000, 0)
    s = sample_vmf_mixture(o, rng.random((len(o), 3)))
    assert np.allclose(np.linalg.norm(s, axis=1), 1.0, atol=1e-9)
    f = lambda d: np.exp(d[:, 0]) + d[:, 1] ** 2 + 0.3 * d[:, 2]  # noqa: E731
    mc = f(s).mean()
    q = uniform_sphere(400_000, rng)
    quad = 4 * np.pi * (f(q) * G.pdf(np.repeat(o[:1], len(q), 0), q)).mean()
    assert abs(mc - quad) < 0.02 * abs(quad), (mc, quad)


def test_one_sample_mis_weight_is_unbiased():
    """E[(cos/pi)/p] = 1 over the hemisphere for p = a p_vmf + (1-a) cos/pi."""
    rng = np.random.default_rng(4)
    n = 300_000
    o = np.repeat(random_outputs(1, 5), n, 0)
    nrm = np.array([0.0, 0.0, 1.0])
    alpha = 0.5
    u = rng.random((n, 4))
    g = sample_vmf_mixture(o, u[:, 1:])
    r1, phi = u[:, 2], 2 * np.pi * u[:, 3]
    c = np.stack([np.cos(phi) * np.sqrt(r1), np.sin(phi) * np.sqrt(r1), np.sqrt(1 - r1)], 1)
    d = np.where((u[:, :1] < alpha), g, c)
    cos = d @ nrm
    p = alpha * G.pdf(o, d) + (1 - alpha) * np.maximum(cos, 0) / np.pi
    w = np.where(cos > 0, (cos / np.pi) / p, 0.0)
    assert abs(w.mean() - 1.0) < 0.01, w.mean()


@pytest.mark.parametrize("bsdf", [0.0, 0.5])
def test_loss_gradient_matches_finite_differences(bsdf):
    """bsdf = 0: KL to the vMF mixture; 0.5: KL to the one-sample MIS
    density the renderer samples."""
    rng = np.random.default_rng(6)
    p = np.zeros(G.NPARAM)
    p[:G.N_GRID] = rng.uniform(-0.3, 0.3, G.N_GRID)
    for ofs, end, fan in ((G.OFS_W1, G.OFS_B1, G.GIN), (G.OFS_W2, G.OFS_B2, G.HID),
                          (G.OFS_W3, G.OFS_B3, G.HID), (G.OFS_W4, G.OFS_B4, G.HID)):
        p[ofs:end] = rng.uniform(-1, 1, end - ofs) * np.sqrt(3.0 / fan)
    p[G.OFS_B1:G.OFS_W2] = 0.1
    p[G.OFS_B4 + 4:G.OFS_B4 + 5 * G.LOBES:5] = 1.0
    n = 48
    rec = np.zeros(n, G.RECORD)
    rec["pos"] = rng.uniform(0, 1, (n, 3))
    rec["nrm"] = uniform_sphere(n, rng)
    rec["dir"] = uniform_sphere(n, rng)
    rec["target"] = rng.uniform(0.1, 2.0, n)
    gmin, ginv = np.zeros(3, np.float32), np.ones(3, np.float32) * np.float32(0.9)
    g = G.grad(p, rec, gmin, ginv, bsdf)
    # check MLP weights, biases and grid entries that the records touch
    q = G.normalise(rec["pos"], gmin, ginv)
    idx, _ = G.corners(q, 5)
    grid_ids = 5 * G.TABLE * G.FEAT + idx[:3, :2].ravel() * G.FEAT
    ids = np.concatenate([rng.integers(G.OFS_W1, G.NPARAM, 40), grid_ids])
    h = 1e-6
    for i in ids:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (G.loss(pp, rec, gmin, ginv, bsdf) - G.loss(pm, rec, gmin, ginv, bsdf)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-5 + 2e-3 * abs(fd), (i, fd, g[i])


def _host_sample(mix, u):
    import ctypes as C

    from nodebleed_amd import lib
    out = np.zeros(3, np.float32)
    m = np.ascontiguousarray(mix, np.float32)
    rc = lib().rend_hip_guide_sample_host(C.c_void_p(m.ctypes.data), float(u[0]), float(u[1]),
                                          float(u[2]), C.c_void_p(out.ctypes.data))
    assert rc == 0
    return out


def _mix_record(pi, mu, k):
    rec = np.zeros(5 * G.LOBES, np.float32)
    rec[0::5], rec[1::5], rec[2::5], rec[3::5], rec[4::5] = pi, mu[:, 0], mu[:, 1], mu[:, 2], k
    return rec


def test_product_sampler_matches_restatement():
    """The device sampling routine (compiled for the host) draws the same
    directions as the numpy restatement for the same uniforms."""
    rng = np.random.default_rng(5)
    o = random_outputs(400, 11)
    u = rng.random((400, 3))
    want = sample_vmf_mixture(o, u)
    pi, mu, k = G.mixture(o)
    for i in range(len(o)):
        got = _host_sample(_mix_record(pi[i], mu[i], k[i]), u[i])
        assert np.all(np.isfinite(got))
        assert np.abs(got - want[i]).max() < 2e-3, (i, got, want[i])


@pytest.mark.parametrize("logk", [-2.0, 2.0, 7.0])
def test_sampler_is_finite_at_the_ends_of_the_unit_interval(logk):
    """u1 = 0 (a 24-bit uniform can be exactly 0) with a concentration whose
    e^{-2k} underflows gives the lobe's mean direction, not log(0) = -inf;
    the largest u1 gives the widest finite angle."""
    k = float(np.exp(logk))
    mu = np.zeros((G.LOBES, 3))
    mu[:, 1] = 1.0
    pi = np.full(G.LOBES, 1.0 / G.LOBES)
    rec = _mix_record(pi, mu, np.full(G.LOBES, k))
    for u1 in (0.0, 1.0 - 2.0 ** -24):
        for u2 in (0.0, 0.3, 1.0 - 2.0 ** -24):
            d = _host_sample(rec, (0.5, u1, u2))
            assert np.all(np.isfinite(d)), (k, u1, u2, d)
            assert abs(np.linalg.norm(d) - 1.0) < 1e-5
            if u1 == 0.0:
                assert np.allclose(d, (0.0, 1.0, 0.0), atol=1e-6)
            else:
                v = 2.0 ** -24
                w = 1 + np.log(v + (1 - v) * np.exp(-2 * k)) / k
                assert abs(d[1] - w) < 1e-4 * max(1.0, abs(w)), (k, d, w)


UNIT_COS, UNIT_RATIO = G.UNIT_COS, G.UNIT_RATIO


def _grad_case(seed=8, n=256):
    rng = np.random.default_rng(seed)
    p = np.zeros(G.NPARAM)
    p[:G.N_GRID] = rng.uniform(-0.5, 0.5, G.N_GRID)
    for ofs, end, fan in ((G.OFS_W1, G.OFS_B1, G.GIN), (G.OFS_W2, G.OFS_B2, G.HID),
                          (G.OFS_W3, G.OFS_B3, G.HID), (G.OFS_W4, G.OFS_B4, G.HID)):
        p[ofs:end] = rng.uniform(-1, 1, end - ofs) * np.sqrt(6.0 / fan)
    p[G.OFS_B1:G.OFS_W2] = 0.05
    p[G.OFS_B4 + 4:G.OFS_B4 + 5 * G.LOBES:5] = 1.0
    p = p.astype(np.float32).astype(np.float64)
    rec = np.zeros(n, G.RECORD)
    rec["pos"] = rng.uniform(0, 1, (n, 3))
    rec["nrm"] = uniform_sphere(n, rng)
    rec["dir"] = uniform_sphere(n, rng)
    rec["target"] = rng.uniform(0.1, 2.0, n)
    gmin, ginv = np.zeros(3, np.float32), np.ones(3, np.float32) * np.float32(0.9)
    return p, rec, gmin, ginv


def test_unit_blocks_catch_a_wrong_lobe_gradient():
    """The per-unit comparison the GPU gradient test applies against the
    BF16-emulating restatement: a mistake in a single lobe's gradient -- its
    log-concentration term 15 % off, one mean-direction component's sign, its
    logit term dropped -- fails exactly that output unit, and the correct
    gradient fails none.  (The same bounds hold the GPU kernel to the
    restatement, test_gpu_guide.py.)"""
    p, rec, gmin, ginv = _grad_case()
    emul = G.grad(p, rec, gmin, ginv, emulate=True)
    assert np.abs(G.unit_blocks(emul)[5 * G.LOBES:]).max() == 0.0

    def kappa(go):
        go[:, 5 * 3 + 4] *= 0.85

    def direction(go):
        go[:, 5 * 5 + 2] *= -1.0

    def logit(go):
        go[:, 5 * 1] = 0.0

    for plant, unit in ((kappa, 19), (direction, 27), (logit, 5)):
        bad = G.grad(p, rec, gmin, ginv, emulate=True, plant=plant)
        cos, ratio = G.block_agreement(bad, emul)
        used = np.arange(5 * G.LOBES)
        fails = used[(cos[used] < UNIT_COS) | (np.abs(ratio[used] - 1) > UNIT_RATIO)]
        assert fails.tolist() == [unit], (plant.__name__, fails)
