"""GPU checks of the learned path-guiding kernels against the numpy
restatement (guide_ref.py): MFMA lane layout, BF16 network forward,
loss gradients, and unbiasedness of guided rendering."""
import ctypes as C
import sys

import numpy as np
import pytest

import guide_ref as G
from nodebleed_amd import Scene, check, lib, ptr
from oracle_lib import rel_mse

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def test_mfma_bf16_layout():
    rng = np.random.default_rng(0)
    A = rng.normal(size=(16, 32)).astype(np.float32)
    B = rng.normal(size=(32, 16)).astype(np.float32)
    D = np.zeros((16, 16), np.float32)
    say("mfma 16x16x32 bf16 layout check")
    check(lib().rend_hip_mfma_check(ptr(A), ptr(B), ptr(D)), "mfma")
    want = G.bf16(A).astype(np.float64) @ G.bf16(B).astype(np.float64)
    assert np.abs(D - want).max() < 1e-4 * (1 + np.abs(want).max())


@pytest.fixture(scope="module")
def guided_scene():
    s = Scene(config=5, detail=0.1, width=64, height=36)
    check(lib().rend_hip_guide_enable(s.rd, 0.5, 4, 2, 1e-2, 7), "guide enable")
    yield s
    s.close()


def random_params(seed):
    rng = np.random.default_rng(seed)
    p = np.zeros(G.NPARAM, np.float32)
    p[:G.N_GRID] = rng.uniform(-0.5, 0.5, G.N_GRID).astype(np.float32)
    for ofs, end, fan in ((G.OFS_W1, G.OFS_B1, G.GIN), (G.OFS_W2, G.OFS_B2, G.HID),
                          (G.OFS_W3, G.OFS_B3, G.HID), (G.OFS_W4, G.OFS_B4, G.HID)):
        p[ofs:end] = (rng.uniform(-1, 1, end - ofs) * np.sqrt(6.0 / fan)).astype(np.float32)
    p[G.OFS_B1:G.OFS_W2] = 0.05
    p[G.OFS_B4 + 4:G.OFS_B4 + 5 * G.LOBES:5] = 1.0
    return p


def points(scene, n, seed):
    rng = np.random.default_rng(seed)
    a = scene.aabbs()
    lo, hi = a["min"].min(0), a["max"].max(0)
    pos = (lo + rng.random((n, 3)) * (hi - lo)).astype(np.float32)
    nrm = rng.normal(size=(n, 3))
    nrm = (nrm / np.linalg.norm(nrm, axis=1, keepdims=True)).astype(np.float32)
    return pos, nrm


def test_forward_matches_cpu(guided_scene):
    s = guided_scene
    L = lib()
    p = random_params(1)
    check(L.rend_hip_guide_params(s.rd, ptr(p), 1))
    pos, nrm = points(s, 1000, 2)
    pn = np.ascontiguousarray(np.concatenate([pos, nrm], 1), np.float32)
    out = np.zeros((len(pos), G.GOUT), np.float32)
    say("guide forward: 1000 points")
    check(L.rend_hip_guide_forward_host(s.rd, ptr(pn), len(pos), ptr(out)))
    gmin, ginv = G.scene_box(s.aabbs())
    half = lambda g: g.astype(np.float16).astype(np.float64)  # noqa: E731
    x = G.encode(p.astype(np.float64), pos, nrm, gmin, ginv, grid_round=half)
    want, _ = G.forward(p.astype(np.float64), x, emulate_bf16=True)
    err = np.abs(out - want) / (1.0 + np.abs(want))
    say(f"  forward: error median {np.median(err):.2e}, 99.9 % {np.quantile(err, 0.999):.2e}, "
        f"max {err.max():.2e}")
    # the restatement rounds where the kernel rounds (FP16 grid, BF16 MFMA
    # operands): outputs agree to FP32 accumulation order, apart from rare
    # BF16 rounding ties decided the other way
    assert np.median(err) < 1e-6 and np.quantile(err, 0.999) < 1e-5 and err.max() < 1e-2, (
        np.median(err), np.quantile(err, 0.999), err.max())


def test_mixture_kernel_matches_cpu(guided_scene):
    """The rendering-time inference kernel (activations kept in registers,
    permuted weight copy) gives the restatement's vMF mixtures."""
    s = guided_scene
    L = lib()
    p = random_params(6)
    check(L.rend_hip_guide_params(s.rd, ptr(p), 1))
    pos, nrm = points(s, 1000, 7)
    pn = np.ascontiguousarray(np.concatenate([pos, nrm], 1), np.float32)
    out = np.zeros((len(pos), G.LOBES, 5), np.float32)
    say("guide mixture kernel: 1000 points")
    check(L.rend_hip_guide_mix_host(s.rd, ptr(pn), len(pos), ptr(out)))
    gmin, ginv = G.scene_box(s.aabbs())
    half = lambda g: g.astype(np.float16).astype(np.float64)  # noqa: E731
    x = G.encode(p.astype(np.float64), pos, nrm, gmin, ginv, grid_round=half)
    o, _ = G.forward(p.astype(np.float64), x, emulate_bf16=True)
    pi, mu, k = G.mixture(o)
    assert np.isfinite(out).all()
    np.testing.assert_allclose(out[..., 0].sum(axis=1), 1.0, atol=1e-5)
    werr = np.abs(out[..., 0] - pi)
    assert np.median(werr) < 1e-7 and werr.max() < 1e-5, (np.median(werr), werr.max())
    dots = (out[..., 1:4] * mu).sum(axis=2)
    assert dots.min() > 1 - 1e-6, dots.min()
    kerr = np.abs(out[..., 4] / k - 1)
    say(f"  mixture: weight error median {np.median(werr):.2e} max {werr.max():.2e}; "
        f"direction dot min {dots.min():.6f}; kappa error median {np.median(kerr):.2e} "
        f"max {kerr.max():.2e}")
    assert np.median(kerr) < 1e-6 and kerr.max() < 1e-5, (np.median(kerr), kerr.max())


@pytest.mark.parametrize("objective", [0, 1])
def test_gradient_matches_cpu(guided_scene, objective):
    """objective 0: KL to the vMF mixture; 1: KL to the MIS density
    0.5 vMF + 0.5 cos/pi (the fixture's mix is 0.5)."""
    s = guided_scene
    L = lib()
    check(L.rend_hip_guide_objective(s.rd, objective))
    p = random_params(3)
    check(L.rend_hip_guide_params(s.rd, ptr(p), 1))
    pos, nrm = points(s, 512, 4)
    rng = np.random.default_rng(5)
    rec = np.zeros(len(pos), G.RECORD)
    rec["pos"], rec["nrm"] = pos, nrm
    d = rng.normal(size=(len(pos), 3))
    rec["dir"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rec["target"] = rng.uniform(0.1, 2.0, len(pos))
    g = np.zeros(G.NPARAM, np.float32)
    loss = C.c_float(0)
    say(f"guide gradient: 512 records, objective {objective}")
    try:
        check(L.rend_hip_guide_grad_host(s.rd, ptr(rec), len(rec), ptr(g), C.byref(loss)))
    finally:
        L.rend_hip_guide_objective(s.rd, 0)
    gmin, ginv = G.scene_box(s.aabbs())
    bsdf = 0.5 if objective else 0.0
    want = G.grad(p.astype(np.float64), rec, gmin, ginv, bsdf)
    # every output unit (a lobe's logit, mean direction or log concentration)
    # against the restatement at the kernels' precision
    emul = G.grad(p.astype(np.float64), rec, gmin, ginv, bsdf, emulate=True)
    cos, ratio = G.block_agreement(g, emul)
    used = np.arange(5 * G.LOBES)
    say(f"  per unit vs emulated: min cosine {cos[used].min():.5f}, max |ratio - 1| "
        f"{np.abs(ratio[used] - 1).max():.4f} (unit {used[np.argmax(np.abs(ratio[used] - 1))]})")
    cx, rx = G.block_agreement(g, want)
    say(f"  per unit vs exact:    min cosine {cx[used].min():.5f}, max |ratio - 1| "
        f"{np.abs(rx[used] - 1).max():.4f}")
    bad = used[(cos[used] < G.UNIT_COS) | (np.abs(ratio[used] - 1) > G.UNIT_RATIO)]
    assert bad.size == 0, [(int(u), float(cos[u]), float(ratio[u])) for u in bad]
    assert np.abs(G.unit_blocks(g)[5 * G.LOBES:]).max() == 0.0
    # each parameter block against the restatement at the kernels' precision
    # (and, for the record, against the exact one); the hash-grid gradient is
    # summed by BF16 atomics in any order, hence its looser bound
    for name, a, b, cmin, rtol in (("W1", G.OFS_W1, G.OFS_B1, 0.99999, 5e-4),
                                   ("b1", G.OFS_B1, G.OFS_W2, 0.99999, 5e-4),
                                   ("W2", G.OFS_W2, G.OFS_B2, 0.99999, 5e-4),
                                   ("b2", G.OFS_B2, G.OFS_W3, 0.99999, 5e-4),
                                   ("W3", G.OFS_W3, G.OFS_B3, 0.99999, 5e-4),
                                   ("b3", G.OFS_B3, G.OFS_W4, 0.99999, 5e-4),
                                   ("W4", G.OFS_W4, G.OFS_B4, 0.99999, 5e-4),
                                   ("b4", G.OFS_B4, G.NPARAM, 0.99999, 5e-4),
                                   ("grid", 0, G.N_GRID, 0.9999, 3e-3)):
        u, v, x = g[a:b].astype(np.float64), emul[a:b], want[a:b]
        cos = u @ v / (np.linalg.norm(u) * np.linalg.norm(v) + 1e-30)
        ratio = np.linalg.norm(u) / (np.linalg.norm(v) + 1e-30)
        cx = u @ x / (np.linalg.norm(u) * np.linalg.norm(x) + 1e-30)
        say(f"  {name}: cosine {cos:.6f} norm ratio {ratio:.5f} (vs exact: cosine {cx:.4f})")
        assert cos > cmin and abs(ratio - 1) < rtol, name
    assert abs(loss.value - G.loss(p.astype(np.float64), rec, gmin, ginv, bsdf)) < 0.05 * abs(
        loss.value) + 1e-3


@pytest.mark.parametrize("target,depth,share", [(0, 1, 1), (1, 2, 0), (0, 5, 1)])
def test_guided_rendering_is_unbiased(target, depth, share):
    """Plain and guided (trained online, one-sample MIS with the BSDF) images
    converge to the same expectation, for either training target, guided
    depth and with or without the per-pixel depth-0 proposal.  The box scene
    keeps Monte Carlo noise low enough for a tight bound: 3072 spp per
    image, 4x4 block means."""
    s = Scene(config=1, detail=0.25, width=64, height=64)
    L = lib()
    try:
        check(L.rend_hip_guide_enable(s.rd, 0.5, 2, 4, 1e-2, 11), "guide enable")
        check(L.rend_hip_guide_sampling(s.rd, depth, share), "sampling")
        check(L.rend_hip_guide_target(s.rd, target), "target")
        say(f"guided render 96 frames x 32 spp (box), target {target} depth {depth} share {share}")
        img_g, _ = render_spp(s, 96)
        info = (C.c_float * 4)()
        check(L.rend_hip_guide_info(s.rd, info))
        assert info[2] > 0, "the network was trained"
        L.rend_hip_guide_disable(s.rd)
        say("plain render 96 frames x 32 spp (box)")
        img_p, _ = render_spp(s, 96)
    finally:
        s.close()
    blk = lambda x: x.reshape(16, 4, 16, 4, 3).mean(axis=(1, 3))  # noqa: E731
    err = rel_mse(blk(img_g), blk(img_p))
    say(f"block relMSE guided vs plain {err:.2e}, mean {img_g.mean():.4f} vs {img_p.mean():.4f}")
    assert err < 1e-5
    assert abs(img_g.mean() / img_p.mean() - 1) < 2e-3


def render_spp(s, frames):
    v = s.view()
    L = s.L
    # 32 spp per frame through the views API
    from nodebleed_amd import RendError  # noqa: F401
    import nodebleed_amd as nb
    cam = s.camera()
    eye = cam[:3]
    tgt = eye + cam[9:12]
    view = np.zeros(1, dtype=[("eye", "<f4", 3), ("target", "<f4", 3), ("vfov", "<f4"),
                              ("focdist", "<f4"), ("focangle", "<f4"), ("w", "<u4"),
                              ("h", "<u4"), ("spp", "<u4")])
    view["eye"], view["target"] = eye, tgt
    view["vfov"] = np.degrees(2 * np.arctan(cam[16] / 2 / cam[15]))
    view["focdist"], view["w"], view["h"], view["spp"] = cam[15], v.width, v.height, 32
    acc_d = L.rend_hip_malloc(v.width * v.height * 3 * 4)
    zero = np.zeros(v.width * v.height * 3, np.float32)
    check(L.rend_hip_memcpy_htod(acc_d, ptr(zero), zero.nbytes))
    check(L.rend_hip_pt_config(5, 3, 0))
    rays = C.c_ulonglong(0)
    for f in range(frames):
        check(L.rend_hip_render_views(s.rd, ptr(view), 1, f * 32, acc_d, None,
                                      C.byref(rays)), "render views")
    out = np.zeros_like(zero)
    check(L.rend_hip_memcpy_dtoh(ptr(out), acc_d, out.nbytes))
    L.rend_hip_free(acc_d)
    del nb
    return out.reshape(v.height, v.width, 3) / (frames * 32), rays.value


def test_training_learns_a_known_lobe(guided_scene):
    """Online training pipeline on the GPU (sampled records, MFMA backward,
    hash-grid scatter, Adam, mirror refresh): fresh records every step, at
    random points of a region, with directions from one vMF lobe (kappa 40),
    make the rendering-time mixture put high density on that lobe's
    direction at new points of the region."""
    s = guided_scene
    L = lib()
    check(L.rend_hip_guide_enable(s.rd, 0.5, 2, 2, 1e-2, 21), "guide re-init")
    check(L.rend_hip_guide_training(s.rd, 32768, 1e-2, 1e-2, 1), "schedule")
    rng = np.random.default_rng(8)
    ab = s.aabbs()
    lo, hi = ab["min"].min(axis=0), ab["max"].max(axis=0)
    c = lo + (hi - lo) * np.array([0.3, 0.2, 0.6])
    mu = np.array([0.6, 0.8, 0.0])
    k = 40.0

    def batch(n):
        pos = c + rng.uniform(-0.1, 0.1, (n, 3)) * (hi - lo)
        u = rng.random(n)
        w = 1 + np.log(u + (1 - u) * np.exp(-2 * k)) / k
        phi = rng.uniform(0, 2 * np.pi, n)
        t1 = np.array([0.0, 0.0, 1.0])
        t2 = np.cross(mu, t1)
        sw = np.sqrt(np.maximum(0, 1 - w * w))
        d = w[:, None] * mu + sw[:, None] * (np.cos(phi)[:, None] * t1 +
                                               np.sin(phi)[:, None] * t2)
        rec = np.zeros(n, G.RECORD)
        rec["pos"], rec["nrm"], rec["dir"] = pos, (0, 1, 0), d
        rec["pdf"], rec["target"] = 1.0, 1.0
        return rec

    loss = C.c_float(0)
    say("guide training on a synthetic lobe: 150 steps of fresh records")
    for _ in range(150):
        rec = batch(32768)
        check(L.rend_hip_guide_train_host(s.rd, ptr(rec), len(rec), 1, C.byref(loss)))
    pos = batch(256)["pos"]
    pn = np.ascontiguousarray(np.concatenate([pos, np.tile([0, 1.0, 0], (256, 1))], 1),
                              np.float32)
    mix = np.zeros((256, G.LOBES, 5), np.float32)
    check(L.rend_hip_guide_mix_host(s.rd, ptr(pn), 256, ptr(mix)))
    pi, m, kap = mix[..., 0], mix[..., 1:4], mix[..., 4]

    def density(dirn):
        cth = (m * dirn).sum(axis=2)
        return (pi * kap / (2 * np.pi * -np.expm1(-2 * kap)) * np.exp(kap * (cth - 1))).sum(axis=1)

    p_mu, p_opp = density(mu), density(-mu)
    say(f"loss {loss.value:.3f}; density at the lobe {np.median(p_mu):.2f}, opposite "
        f"{np.median(p_opp):.4f}, uniform {1 / (4 * np.pi):.4f}")
    # a vMF with kappa 40 has density 6.4 at its mean; uniform 0.08
    assert np.median(p_mu) > 3.0 and np.median(p_opp) < 0.05


def test_record_cdf_matches_cumsum():
    """Fixed-point prefix sums of training-record targets (tile sums, scan of
    the tile sums, tile scans) equal an exact integer cumulative sum of
    floor(v * 2^(60 - e)), across tile boundaries and with zero runs; the
    scale keeps the total below 2^62."""
    L = lib()
    rng = np.random.default_rng(4)
    for n in (1, 100, 4095, 4096, 4097, 70_001, (1 << 20) + 3):
        v = rng.exponential(1.0, n).astype(np.float32)
        v[rng.random(n) < 0.3] = 0.0
        v[0] = 0.0
        wsum = np.float32(v.astype(np.float64).sum())
        say(f"record cdf n={n}")
        out = np.zeros(n, np.uint64)
        check(L.rend_hip_guide_cdf_host(ptr(v), n, wsum, ptr(out)), "cdf")
        e = np.frexp(max(float(wsum), 1e-30))[1]
        q = np.floor(v.astype(np.float64) * 2.0 ** (60 - e)).astype(np.uint64)
        want = np.cumsum(q, dtype=np.uint64)
        assert np.array_equal(out, want), n
        assert int(want[-1]) < 1 << 62


def test_importance_sampled_training_learns_a_known_lobe(guided_scene):
    """Records with uniformly random directions whose targets follow one vMF
    lobe (kappa 40): drawing records in proportion to target must teach the
    mixture that lobe, as the uniformly drawn, target-weighted steps do."""
    s = guided_scene
    L = lib()
    check(L.rend_hip_guide_enable(s.rd, 0.5, 2, 2, 1e-2, 22), "guide re-init")
    check(L.rend_hip_guide_training(s.rd, 32768, 1e-2, 1e-2, 1), "schedule")
    rng = np.random.default_rng(9)
    ab = s.aabbs()
    lo, hi = ab["min"].min(axis=0), ab["max"].max(axis=0)
    c = lo + (hi - lo) * np.array([0.7, 0.4, 0.3])
    mu = np.array([0.0, 0.6, -0.8])
    k = 40.0
    loss = C.c_float(0)
    say("guide training, records drawn by target: 150 steps")
    for _ in range(150):
        n = 65536
        d = rng.normal(size=(n, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        rec = np.zeros(n, G.RECORD)
        rec["pos"] = c + rng.uniform(-0.1, 0.1, (n, 3)) * (hi - lo)
        rec["nrm"], rec["dir"], rec["pdf"] = (0, 1, 0), d, 1 / (4 * np.pi)
        rec["target"] = np.exp(k * (d @ mu - 1)) * 4 * np.pi
        check(L.rend_hip_guide_train_host(s.rd, ptr(rec), n, 1, C.byref(loss)))
    pos = c + rng.uniform(-0.1, 0.1, (256, 3)) * (hi - lo)
    pn = np.ascontiguousarray(np.concatenate([pos, np.tile([0, 1.0, 0], (256, 1))], 1),
                              np.float32)
    mix = np.zeros((256, G.LOBES, 5), np.float32)
    check(L.rend_hip_guide_mix_host(s.rd, ptr(pn), 256, ptr(mix)))
    pi, m, kap = mix[..., 0], mix[..., 1:4], mix[..., 4]

    def density(dirn):
        cth = (m * dirn).sum(axis=2)
        return (pi * kap / (2 * np.pi * -np.expm1(-2 * kap)) * np.exp(kap * (cth - 1))).sum(axis=1)

    p_mu, p_opp = density(mu), density(-mu)
    say(f"loss {loss.value:.3f}; density at the lobe {np.median(p_mu):.2f}, opposite "
        f"{np.median(p_opp):.4f}")
    assert np.median(p_mu) > 3.0 and np.median(p_opp) < 0.05


def test_warmup_schedule_counts_adam_steps():
    """rend_hip_guide_warmup: the first `frames` frames after enabling train
    with the warm-up step count, later frames with the per-frame count; with
    training frozen (0 steps per frame) no step is taken, warm-up or not."""
    s = Scene(config=5, detail=0.1, width=64, height=36, spp=4)
    L = lib()
    info = (C.c_float * 4)()
    try:
        for warm, want in (((16, 64), 16 * 64 + 4 * 8), ((0, 0), 20 * 8), ((3, 4), 20 * 8)):
            check(L.rend_hip_guide_enable(s.rd, 0.5, 4, 8, 1e-2, 7), "guide enable")
            check(L.rend_hip_guide_warmup(s.rd, *warm), "warm-up")
            say(f"warm-up {warm}: 20 frames")
            s.render(frames=20, maxdepth=3)
            check(L.rend_hip_guide_info(s.rd, info), "info")
            assert int(info[2]) == want, (warm, info[2])
        check(L.rend_hip_guide_enable(s.rd, 0.5, 4, 0, 1e-2, 7), "guide enable frozen")
        s.render(frames=3, maxdepth=3)
        check(L.rend_hip_guide_info(s.rd, info), "info")
        assert int(info[2]) == 0
        assert L.rend_hip_guide_warmup(s.rd, 1, 5000) != 0
    finally:
        L.rend_hip_guide_disable(s.rd)
        s.close()
