"""Learned path guide on the GPU against the FP64 oracle (oracle/or_guide.c):
network evaluation and gradients (BF16 emulated and real-valued), the Adam
update, learning a known target, training-sample collection, unbiasedness
of guided renders, and the gain over the plain sampler."""
import numpy as np
import pytest

import devilray_amd as dr
import oracle
from devilray_amd import scenes

pytestmark = pytest.mark.gpu

NP = dr.GUIDE_N_PARAMS
IN = [32, 64, 64, 64, 64]
OUT = [64, 64, 64, 64, 48]
BOUNDS = np.array([-2.0, -1.0, -3.0, 2.0, 3.0, 1.0], np.float32)


def layer_slices():
    off, res = 0, []
    for i, o in zip(IN, OUT):
        res.append((slice(off, off + o * i), slice(off + o * i, off + o * i + o)))
        off += o * i + o
    return res


def random_params(rng, out_scale=0.5):
    p = np.zeros(NP, np.float32)
    for l, (ws, bs) in enumerate(layer_slices()):
        lim = np.sqrt(6.0 / IN[l]) * (out_scale if l == 4 else 1.0)
        p[ws] = rng.uniform(-lim, lim, OUT[l] * IN[l])
        p[bs] = rng.normal(0, 0.05, OUT[l])
    return p


def unit(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def random_samples(rng, n, bounds=BOUNDS):
    s = np.zeros(n, dr.GUIDE_SAMPLE_DTYPE)
    s["pos"] = rng.uniform(bounds[:3], bounds[3:], (n, 3))
    nrm = unit(rng.normal(size=(n, 3)))
    s["normal"] = nrm
    s["wo"] = unit(rng.normal(size=(n, 3)))
    wi = unit(rng.normal(size=(n, 3)))
    s["wi"] = np.where((np.sum(wi * nrm, 1) < 0)[:, None], -wi, wi)
    s["weight"] = rng.exponential(1.0, n)
    return s


def mixture_pdf(head, w):
    """q(w) of decoded heads (n x 44) at directions w (n x 3)."""
    lob = head[:, :40].reshape(-1, 8, 5).astype(np.float64)
    mu, k, pi = lob[..., :3], lob[..., 3], lob[..., 4]
    c = np.einsum("nkd,nd->nk", mu, w)
    f = k / (2 * np.pi * -np.expm1(-2 * k)) * np.exp(k * (c - 1))
    return np.sum(pi * f, 1)


@pytest.fixture(scope="module")
def guide(gpu_ctx):
    g = dr.Guide(gpu_ctx, BOUNDS, capacity=1 << 16, seed=5)
    yield g
    g.close()


def test_eval_matches_oracle(guide):
    rng = np.random.default_rng(21)
    p = random_params(rng)
    guide.set_params(p)
    assert np.array_equal(guide.get_params(), p)
    s = random_samples(rng, 2000)
    h = guide.eval(s)
    _, h16 = oracle.guide_forward(p.astype(np.float64), BOUNDS, s, emulate_bf16=True)
    _, h64 = oracle.guide_forward(p.astype(np.float64), BOUNDS, s, emulate_bf16=False)
    lob, lob16 = h[:, :40].reshape(-1, 8, 5), h16[:, :40].reshape(-1, 8, 5)
    cosang = np.sum(lob[..., :3] * lob16[..., :3], -1)
    print(f"min cos {np.min(cosang):.7f}, mean cos {np.mean(cosang):.9f}, "
          f"max kappa rel {np.max(np.abs(lob[..., 3] / lob16[..., 3] - 1)):.2e}, "
          f"max weight abs {np.max(np.abs(lob[..., 4] - lob16[..., 4])):.2e}", flush=True)
    # the device sums in FP32 in MFMA order, the oracle in FP64: a hidden
    # activation next to a BF16 rounding boundary can round one BF16 ulp
    # (2^-8) apart, which turns a lobe's mean direction by up to about a degree
    assert np.min(cosang) > 1 - 5e-4
    assert np.mean(cosang) > 1 - 1e-6
    assert np.max(np.abs(lob[..., 3] / lob16[..., 3] - 1)) < 5e-3
    assert np.max(np.abs(lob[..., 4] - lob16[..., 4])) < 2e-3
    assert np.max(np.abs(h[:, 40] - h16[:, 40])) < 2e-3
    assert np.mean(np.abs(lob[..., 4] - lob16[..., 4])) < 1e-4
    # against the real-valued network: the BF16 rounding only
    assert np.mean(np.abs(lob[..., 4] - h64[:, :40].reshape(-1, 8, 5)[..., 4])) < 5e-3
    assert np.all((h[:, 40] >= 0.1) & (h[:, 40] <= 0.9))


def test_gradient_matches_oracle(guide):
    rng = np.random.default_rng(22)
    p = random_params(rng)
    guide.set_params(p)
    s = random_samples(rng, 700)  # not a multiple of the 64-sample tile
    s["weight"][::7] = 0.0
    g, sw = guide.gradient(s)
    g16, sw16 = oracle.guide_grad(p.astype(np.float64), BOUNDS, s, emulate_bf16=True)
    g64, _ = oracle.guide_grad(p.astype(np.float64), BOUNDS, s, emulate_bf16=False)
    assert abs(sw - sw16) <= 1e-5 * sw16
    for l, (ws, bs) in enumerate(layer_slices()):
        for sl in (ws, bs):
            err = np.linalg.norm(g[sl] - g16[sl]) / np.linalg.norm(g16[sl])
            assert err < 2e-2, (l, err)
    cos = g @ g64 / np.linalg.norm(g) / np.linalg.norm(g64)
    assert cos > 0.995
    # deterministic: the same call gives the same bits
    g2, _ = guide.gradient(s)
    assert np.array_equal(g, g2)


def test_gradient_is_additive_over_tiles(guide):
    """A batch of more than 256 tiles makes each training block loop over
    several tiles; its gradient equals the sum of the gradients of 4096-sample
    parts (one tile per block) up to FP32 summation order."""
    rng = np.random.default_rng(26)
    guide.set_params(random_params(rng))
    s = random_samples(rng, 2 * 32768 + 777)
    s["weight"][::11] = 0.0
    g, sw = guide.gradient(s)
    parts = [guide.gradient(s[k:k + 4096]) for k in range(0, len(s), 4096)]
    gs = np.sum([q[0] for q in parts], axis=0, dtype=np.float64)
    sws = float(sum(q[1] for q in parts))
    assert abs(sw - sws) <= 1e-4 * sws
    err = np.linalg.norm(g - gs) / np.linalg.norm(gs)
    assert err < 1e-4, err


def test_second_moment_selection_gradient_matches_oracle(guide):
    """DR_GUIDE_SELECT_VARIANCE: the device gradient equals the oracle's
    second-moment objective (selection row trained alone) layer by layer."""
    rng = np.random.default_rng(24)
    p = random_params(rng)
    guide.set_params(p)
    s = random_samples(rng, 700)
    s["weight"][::7] = 0.0
    s["pdf"] = rng.uniform(0.05, 2.0, len(s))
    guide.set_selection_objective(dr.DR_GUIDE_SELECT_VARIANCE)
    try:
        g, _ = guide.gradient(s)
    finally:
        guide.set_selection_objective(dr.DR_GUIDE_SELECT_KL)
    g16, _ = oracle.guide_grad(p.astype(np.float64), BOUNDS, s, emulate_bf16=True, objective=1)
    g0, _ = oracle.guide_grad(p.astype(np.float64), BOUNDS, s, emulate_bf16=True, objective=0)
    for l, (ws, bs) in enumerate(layer_slices()):
        for sl in (ws, bs):
            err = np.linalg.norm(g[sl] - g16[sl]) / np.linalg.norm(g16[sl])
            assert err < 2e-2, (l, err)
    assert np.linalg.norm(g16 - g0) > 0.05 * np.linalg.norm(g0)  # the objectives differ here
    with pytest.raises(dr.DevilRayError):
        guide.set_selection_objective(7)


def test_adam_step(guide):
    rng = np.random.default_rng(23)
    p = random_params(rng)
    guide.set_params(p)
    s = random_samples(rng, 1000)
    lr, b1, b2, eps = 1e-2, 0.9, 0.99, 1e-8
    m = np.zeros(NP, np.float64)
    v = np.zeros(NP, np.float64)
    w = p.astype(np.float64)
    for t in (1, 2, 3):
        g, sw = guide.gradient(s)
        gn = g.astype(np.float64) / sw
        m = b1 * m + (1 - b1) * gn
        v = b2 * v + (1 - b2) * gn * gn
        w = w - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
        guide.train_step(s)
        got = guide.get_params().astype(np.float64)
        assert np.max(np.abs(got - w)) < 1e-5, t
        w = got  # continue from the device state (FP32 rounding)


def test_learning_rate_schedule(gpu_ctx):
    """dr_guide_set_learning_rate: later steps use the new step size (Adam
    closed form with a per-step lr), and invalid rates are refused."""
    g = dr.Guide(gpu_ctx, BOUNDS, capacity=1 << 16, seed=6)
    rng = np.random.default_rng(29)
    p = random_params(rng)
    g.set_params(p)
    s = random_samples(rng, 1000)
    b1, b2, eps = 0.9, 0.99, 1e-8
    m = np.zeros(NP, np.float64)
    v = np.zeros(NP, np.float64)
    w = p.astype(np.float64)
    for t, lr in ((1, 1e-2), (2, 2.5e-3), (3, 5e-4)):
        g.set_learning_rate(lr)
        grad, sw = g.gradient(s)
        gn = grad.astype(np.float64) / sw
        m = b1 * m + (1 - b1) * gn
        v = b2 * v + (1 - b2) * gn * gn
        w = w - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
        g.train_step(s)
        got = g.get_params().astype(np.float64)
        assert np.max(np.abs(got - w)) < 1e-5, (t, lr)
        w = got
    for bad in (0.0, -1e-3, float("nan"), 2.0):
        with pytest.raises(dr.DevilRayError):
            g.set_learning_rate(bad)
    g.close()


def test_learns_a_known_target(gpu_ctx):
    """Samples weighted by a vMF target (kappa 20) per position: the trained
    mixture's cross-entropy against the target drops by more than 1 nat."""
    g = dr.Guide(gpu_ctx, BOUNDS, capacity=1 << 14, seed=9)
    rng = np.random.default_rng(24)
    centers = rng.uniform(BOUNDS[:3] * 0.5, BOUNDS[3:] * 0.5, (4, 3)).astype(np.float32)
    normals = unit(rng.normal(size=(4, 3)))
    targets = unit(normals + 0.6 * unit(rng.normal(size=(4, 3))))
    kap = 20.0

    def batch(n):
        k = rng.integers(0, 4, n)
        s = np.zeros(n, dr.GUIDE_SAMPLE_DTYPE)
        s["pos"] = centers[k]
        s["normal"] = normals[k]
        s["wo"] = normals[k]
        # cosine-distributed proposals about the normal
        u1, u2 = rng.random(n), rng.random(n)
        z = np.sqrt(1 - u1)
        r = np.sqrt(u1)
        t = unit(np.cross(normals[k], [0.3, 0.5, 0.8]))
        b = np.cross(normals[k], t)
        wi = z[:, None] * normals[k] + r[:, None] * (np.cos(2 * np.pi * u2)[:, None] * t +
                                                     np.sin(2 * np.pi * u2)[:, None] * b)
        s["wi"] = wi
        c = np.sum(wi * targets[k], 1)
        f = kap / (2 * np.pi * -np.expm1(-2 * kap)) * np.exp(kap * (c - 1))
        s["weight"] = f / (z / np.pi)
        return s

    def cross_entropy():
        xs = []
        for k in range(4):
            e = np.zeros(2000, dr.GUIDE_SAMPLE_DTYPE)
            e["pos"] = centers[k]
            e["normal"] = normals[k]
            e["wo"] = normals[k]
            wi = np.array([oracle.vmf_sample(targets[k], kap, a, b) for a, b in rng.random((2000, 2))])
            h = g.eval(e)
            q = mixture_pdf(h, wi)
            bb = np.maximum(0, wi @ normals[k]) / np.pi
            pm = h[:, 40] * q + (1 - h[:, 40]) * bb
            xs.append(-np.mean(np.log(pm)))
        return float(np.mean(xs))

    before = cross_entropy()
    for _ in range(300):
        g.train_step(batch(4096))
    after = cross_entropy()
    g.close()
    assert after < before - 1.0, (before, after)


def _cornell(gpu_ctx, res=48):
    sd = scenes.cornell_c0()
    sc = dr.Scene(gpu_ctx, sd.vertices, sd.indices, sd.tri_material, sd.dr_materials())
    sc.build_bvh()
    c = sd.camera
    cam = dr.camera_lookat(c["eye"], c["target"], c["up"], c["fov"], res, res)
    return sc, cam


def test_samples_record_their_pdf(gpu_ctx):
    """Each staged sample carries the density its direction was drawn from:
    cos/pi for BSDF sampling (no guide in use), and alpha q + (1 - alpha) cos/pi
    >= 0.1 cos/pi once the guide samples (alpha <= 0.9)."""
    print("[case] sample pdf", flush=True)
    sc, cam = _cornell(gpu_ctx, res=32)
    g = dr.Guide(gpu_ctx, sc.bounds(), capacity=1 << 14, seed=5)
    sc.render_guided(cam, g, dr.DR_GUIDE_COLLECT, spp=4, seed=1)
    s, _ = g.copy_samples()
    assert len(s) > 100
    b = np.maximum(np.sum(s["normal"].astype(np.float64) * s["wi"], 1), 0.0) / np.pi
    np.testing.assert_allclose(s["pdf"], b, rtol=1e-5, atol=1e-6)
    g.train_step()
    sc.render_guided(cam, g, dr.DR_GUIDE_USE | dr.DR_GUIDE_COLLECT, spp=4, seed=2)
    s, _ = g.copy_samples()
    b = np.maximum(np.sum(s["normal"].astype(np.float64) * s["wi"], 1), 0.0) / np.pi
    assert len(s) > 100 and np.all(s["pdf"] > 0.0)
    assert np.all(s["pdf"] >= 0.1 * b * (1.0 - 1e-4) - 1e-7)
    g.close()
    sc.close()


def test_collection(gpu_ctx):
    sc, cam = _cornell(gpu_ctx)
    g = dr.Guide(gpu_ctx, sc.bounds(), capacity=1 << 13, seed=1)
    sc.render_guided(cam, g, dr.DR_GUIDE_COLLECT, spp=4, seed=3)
    smp, total = g.copy_samples()
    assert total > len(smp) * 0.5 and len(smp) == min(total, 4 << 13)
    w = smp["weight"]
    assert np.all(np.isfinite(w)) and np.all(w >= 0) and np.mean(w > 0) > 0.2
    assert np.allclose(np.linalg.norm(smp["normal"], axis=1), 1, atol=1e-5)
    cs = np.sum(smp["normal"] * smp["wi"], 1)
    assert np.all(cs[w > 0] > 0)
    b = sc.bounds()
    assert np.all(smp["pos"] >= b[:3] - 1e-4) and np.all(smp["pos"] <= b[3:] + 1e-4)
    # the image of a collecting render equals the plain render (collection
    # draws no random numbers the estimator uses)
    a = sc.render(cam, spp=4, seed=3)
    a2 = sc.render_guided(cam, g, dr.DR_GUIDE_COLLECT, spp=4, seed=3)
    assert np.array_equal(a, a2)
    # a fraction keeps about that share
    g.set_options(0.25)
    sc.render_guided(cam, g, dr.DR_GUIDE_COLLECT, spp=4, seed=3)
    _, t2 = g.copy_samples(0)
    assert abs(t2 / total - 0.25) < 0.02
    g.close()
    sc.close()


def test_stage_order_is_reproducible(gpu_ctx):
    """Two collecting renders of the same samples stage the same samples in
    the same order, (path sample, vertex), although their threads take
    stage slots in whatever order the device schedules them (16,384 paths
    in 64 blocks of 256 per bounce)."""
    print("[case] stage order", flush=True)
    sc, cam = _cornell(gpu_ctx, res=64)
    g = dr.Guide(gpu_ctx, sc.bounds(), capacity=1 << 16, seed=1)
    runs = []
    for _ in range(2):
        sc.render_guided(cam, g, dr.DR_GUIDE_COLLECT, spp=4, seed=7)
        smp, total = g.copy_samples()
        assert total == len(smp) and 0 < total <= 4 << 16  # the stage did not fill
        runs.append(smp)
    assert runs[0].tobytes() == runs[1].tobytes()
    # the order hashes the path sample: its first half is spread over the
    # image like its second half (in raster order the first half would be the
    # top rows, whose first hits lie higher in the box)
    s = runs[0]
    h = len(s) // 2
    lo, hi = sc.bounds()[:3], sc.bounds()[3:]
    d = np.abs(s["pos"][:h].mean(0) - s["pos"][h:].mean(0)) / (hi - lo)
    assert np.all(d < 0.02), d
    g.close()
    sc.close()


def test_training_is_reproducible(gpu_ctx):
    """Online training twice from the same guide seed on the same render
    seeds: the networks and the images are the same bit for bit (the
    training ring is filled in sample order, and the collection fraction is
    sized so that the stage never fills)."""
    print("[case] training reproducible", flush=True)
    from devilray_amd.guiding import ProgressiveRenderer
    sc, cam = _cornell(gpu_ctx, res=48)
    out = []
    for _ in range(2):
        g = dr.Guide(gpu_ctx, sc.bounds(), capacity=1 << 14, seed=4)
        r = ProgressiveRenderer(sc, cam, g, spp_per_pass=8, max_depth=8, seed=21, train_steps=2)
        for _ in range(5):
            r.step()
        _, total = g.copy_samples(0)
        assert 0 < total <= 4 * g.capacity
        out.append((g.get_params(), r.image()))
        r.close()
        g.close()
    sc.close()
    assert out[0][0].tobytes() == out[1][0].tobytes()
    assert out[0][1].tobytes() == out[1][1].tobytes()


@pytest.mark.parametrize("trained", [False, True])
def test_guided_render_is_unbiased(gpu_ctx, trained):
    """Mean of 8 independent guided renders vs 8 plain renders (Cornell box):
    per-pixel z-scores of the difference are standard normal."""
    sc, cam = _cornell(gpu_ctx, res=32)
    g = dr.Guide(gpu_ctx, sc.bounds(), capacity=1 << 16, seed=2)
    if trained:
        for it in range(30):
            sc.render_guided(cam, g, dr.DR_GUIDE_USE * (it > 0) | dr.DR_GUIDE_COLLECT, spp=16, seed=100 + it)
            g.train_step()
    reps, spp = 8, 64
    plain = np.stack([sc.render(cam, spp=spp, seed=1000 + r) for r in range(reps)])
    guided = np.stack([sc.render_guided(cam, g, dr.DR_GUIDE_USE, spp=spp, seed=2000 + r) for r in range(reps)])
    pl = plain[..., :3].sum(-1) / spp
    gd = guided[..., :3].sum(-1) / spp
    se = np.sqrt(pl.var(0, ddof=1) / reps + gd.var(0, ddof=1) / reps)
    ok = se > 0
    z = (pl.mean(0) - gd.mean(0))[ok] / se[ok]
    assert abs(np.mean(z)) < 4 / np.sqrt(z.size) * 2
    assert np.mean(np.abs(z) > 3.5) < 0.02
    # whole-image mean
    tot_p, tot_g = pl.mean((1, 2)), gd.mean((1, 2))
    t = (tot_p.mean() - tot_g.mean()) / np.sqrt(tot_p.var(ddof=1) / reps + tot_g.var(ddof=1) / reps)
    assert abs(t) < 4.5
    g.close()
    sc.close()


@pytest.fixture(scope="module")
def indirect_scene(gpu_ctx):
    sd = scenes.many_lights(n_spheres=60, subdiv=3, n_emissive=5000, seed=4, extent=20.0, partition=True)
    sc = dr.Scene(gpu_ctx, sd.vertices, sd.indices, sd.tri_material, sd.dr_materials())
    sc.build_bvh()
    c = sd.camera
    cam = dr.camera_lookat(c["eye"], c["target"], c["up"], c["fov"], 96, 54)
    yield sc, cam
    sc.close()


def test_indirect_scene_is_indirect(indirect_scene):
    """Light reaching the camera side after one bounce (max depth 2: camera
    vertex + NEE) is under 10% of the total."""
    sc, cam = indirect_scene
    direct = sc.render(cam, spp=512, max_depth=2, rr_depth=2, seed=1)
    full = sc.render(cam, spp=512, max_depth=12, rr_depth=3, seed=1)
    d = direct[..., :3].sum() / direct[..., 3].sum()
    f = full[..., :3].sum() / full[..., 3].sum()
    print("direct share", d / f)
    assert f > 0 and d / f < 0.1


def test_guide_reduces_error_hidden_light(gpu_ctx):
    """Floor lit only through a small reflector (next-event estimation from
    the floor always fails): the online-trained guide learns to aim at the
    reflector, so at equal samples per pixel its squared error on the floor
    against a high-spp plain reference is far below the plain sampler's
    (median and mean over floor pixels, averaged over three seeds each)."""
    from devilray_amd.guiding import ProgressiveRenderer
    sd = scenes.hidden_light_reflector()
    sc = dr.Scene(gpu_ctx, sd.vertices, sd.indices, sd.tri_material, sd.dr_materials())
    sc.build_bvh()
    c = sd.camera
    cam = dr.camera_lookat(c["eye"], c["target"], c["up"], c["fov"], 64, 64)
    ref = sc.render(cam, spp=65536, max_depth=8, rr_depth=3, seed=77)
    ref = ref[..., :3] / ref[..., 3:4]

    # score the floor: pixels that see the emitter or the reflector carry
    # edge (anti-aliasing) noise that no bounce sampler changes
    floor = ref.mean(-1) < np.quantile(ref.mean(-1), 0.95)

    def err(guide, seed):
        r = ProgressiveRenderer(sc, cam, guide, spp_per_pass=16, max_depth=8, seed=seed, train_steps=16)
        for _ in range(16):
            r.step()
        e = ((r.image() - ref) ** 2).mean(-1)[floor]
        r.close()
        return np.median(e), e.mean()

    plain = [err(None, 900 + k) for k in range(3)]
    guided = []
    for k in range(3):
        g = dr.Guide(gpu_ctx, sc.bounds(), capacity=1 << 16, seed=3 + k)
        guided.append(err(g, 950 + k))
        g.close()
    sc.close()
    mp, mg = np.mean([p[0] for p in plain]), np.mean([q[0] for q in guided])
    ap, ag = np.mean([p[1] for p in plain]), np.mean([q[1] for q in guided])
    print(f"median sq. error plain {mp:.3e} guided {mg:.3e} ratio {mg / mp:.3f}; mean ratio {ag / ap:.3f}")
    assert mg < 0.25 * mp
    # The mean over the floor is carried by a few pixels where a cosine-sampled
    # bounce found the reflector with a small pdf, so it moves with the seeds
    # (0.24-0.50 of plain has been seen).
    assert ag < 0.75 * ap
