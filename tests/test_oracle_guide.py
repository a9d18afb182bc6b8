"""Pins the FP64 guide oracle (oracle/or_guide.c) before it checks the device:
vMF normalisation by quadrature, Jakob's sampler against the closed-form CDF
and mean cosine, the hand-derived gradient against central differences of
the loss, and the input encoding at points where its values are known."""
import numpy as np
import pytest

import oracle

NP = oracle.GUIDE_N_PARAMS
IN = [32, 64, 64, 64, 64]
OUT = [64, 64, 64, 64, 48]


def layer_offsets():
    off, res = 0, []
    for i, o in zip(IN, OUT):
        res.append((off, off + o * i))
        off += o * i + o
    assert off == NP
    return res


def random_params(rng, out_scale=0.3):
    p = np.zeros(NP)
    for l, (wo, bo) in enumerate(layer_offsets()):
        lim = np.sqrt(6.0 / IN[l]) * (out_scale if l == 4 else 1.0)
        p[wo:bo] = rng.uniform(-lim, lim, OUT[l] * IN[l])
        p[bo:bo + OUT[l]] = rng.normal(0, 0.05, OUT[l])
    return p


def unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def random_samples(rng, n, bounds):
    s = np.zeros(n, oracle.GUIDE_SAMPLE_DTYPE)
    lo, hi = np.array(bounds[:3]), np.array(bounds[3:])
    s["pos"] = rng.uniform(lo, hi, (n, 3))
    nrm = unit(rng.normal(size=(n, 3)))
    s["normal"] = nrm
    s["wo"] = unit(rng.normal(size=(n, 3)))
    wi = unit(rng.normal(size=(n, 3)))
    wi = np.where((np.sum(wi * nrm, 1) < 0)[:, None], -wi, wi)  # upper hemisphere
    s["wi"] = wi
    s["weight"] = rng.exponential(1.0, n)
    return s


def fibonacci_sphere(n):
    i = np.arange(n) + 0.5
    z = 1 - 2 * i / n
    phi = np.pi * (1 + 5 ** 0.5) * i
    r = np.sqrt(1 - z * z)
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], 1)


BOUNDS = [-2.0, -1.0, -3.0, 2.0, 3.0, 1.0]


@pytest.mark.parametrize("kappa", [0.05, 1.0, 10.0, 100.0])
def test_vmf_pdf_integrates_to_one(kappa):
    rng = np.random.default_rng(1)
    mu = unit(rng.normal(size=3))
    pts = fibonacci_sphere(200_000)
    vals = np.array([oracle.vmf_pdf(mu, kappa, w) for w in pts])
    integral = vals.mean() * 4 * np.pi
    assert abs(integral - 1.0) < 2e-3


@pytest.mark.parametrize("kappa", [0.5, 5.0, 50.0])
def test_vmf_sampler_matches_density(kappa):
    rng = np.random.default_rng(int(kappa * 10))
    mu = unit(rng.normal(size=3))
    n = 20_000
    xs = rng.random((n, 2))
    w = np.array([oracle.vmf_sample(mu, kappa, a, b) for a, b in xs])
    assert np.allclose(np.linalg.norm(w, axis=1), 1.0, atol=1e-12)
    c = w @ mu
    # mean cosine coth(k) - 1/k
    mc = 1 / np.tanh(kappa) - 1 / kappa
    assert abs(c.mean() - mc) < 4 * c.std() / np.sqrt(n) + 1e-12
    # Kolmogorov-Smirnov against P(c <= t) = (e^{k(t-1)} - e^{-2k}) / (1 - e^{-2k})
    cs = np.sort(c)
    cdf = (np.exp(kappa * (cs - 1)) - np.exp(-2 * kappa)) / (1 - np.exp(-2 * kappa))
    emp = np.arange(1, n + 1) / n
    assert np.max(np.abs(cdf - emp)) < 1.63 / np.sqrt(n)  # 1% level
    # azimuth uniform: the component orthogonal to mu averages to zero
    perp = w - c[:, None] * mu
    assert np.all(np.abs(perp.mean(0)) < 4 * perp.std(0) / np.sqrt(n) + 1e-9)


def test_encoding_known_points():
    s = np.zeros(1, oracle.GUIDE_SAMPLE_DTYPE)
    s["pos"] = BOUNDS[:3]  # u = 0 on every axis
    s["normal"] = (0.0, 0.6, 0.8)
    s["wo"] = (1.0, 0.0, 0.0)
    f = oracle.guide_encode(BOUNDS, s)
    for d in range(3):
        assert np.allclose(f[8 * d:8 * d + 8:2], 0.0, atol=1e-7)
        assert np.allclose(f[8 * d + 1:8 * d + 8:2], 1.0, atol=1e-7)
    assert np.allclose(f[24:27], (0, 0.6, 0.8), atol=1e-7)
    assert np.allclose(f[27:30], (1, 0, 0))
    assert f[30] == 1.0 and f[31] == 0.0
    # sin/cos(2^(k+1) pi u), k = 0..3: periods 1, 1/2, 1/4 and 1/8 of the box
    s["pos"] = [(a + b) / 2 for a, b in zip(BOUNDS[:3], BOUNDS[3:])]  # u = 1/2
    f = oracle.guide_encode(BOUNDS, s)
    want = [0, -1, 0, 1, 0, 1, 0, 1]
    for d in range(3):
        assert np.allclose(f[8 * d:8 * d + 8], want, atol=1e-7)
    s["pos"] = [a + 0.25 * (b - a) for a, b in zip(BOUNDS[:3], BOUNDS[3:])]  # u = 1/4
    f = oracle.guide_encode(BOUNDS, s)
    want = [1, 0, 0, -1, 0, 1, 0, 1]
    assert np.allclose(f[0:8], want, atol=1e-7)
    s["pos"] = [a + 0.125 * (b - a) for a, b in zip(BOUNDS[:3], BOUNDS[3:])]  # u = 1/8
    f = oracle.guide_encode(BOUNDS, s)
    r = np.sqrt(0.5)
    want = [r, r, 1, 0, 0, -1, 0, 1]
    assert np.allclose(f[0:8], want, atol=1e-7)


def test_head_is_a_normalised_mixture():
    rng = np.random.default_rng(3)
    p = random_params(rng, out_scale=1.0)
    s = random_samples(rng, 8, BOUNDS)
    raw, head = oracle.guide_forward(p, BOUNDS, s)
    pts = fibonacci_sphere(100_000)
    for h in head:
        mu = h[:40].reshape(8, 5)
        assert np.allclose(np.linalg.norm(mu[:, :3], axis=1), 1.0)
        assert abs(mu[:, 4].sum() - 1.0) < 1e-12
        assert np.all(mu[:, 3] >= np.exp(-3) - 1e-12) and np.all(mu[:, 3] <= np.exp(7) + 1e-9)
        assert 0.1 <= h[40] <= 0.9
        # mixture density integrates to one (quadrature, lobes with kappa <= 200)
        if mu[:, 3].max() < 200:
            q = sum(m[4] * (lambda c, k: k / (2 * np.pi * -np.expm1(-2 * k)) * np.exp(k * (c - 1)))(pts @ m[:3], m[3])
                    for m in mu)
            assert abs(q.mean() * 4 * np.pi - 1.0) < 3e-3


def test_gradient_matches_central_differences():
    rng = np.random.default_rng(7)
    p = random_params(rng)
    s = random_samples(rng, 48, BOUNDS)
    g, sw = oracle.guide_grad(p, BOUNDS, s)
    assert abs(sw - s["weight"].astype(np.float64).sum()) < 1e-9
    offs = layer_offsets()
    picks = []
    for l, (wo, bo) in enumerate(offs):
        picks += list(rng.choice(np.arange(wo, bo), 10, replace=False))
        picks += list(rng.choice(np.arange(bo, bo + OUT[l]), 4, replace=False))
    # every head component of the last layer's bias: mu, log kappa, logit, selection
    bo4 = offs[4][1]
    for lobe in (0, 5):
        g_, q0 = lobe // 2, (lobe % 2) * 5
        for c in range(5):
            q = q0 + c
            picks.append(bo4 + (q // 4) * 16 + 4 * g_ + q % 4)
    picks.append(bo4 + 34)
    for i in picks:
        h = 1e-6 * max(1.0, abs(p[i]))
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (oracle.guide_loss(pp, BOUNDS, s, frozen=p) - oracle.guide_loss(pm, BOUNDS, s, frozen=p)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(fd)), (i, fd, g[i])


def test_second_moment_selection_gradient():
    """Objective 1: the selection row's gradient is the central difference of
    sum w^2 pdf / (alpha q + (1 - alpha) b); every other parameter gets the
    gradient of the lobe term alone (objective 2), so the hidden layers see no
    delta from the selection row; and at pdf = the current mixture pdf the
    selection gradient equals the KL's times the weight."""
    rng = np.random.default_rng(17)
    p = random_params(rng, out_scale=1.0)
    s = random_samples(rng, 40, BOUNDS)
    s["pdf"] = rng.uniform(0.05, 2.0, len(s))
    g1, _ = oracle.guide_grad(p, BOUNDS, s, objective=1)
    g0, _ = oracle.guide_grad(p, BOUNDS, s, objective=0)
    offs = layer_offsets()
    wo4, bo4 = offs[4]
    sel = [bo4 + 34] + list(wo4 + 64 * 34 + rng.choice(64, 12, replace=False))
    assert np.any(g1[sel] != g0[sel])
    picks = list(sel)
    for l, (wo, bo) in enumerate(offs):
        picks += list(rng.choice(np.arange(wo, bo), 8, replace=False))
        picks += list(rng.choice(np.arange(bo, bo + OUT[l]), 3, replace=False))
    for i in picks:
        h = 1e-6 * max(1.0, abs(p[i]))
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        obj = 1 if i in sel else 2
        fd = (oracle.guide_loss(pp, BOUNDS, s, frozen=p, objective=obj)
              - oracle.guide_loss(pm, BOUNDS, s, frozen=p, objective=obj)) / (2 * h)
        assert abs(fd - g1[i]) <= 1e-5 * max(1.0, abs(fd)), (i, obj, fd, g1[i])
    # d/dalpha (w^2 ps / pm) = -w (q - b) / pm * (w ps / pm): with ps = pm
    # (a sample drawn from the current mixture) it is the KL term times w
    _, head = oracle.guide_forward(p, BOUNDS, s)
    one = s[:1].copy()
    q = sum(head[0][5 * j + 4] * oracle.vmf_pdf(head[0][5 * j:5 * j + 3], head[0][5 * j + 3], one["wi"][0])
            for j in range(8))
    b = max(0.0, float(np.dot(one["normal"][0], one["wi"][0]))) / np.pi
    one["pdf"] = head[0][40] * q + (1 - head[0][40]) * b
    a1, _ = oracle.guide_grad(p, BOUNDS, one, objective=1)
    a0, _ = oracle.guide_grad(p, BOUNDS, one, objective=0)
    w = float(one["weight"][0])
    assert np.isclose(a1[bo4 + 34], w * a0[bo4 + 34], rtol=1e-5, atol=1e-12)


def test_unused_outputs_get_no_gradient():
    rng = np.random.default_rng(9)
    p = random_params(rng)
    s = random_samples(rng, 32, BOUNDS)
    g, _ = oracle.guide_grad(p, BOUNDS, s)
    bo4 = layer_offsets()[4][1]
    unused = [35, 38, 39, 42, 43, 46, 47]
    assert np.all(g[bo4 + np.array(unused)] == 0.0)
    wo4 = layer_offsets()[4][0]
    for r in unused:
        assert np.all(g[wo4 + 64 * r: wo4 + 64 * (r + 1)] == 0.0)


def test_bf16_emulation_stays_close():
    rng = np.random.default_rng(11)
    p = random_params(rng)
    s = random_samples(rng, 64, BOUNDS)
    g64, _ = oracle.guide_grad(p, BOUNDS, s, emulate_bf16=False)
    g16, _ = oracle.guide_grad(p, BOUNDS, s, emulate_bf16=True)
    cos = g64 @ g16 / np.linalg.norm(g64) / np.linalg.norm(g16)
    assert cos > 0.995
    raw64, _ = oracle.guide_forward(p, BOUNDS, s, False)
    raw16, _ = oracle.guide_forward(p, BOUNDS, s, True)
    assert np.max(np.abs(raw64 - raw16)) < 0.05 * max(1.0, np.max(np.abs(raw64)))
