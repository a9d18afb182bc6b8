"""The reference's lighting tests (test/src/test_lighting.cpp:33-150)
restated against the oracle integrator's own lighting terms (the functions
its path loop calls; the device's shading kernel mirrors them and is
parity-checked against the oracle's images)."""
import numpy as np

import oracle


def test_specular_bounce_returns_raw_emission():  # :33-40
    e, beta = np.float32([2, 4, 6]), np.float32([1, 0.5, 0.25])
    assert np.allclose(e * beta * oracle.emission_weight(True, 5.0, 3.0), [2, 2, 1.5])


def test_non_specular_weights_by_power_heuristic():  # :42-51
    w = oracle.emission_weight(False, 3.0, 1.0)
    assert abs(w - 9.0 / 10.0) < 1e-7 and w == oracle.power_heuristic(3.0, 1.0)


def test_zero_pdf_non_specular_contributes_nothing():  # :53-59
    assert oracle.emission_weight(False, 0.0, 0.0) == 0.0
    assert oracle.emission_weight(False, 0.0, 2.0) == 0.0 and oracle.emission_weight(False, 2.0, 0.0) == 1.0


def test_next_bounce_pdfs_match_analytic_formula():  # :61-82
    # vertex at the origin facing +z, a light point 2 above it facing down, area pdf 1
    ok, _, pl, pb = oracle.direct_light([0, 0, 0], [0, 0, 1], [1, 1, 1], [0, 0, 2], [0, 0, -1], [1, 1, 1], 1.0)
    assert ok and abs(pb - 1 / np.pi) < 1e-6          # the cosine bounce: cos / pi
    assert abs(pl - 4.0) < 1e-6                      # area pdf x d^2 / |cos_l|
    ok, _, pl, _ = oracle.direct_light([0, 0, 0], [0, 0, 1], [1, 1, 1], [0, 1, 1], [0, 0, -1], [1, 1, 1], 1.0)
    assert ok and abs(pl - 2.0 / np.sqrt(0.5)) < 1e-5   # d^2 = 2, |cos_l| = 1/sqrt 2


def test_unoccluded_matches_rendering_equation():  # :84-101
    ok, ld, _, _ = oracle.direct_light([0, 0, 0], [0, 0, 1], [0.8] * 3, [0, 0, 2], [0, 0, -1], [1, 1, 1], 0.5)
    expected = 1.0 * (0.8 / np.pi) * (1.0 * 1.0 / 4.0) / 0.5
    assert ok and np.allclose(ld, expected, atol=1e-6)


def test_back_facing_surface_contributes_nothing():  # :103-113
    ok, ld, pl, pb = oracle.direct_light([0, 0, 0], [0, 0, -1], [0.8] * 3, [0, 0, 2], [0, 0, -1], [1, 1, 1], 0.5)
    assert not ok and np.all(ld == 0) and pl == 0 and pb == 0


def test_occluder_blocks_the_shadow_ray():  # :115-130
    # the integrator's shadow test is an any-hit ray towards the light point
    v = np.float32([[-1, -1, 1], [1, -1, 1], [0, 1, 1]])   # a triangle at z = 1
    f = np.uint32([[0, 1, 2]])
    ray = np.zeros(2, oracle.RAY_DTYPE)
    ray["dz"] = 1.0
    ray["tmax"] = [2.0, 0.5]                                 # to the light at z = 2, and stopping short
    assert list(oracle.brute_any(v, f, ray)) == [1, 0]


def test_light_points_lie_on_the_emitter_with_area_pdf():  # :132-150
    # a 2 x 2 emitting quad of two triangles: each is chosen with probability
    # 1/2 (equal areas) and sampled uniformly, so the area pdf is 1/4 everywhere
    v = np.float32([[-1, -1, 3], [1, -1, 3], [1, 1, 3], [-1, 1, 3]])
    f = np.uint32([[0, 1, 2], [0, 2, 3]])
    mats = oracle.materials_array([(0, (0, 0, 0), (5, 5, 5), 1.0)])
    ent, pdf = oracle.lights_build(v, f, np.zeros(2, np.int32), mats)
    assert np.allclose(pdf, 0.25, rtol=1e-6)
    rng = np.random.default_rng(3)
    for t in range(2):
        a, b, c = v[f[t]]
        for u1, u2 in rng.random((200, 2), dtype=np.float32):
            x = oracle.sample_triangle(a, b, c, u1, u2)
            assert abs(x[2] - 3.0) < 1e-6 and np.all(np.abs(x[:2]) <= 1 + 1e-6)
