"""The reference's sampler tests (test/src/test_sampling.cpp:18-102) restated
against the oracle's samplers, which the device shading kernel mirrors (the
GPU renders are parity-checked against the oracle): unit directions, a mean
cosine of 2/3 for the cosine-weighted bounce, triangle points inside the
triangle with the centroid as their mean. 5000 samples as in the reference;
the random numbers come from numpy (seed 0) instead of its DeterministicRng."""
import numpy as np

import oracle

N = 5000


def uniforms(seed=0):
    return np.random.default_rng(seed).random((N, 2), dtype=np.float32)


def test_sphere_samples_are_unit_with_zero_mean():
    w = np.array([oracle.sample_sphere(a, b) for a, b in uniforms()])
    assert np.allclose(np.linalg.norm(w, axis=1), 1.0, atol=1e-5)
    assert np.linalg.norm(w.mean(0)) < 0.05


def test_cosine_samples_unit_in_hemisphere_mean_cos_two_thirds():
    for n in (np.float32([0, 0, 1]), (np.float32([1, 2, 3]) / np.linalg.norm([1, 2, 3])).astype(np.float32)):
        w = np.array([oracle.sample_cosine(n, a, b) for a, b in uniforms(1)])
        assert np.allclose(np.linalg.norm(w, axis=1), 1.0, atol=1e-5)
        cos = w @ n
        assert np.all(cos >= -1e-5)
        assert abs(cos.mean() - 2.0 / 3.0) < 0.03
        # the density is cos / pi: E[cos^2] = 1/2 (uniform hemisphere would give 1/3)
        assert abs((cos ** 2).mean() - 0.5) < 0.03


def test_triangle_points_inside_with_centroid_mean():
    a, b, c = np.float32([0, 0, 0]), np.float32([1, 0, 0]), np.float32([0, 1, 0])
    p = np.array([oracle.sample_triangle(a, b, c, u, v) for u, v in uniforms(2)])
    assert np.all(p[:, 0] >= -1e-6) and np.all(p[:, 1] >= -1e-6)
    assert np.all(p[:, 0] + p[:, 1] <= 1 + 1e-6) and np.allclose(p[:, 2], 0.0, atol=1e-6)
    assert np.allclose(p.mean(0), [1 / 3, 1 / 3, 0], atol=0.02)
    # area-uniform: the fraction in the half of the triangle nearer A (x + y < 1/sqrt 2) is 1/2
    assert abs(np.mean(p[:, 0] + p[:, 1] < np.sqrt(0.5)) - 0.5) < 0.03
