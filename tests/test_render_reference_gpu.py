"""The device render with the C++ facade's defaults (the reference's
integrator settings: 10 vertices, no Russian roulette, Schlick Fresnel,
uniformly jittered pixel samples) against the oracle's restatement of the
reference's own integrator (sampleColor in reference semantics: path first,
then NEE; IOR stack; cast over objects with median-split BBHs), which the CPU
tests pin with the reference's furnace and its umbrella PNG.

The two use independent random streams, so the bar is statistical, at high
spp, over 8 independent batches of each:
  - the image means agree: |z| < 3;
  - per 4 x 4 block and channel, the difference over its standard error
    (Student-t with Welch degrees of freedom, mapped to a normal score)
    looks like noise (mean within 0.12 of 0, sd in (0.8, 1.2), under 1%
    beyond 3; 3% for Cornell, whose glass caustic has heavy tails);
  - per-pixel relative tolerance of the luminance: median, 95th and 99th
    percentiles of |device - reference| / reference below the stated bounds
    (Cornell 0.04 / 0.15 / 0.30 at 8192 against 2048 spp; umbrella 0.10 /
    0.30 / 0.45 at 4096 against 384 spp, where the reference side's noise
    sets them).
"""
import numpy as np
import pytest

import devilray_amd as dr
import oracle
from devilray_amd import scenes
from stats_helpers import assert_noise_like, batch_z, mean_z

pytestmark = pytest.mark.gpu


def gpu_batches(ctx, sc, cam, spp, k, seed0):
    s = dr.Scene(ctx, sc.vertices, sc.indices, sc.tri_material, sc.dr_materials())
    s.build_bvh()
    out = []
    for j in range(k):
        acc = s.render(cam, spp=spp, max_depth=10, rr_depth=10, fresnel=1, seed=seed0 + j)
        out.append(acc[..., :3] / acc[..., 3:])
    s.close()
    return out


def ref_batches(sc, cam, spp, k, seed0):
    rs = oracle.RefScene(sc.vertices, sc.indices, sc.object_ranges())
    ocam = oracle.Camera.from_buffer_copy(bytes(cam))
    out = []
    for j in range(k):
        acc = oracle.render_ref(rs, sc.tri_material, oracle.materials_array(sc.materials), ocam, spp=spp,
                                seed=seed0 + j)
        out.append(acc[..., :3] / acc[..., 3:])
    return out


def compare(gpu, ref, label, q50, q95, q99, beyond3=0.01):
    z = mean_z(gpu, ref)
    print(f"{label}: image means {np.mean(gpu):.6f} / {np.mean(ref):.6f}, z {z:.2f}", flush=True)
    assert abs(z) < 3
    assert_noise_like(batch_z(gpu, ref), label, beyond3=beyond3)
    lg, lr = np.mean(gpu, axis=0).mean(-1), np.mean(ref, axis=0).mean(-1)
    rel = np.abs(lg - lr) / np.maximum(lr, 1e-3)
    qs = np.quantile(rel, [0.5, 0.95, 0.99])
    print(f"{label}: per-pixel relative difference quantiles 50/95/99 {qs}", flush=True)
    assert qs[0] <= q50 and qs[1] <= q95 and qs[2] <= q99


def test_cornell_matches_reference_integrator(gpu_ctx):
    print("[case] cornell, facade defaults vs reference integrator", flush=True)
    sc = scenes.cornell_c0()
    c = sc.camera
    cam = dr.camera_lookat(c["eye"], c["target"], c["up"], c["fov"], 64, 64)
    gpu = gpu_batches(gpu_ctx, sc, cam, 1024, 8, 1000)
    ref = ref_batches(sc, cam, 256, 8, 2000)
    # the glass sphere's caustic is reached only by BSDF sampling: rare bright
    # samples give its blocks heavier tails than normal noise (about 1% of
    # blocks beyond 3 where 0.3% would be normal); a bias would move the mean
    # or the spread of the block scores instead
    compare(gpu, ref, "cornell", 0.04, 0.15, 0.30, beyond3=0.03)


def test_umbrella_matches_reference_integrator(gpu_ctx):
    print("[case] umbrella, facade defaults vs reference integrator", flush=True)
    sc = scenes.triangle_umbrella()
    cam = dr.camera_reference(128, 128, 50.0, 0.2)
    gpu = gpu_batches(gpu_ctx, sc, cam, 512, 8, 3000)
    ref = ref_batches(sc, cam, 48, 8, 4000)
    compare(gpu, ref, "umbrella", 0.10, 0.30, 0.45)
