"""Statistics shared by the image-comparison tests."""
import math

import numpy as np


def batch_z(batches_a, batches_b, block=4):
    """Per block and channel: the difference of two estimators' block means
    over its standard error, estimated from the spread of K independent
    batch images of each (K x h x w x c). With few batches that ratio is
    Student-t distributed (Welch-Satterthwaite degrees of freedom), so it is
    mapped to the standard normal quantile of the same probability."""
    from scipy import stats

    a, b = np.stack(batches_a), np.stack(batches_b)
    ka, h, w, c = a.shape
    kb = b.shape[0]
    blk = lambda x: x.reshape(x.shape[0], h // block, block, w // block, block, c).mean(axis=(2, 4))  # noqa: E731
    ba, bb = blk(a), blk(b)
    va, vb = ba.var(axis=0, ddof=1) / ka, bb.var(axis=0, ddof=1) / kb
    keep = (va + vb) > 0
    va, vb = va[keep], vb[keep]
    t = (ba.mean(axis=0) - bb.mean(axis=0))[keep] / np.sqrt(va + vb)
    df = (va + vb) ** 2 / (va ** 2 / (ka - 1) + vb ** 2 / (kb - 1))
    # upper-tail probabilities keep their precision far out in the tails
    return (np.sign(t) * stats.norm.isf(stats.t.sf(np.abs(t), df))).ravel()


def assert_noise_like(z, label, mean_tol=0.12, sd_range=(0.8, 1.2), beyond3=0.01):
    """z-scores that look like standard normal noise: no shift, unit spread,
    few beyond 3."""
    print(f"{label}: block z mean {z.mean():.3f} sd {z.std():.3f} max {np.abs(z).max():.2f} "
          f"beyond 3: {np.mean(np.abs(z) > 3):.4f}", flush=True)
    assert abs(z.mean()) < mean_tol
    assert sd_range[0] < z.std() < sd_range[1]
    assert np.mean(np.abs(z) > 3) < beyond3


def mean_z(batches_a, batches_b):
    """z-score of the difference of the two image means, from the batch spread."""
    ma = [float(np.mean(x)) for x in batches_a]
    mb = [float(np.mean(x)) for x in batches_b]
    se = math.sqrt(np.var(ma, ddof=1) / len(ma) + np.var(mb, ddof=1) / len(mb))
    return (np.mean(ma) - np.mean(mb)) / se
