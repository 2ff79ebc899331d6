"""Per-pixel statistics shared by the path-tracing parity tests.

Accumulators hold, per pixel, the sum and the sum of squares of each
channel and the sample count (include/rrhip.h RRHIP_ACCUM_STRIDE)."""
import numpy as np


def mean_var(acc):
    n = acc[:, 6:7]
    mean = acc[:, :3] / n
    var = np.maximum(acc[:, 3:6] / n - mean * mean, 0.0) * n / np.maximum(n - 1, 1)
    return mean, var, n


def sigma_test(acc_a, acc_b):
    """z-scores of the per-pixel, per-channel mean difference, taking the
    two estimates as independent (conservative when they share PCG32
    streams and so are positively correlated)."""
    ma, va, na = mean_var(acc_a)
    mb, vb, nb = mean_var(acc_b)
    se = np.sqrt(va / na + vb / nb)
    d = np.abs(ma - mb)
    z = np.where(se > 0, d / np.where(se > 0, se, 1), np.where(d > 1e-6, np.inf, 0.0))
    return z


def image_mean_z(acc_a, acc_b):
    """z-score of the difference of the whole-image channel means."""
    ma, va, na = mean_var(acc_a)
    mb, vb, nb = mean_var(acc_b)
    npx = ma.shape[0]
    da = ma.mean(0) - mb.mean(0)
    se = np.sqrt((va / na).sum(0) + (vb / nb).sum(0)) / npx
    return np.abs(da) / se
