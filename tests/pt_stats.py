"""Per-pixel statistics shared by the path-tracing parity tests.

Accumulators hold, per pixel, the sum and the sum of squares of each
channel and the sample count (include/rrhip.h RRHIP_ACCUM_STRIDE).

The z-scores below take the two accumulators as independent estimates.
Renders that share PCG32 streams (same pixel, same sample index) are
positively correlated, which shrinks the true spread of their difference
and makes these z-scores too small, so the oracle side of every GPU
comparison is rendered with `oracle_render`, whose sample indices are
offset by ORACLE_SAMPLE_BASE: PCG32 seeds mix the full 64-bit sample index
(sampler.rs:24-32), so the two renders draw independent streams.
"""
import numpy as np

ORACLE_SAMPLE_BASE = 1 << 48


def oracle_render(desc, params, start, count, **kw):
    """orc.render over samples [start, start + count) seeded independently
    of a GPU render of the same range."""
    import orc
    import rrhip
    p = rrhip.RenderParams.from_buffer_copy(params)
    p.sample_base = params.sample_base + ORACLE_SAMPLE_BASE
    return orc.render(desc, p, start, count, **kw)


def oracle_render_pixels(desc, params, pixels, start, count, **kw):
    """orc.render_pixels (a tile at its whole-image seeds) seeded
    independently of a GPU render of the same range."""
    import orc
    import rrhip
    p = rrhip.RenderParams.from_buffer_copy(params)
    p.sample_base = params.sample_base + ORACLE_SAMPLE_BASE
    return orc.render_pixels(desc, p, pixels, start, count, **kw)


def mean_var(acc):
    n = acc[:, 6:7]
    mean = acc[:, :3] / n
    var = np.maximum(acc[:, 3:6] / n - mean * mean, 0.0) * n / np.maximum(n - 1, 1)
    return mean, var, n


def sigma_test(acc_a, acc_b):
    """z-scores of the per-pixel, per-channel mean difference of two
    independent estimates."""
    ma, va, na = mean_var(acc_a)
    mb, vb, nb = mean_var(acc_b)
    se = np.sqrt(va / na + vb / nb)
    d = np.abs(ma - mb)
    z = np.where(se > 0, d / np.where(se > 0, se, 1), np.where(d > 1e-6, np.inf, 0.0))
    return z


def image_mean_z(acc_a, acc_b):
    """z-score of the difference of the whole-image channel means of two
    independent estimates."""
    ma, va, na = mean_var(acc_a)
    mb, vb, nb = mean_var(acc_b)
    npx = ma.shape[0]
    da = ma.mean(0) - mb.mean(0)
    se = np.sqrt((va / na).sum(0) + (vb / nb).sum(0)) / npx
    return np.abs(da) / se
