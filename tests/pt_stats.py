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

Ray origins: the GPU's default error-bounded origins start rays on the
surface up to a push far below render.rs's fixed 1e-4 (pt/mod.rs:30). That
fixed push has effects of its own — it darkens the box room by 0.33 %
(the same seeds rendered by the oracle with pushes of 1e-4 and 1e-9) —
so a GPU render with error-bounded origins is compared with the oracle
pushing by EXACT_ORIGIN_PUSH (f64 hit points near unit scale are exact to
~1e-16, so 1e-9 only keeps the reference's t > 1e-9 / 1e-6 tests from
seeing the surface just left), and a GPU render with render.rs's push
(params.ray_offset > 0) with the oracle pushing by the same amount.
"""
import numpy as np

ORACLE_SAMPLE_BASE = 1 << 48
EXACT_ORIGIN_PUSH = 1e-9


def _oracle_params(params):
    import rrhip
    p = rrhip.RenderParams.from_buffer_copy(params)
    p.sample_base = params.sample_base + ORACLE_SAMPLE_BASE
    p.ray_offset = params.ray_offset if params.ray_offset > 0 else EXACT_ORIGIN_PUSH
    return p


def oracle_render(desc, params, start, count, **kw):
    """orc.render over samples [start, start + count) seeded independently
    of a GPU render of the same range, with the GPU render's origin rule."""
    import orc
    return orc.render(desc, _oracle_params(params), start, count, **kw)


def oracle_render_pixels(desc, params, pixels, start, count, **kw):
    """orc.render_pixels (a tile at its whole-image seeds) seeded
    independently of a GPU render of the same range, with the GPU render's
    origin rule."""
    import orc
    return orc.render_pixels(desc, _oracle_params(params), pixels, start, count, **kw)


def mean_var(acc):
    n = acc[:, 6:7]
    mean = acc[:, :3] / n
    var = np.maximum(acc[:, 3:6] / n - mean * mean, 0.0) * n / np.maximum(n - 1, 1)
    return mean, var, n


def sigma_test(acc_a, acc_b):
    """z-scores of the per-pixel, per-channel mean difference of two
    independent estimates. The standard error has a floor at 4e-7 of the
    pixel value, the f32 rounding of a radiance: a pixel every sample of
    which sees the same constant (a background that f32 does not hold
    exactly) has an f32 renderer's exact zero variance against an f64
    renderer's ~1e-17 of rounding noise, and differs by that rounding."""
    ma, va, na = mean_var(acc_a)
    mb, vb, nb = mean_var(acc_b)
    se = np.sqrt(va / na + vb / nb + (4e-7 * np.maximum(np.abs(ma), np.abs(mb))) ** 2)
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
