"""The oracle's adaptive sampling (pt::render_adaptive, pt/mod.rs:285-336)
pinned by: the rule restated in plain Python (Welford over the oracle's own
per-sample values) deciding every pixel's count exactly, the accumulated
sums being exactly those of the first n samples, the closed cases (zero
variance stops at the 32-sample warm-up, zero tolerance never stops,
max_spp below the warm-up is a plain render), and render.rs's own test
(pt_tests.rs:98-115: savings on a direct-light scene, warm-up floor, mean
within 3 % of a fixed-spp render)."""
import numpy as np

import orc
import rrhip
import scenes

LUM = np.array([0.2126, 0.7152, 0.0722])


def per_sample_values(desc, params, n):
    """Firefly-clamped rgb of samples 0..n-1 of every pixel: (n, npix, 3)."""
    out = []
    for s in range(n):
        acc, _ = orc.render(desc, params, s, 1)
        out.append(acc[:, :3].copy())
    return np.stack(out)


def welford_counts(vals, max_spp, tol):
    """pt/mod.rs:300-327 written out per pixel."""
    counts = np.empty(vals.shape[1], np.int64)
    for p in range(vals.shape[1]):
        mean, m2, n = 0.0, 0.0, 0
        while n < max(max_spp, 1):
            lum = 0.2126 * vals[n, p, 0] + 0.7152 * vals[n, p, 1] + 0.0722 * vals[n, p, 2]
            n += 1
            delta = lum - mean
            mean += delta / n
            m2 += delta * (lum - mean)
            if n >= 32 and n % 16 == 0:
                var = m2 / (n - 1.0)
                if 1.96 * np.sqrt(var / n) < tol * max(mean, 0.05):
                    break
        counts[p] = n
    return counts


def test_rule_and_sums_match_a_plain_restatement():
    s = scenes.sphere_zoo(20, 12)
    desc = s.desc()
    params = rrhip.RenderParams.make(max_depth=6)
    max_spp, tol = 96, 0.08
    vals = per_sample_values(desc, params, max_spp)
    want = welford_counts(vals, max_spp, tol)
    acc = orc.render_adaptive(desc, params, max_spp, tol)
    assert np.array_equal(acc[:, 6].astype(np.int64), want)
    assert len(np.unique(want)) >= 3, np.unique(want)  # the case exercises several stopping points
    for p in range(acc.shape[0]):
        n = want[p]
        sums = np.zeros(3)
        for k in range(n):
            sums = sums + vals[k, p]
        assert np.array_equal(acc[p, :3], sums)
    # column 7: sum of squared luminance, as orc_render writes it
    plain, _ = orc.render(desc, params, 0, 40)
    lum = np.einsum("kpc,c->kp", vals[:40], LUM)
    assert np.allclose(plain[:, 7], (lum ** 2).sum(0), rtol=1e-12, atol=1e-300)


def test_closed_cases():
    s = scenes.sphere_zoo(12, 8)
    desc = s.desc()
    params = rrhip.RenderParams.make(max_depth=4)
    # zero tolerance: ci < 0 never holds, every pixel takes max_spp
    acc = orc.render_adaptive(desc, params, 64, 0.0)
    assert (acc[:, 6] == 64).all()
    # below the warm-up no check happens: a plain render
    acc = orc.render_adaptive(desc, params, 20, 10.0)
    plain, _ = orc.render(desc, params, 0, 20)
    assert np.array_equal(acc, plain)
    # a black scene (no lights): zero mean and variance stop at 32
    dark = scenes.Scene(8, 8, 40.0, eye=(0, 0, 0), look_at=(0, 0, 1))
    dark.add_sphere((0, 0, 5), 1.0, dark.add_material(scenes.matte((0.5, 0.5, 0.5))))
    acc = orc.render_adaptive(dark.desc(), params, 256, 0.01)
    assert (acc[:, 6] == 32).all() and not acc[:, :3].any()


def test_adaptive_saves_samples_and_keeps_the_mean():
    """pt_tests.rs:98-115 on the sphere zoo at 32 x 32."""
    s = scenes.sphere_zoo(32, 32)
    desc = s.desc()
    params = rrhip.RenderParams.make(max_depth=8)
    acc = orc.render_adaptive(desc, params, 256, 0.05)
    avg = acc[:, 6].mean()
    assert 32.0 <= avg < 160.0, avg
    ref, _ = orc.render(desc, params, 0, 256)
    ma = (acc[:, :3] / acc[:, 6:7]).mean(0)
    mr = (ref[:, :3] / ref[:, 6:7]).mean(0)
    rel = np.abs(ma - mr).sum() / max(mr.sum(), 0.05)
    assert rel < 0.03, (ma, mr, rel)
