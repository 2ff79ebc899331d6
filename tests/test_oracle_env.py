"""The oracle's lat-long environment map (EnvMap, src/scene/envmap.rs)
pinned by the reference's own test (envmap.rs:150-199: pdf integrates to 1
over the sphere within 0.03, sample -> eval round trip, the bright patch
takes > 150 of 200 samples) and by what the construction fixes: a constant
map has pdf h sin(pi / 2h) / (2 pi^2) everywhere (the sum of sin over the
row centres is 1 / sin(pi / 2h)); sampled texels follow lum * sin(theta)
(a binomial test per texel); the returned pdf is the sampled texel's
weight over the Jacobian; the mean luminance of a constant map is its
luminance times the row sum over h (2 / pi)."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import orc  # noqa: E402
import rgbe  # noqa: E402


def sphere_grid(nt, npp):
    th = (np.arange(nt) + 0.5) / nt * np.pi
    ph = (np.arange(npp) + 0.5) / npp * 2 * np.pi
    T, P = np.meshgrid(th, ph, indexing="ij")
    d = np.stack([np.sin(T) * np.sin(P), np.cos(T), -np.sin(T) * np.cos(P)], -1).reshape(-1, 3)
    dw = (np.sin(T) * (np.pi / nt) * (2 * np.pi / npp)).reshape(-1)
    return d, dw


def patch_map():
    w, h = 32, 16
    m = np.full((h, w, 3), 0.05, np.float32)
    m[4:8, 10:16] = 5.0
    return m


def test_reference_envmap_test():
    m = patch_map()
    d, dw = sphere_grid(128, 256)
    pdf, _ = orc.env_pdf_eval(m, d)
    assert abs((pdf * dw).sum() - 1.0) < 0.03
    rng = np.random.default_rng(7)
    dirs, rad, p, _ = orc.env_sample(m, rng.random((200, 2)))
    assert np.all(p > 0)
    _, ev = orc.env_pdf_eval(m, dirs)
    np.testing.assert_allclose(ev, rad, atol=1e-9)
    assert (rad[:, 0] > 1.0).sum() > 150


def test_constant_map_closed_form():
    h, w = 24, 48
    m = np.full((h, w, 3), 2.0, np.float32)
    d, _ = sphere_grid(37, 71)
    pdf, rad = orc.env_pdf_eval(m, d)
    want = h * np.sin(np.pi / (2 * h)) / (2 * np.pi ** 2)
    np.testing.assert_allclose(pdf, want, rtol=1e-12)
    np.testing.assert_allclose(rad, 2.0)
    _, _, _, ml = orc.env_sample(m, np.zeros((1, 2)))
    lum = 2.0 * (0.2126 + 0.7152 + 0.0722)
    np.testing.assert_allclose(ml, lum * w / np.sin(np.pi / (2 * h)) / (w * h * 2 / np.pi), rtol=1e-12)


def test_sampled_texels_follow_sin_weighted_luminance():
    rng = np.random.default_rng(3)
    h, w = 8, 16
    m = (rng.random((h, w, 3)) * 3).astype(np.float32)
    n = 400000
    dirs, rad, pdf, _ = orc.env_sample(m, rng.random((n, 2)))
    u = (np.arctan2(dirs[:, 0], -dirs[:, 2]) / (2 * np.pi)) % 1.0
    v = np.arccos(np.clip(dirs[:, 1], -1, 1)) / np.pi
    x = np.minimum((u * w).astype(int), w - 1)
    y = np.minimum((v * h).astype(int), h - 1)
    counts = np.bincount(y * w + x, minlength=w * h).astype(float)
    md = m.astype(np.float64)
    lum = 0.2126 * md[..., 0] + 0.7152 * md[..., 1] + 0.0722 * md[..., 2]
    sin_t = np.sin(np.pi * (np.arange(h) + 0.5) / h)[:, None]
    p = (lum * sin_t).reshape(-1)
    p /= p.sum()
    z = np.abs(counts - n * p) / np.sqrt(n * p * (1 - p))
    assert z.max() < 4.5
    # the pdf returned is the texel's weight over the 2 pi^2 sin(theta) Jacobian
    k = y * w + x
    want = p[k] * w * h / (2 * np.pi ** 2 * sin_t[y, 0])
    np.testing.assert_allclose(pdf, want, rtol=1e-9)
    np.testing.assert_allclose(rad, m[y, x], rtol=0, atol=0)


def test_sky_fixture_pdf_normalised():
    sky = rgbe.read_hdr(os.path.join(ROOT, "tests", "golden", "sky.hdr"))
    assert sky.shape == (256, 512, 3) and sky.max() > 100
    d, dw = sphere_grid(512, 1024)
    pdf, _ = orc.env_pdf_eval(sky, d)
    assert abs((pdf * dw).sum() - 1.0) < 0.01
