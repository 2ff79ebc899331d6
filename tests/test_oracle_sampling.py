"""Pins the oracle's PCG32 and Perlin/fBm (and the product generator's use
of them) against known answers.

PCG32: the published pcg32 demo stream and an independent pure-Python
integer restatement (tests/golden/pcg32_kat.json) of render.rs
src/raytracer/pt/sampler.rs:12-42. Perlin: the reference's own KAT
(src/geometry/displace.rs:130-144) plus properties the construction fixes.
"""
import json
import os

import numpy as np

import orc

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "pcg32_kat.json")


def test_pcg32_published_demo_stream():
    kat = json.load(open(GOLDEN))["published_demo"]
    got = orc.pcg32_seq(kat["seed"], kat["stream"], len(kat["u32"]))
    assert got.tolist() == kat["u32"]


def test_pcg32_pixel_sample_streams():
    for row in json.load(open(GOLDEN))["pixel_sample"]:
        got = orc.pcg32_pixel_sample(row["pixel"], row["sample"], len(row["u32"]))
        assert got.tolist() == row["u32"], (row["pixel"], row["sample"])


def test_pcg32_deterministic_and_distinct():
    # sampler.rs:64-88 (uniform mean, per-seed determinism)
    a = orc.pcg32_pixel_sample(7, 3, 8)
    b = orc.pcg32_pixel_sample(7, 3, 8)
    c = orc.pcg32_pixel_sample(7, 4, 8)
    assert (a == b).all() and not (a == c).all()
    u = orc.pcg32_seq(42, 54, 100000).astype(np.float64) / 2.0**32
    assert abs(u.mean() - 0.5) < 0.005


def test_perlin_zero_at_lattice_and_bounded():
    # displace.rs:130-144
    lattice = np.array([[i, j, 0.0] for i in range(-3, 3) for j in range(-3, 3)])
    assert np.abs(orc.perlin(lattice)).max() < 1e-12
    t = np.arange(1000) * 0.113
    pts = np.stack([t, t * 0.7 + 0.3, t * 1.3 + 0.9], axis=1)
    mx = np.abs(orc.perlin(pts)).max()
    assert 0.05 < mx <= 1.3


def test_perlin_is_trilinear_blend_of_corner_gradients():
    # At the centre of a cell every fade weight is 1/2, so the value is the
    # mean of the eight corner dot products g_k . (f - k) with f = 1/2; each
    # is a sum of two +-1/2 terms, so the value lies on the 1/8 grid.
    pts = np.array([[i + 0.5, j + 0.5, k + 0.5] for i in range(-2, 2) for j in range(-2, 2)
                    for k in range(-2, 2)])
    v = orc.perlin(pts)
    assert np.allclose(v * 8, np.round(v * 8), atol=1e-12)
    assert np.abs(v).max() <= 1.0 + 1e-12


def test_fbm_octave_sum():
    rng = np.random.default_rng(3)
    pts = rng.uniform(-5, 5, size=(200, 3))
    off = (10.0, 20.0, 30.0)
    one = orc.fbm(pts, frequency=2.0, octaves=1, offset=off)
    assert np.array_equal(one, orc.perlin(pts * 2.0 + np.array(off)))
    three = orc.fbm(pts, frequency=2.0, octaves=3, gain=0.5, lacunarity=2.0, offset=off)
    manual = sum(0.5**k * orc.perlin(pts * (2.0 * 2.0**k) + np.array(off)) for k in range(3))
    assert np.allclose(three, manual, rtol=0, atol=1e-15)
