"""The device PCG32 stream, bit for bit: the draws the camera-path kernel's
seeding produces on the GPU (rrhip_pcg32_draws) against the golden vectors
of tests/golden/pcg32_kat.json (render.rs sampler.rs:12-42, restated in
Python integers) and, over 65,536 random (pixel, sample) pairs, against a
vectorised numpy restatement with wrapping u64 arithmetic."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def np_pcg32_draws(pixel, sample, count):
    """sampler.rs:12-42 over arrays (numpy u64 arithmetic wraps mod 2^64)."""
    with np.errstate(over="ignore"):
        def mix(z):
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            return z ^ (z >> np.uint64(31))
        mult = np.uint64(6364136223846793005)
        seed = mix(pixel * np.uint64(0x9E3779B97F4A7C15) ^ sample)
        inc = ((mix(sample) | np.uint64(1)) << np.uint64(1)) | np.uint64(1)
        state = np.zeros_like(pixel)

        def step(state):
            old = state
            new = old * mult + inc
            xs = (((old >> np.uint64(18)) ^ old) >> np.uint64(27)).astype(np.uint32)
            rot = (old >> np.uint64(59)).astype(np.uint32)
            out = (xs >> rot) | (xs << ((np.uint32(32) - rot) & np.uint32(31)))
            return new, out
        state, _ = step(state)
        state = state + seed
        state, _ = step(state)
        outs = []
        for _ in range(count):
            state, o = step(state)
            outs.append(o)
    return np.stack(outs, 1)


def test_device_stream_equals_the_golden_vectors(ctx):
    kat = json.load(open(os.path.join(HERE, "golden", "pcg32_kat.json")))
    rows = kat["pixel_sample"]
    px = np.array([r["pixel"] for r in rows], np.uint64)
    sm = np.array([r["sample"] for r in rows], np.uint64)
    got = ctx.pcg32_draws(px, sm, 8)
    want = np.array([r["u32"] for r in rows], np.uint32)
    assert np.array_equal(got, want), (got, want)
    assert np.array_equal(np_pcg32_draws(px, sm, 8), want)


def test_device_stream_equals_numpy_on_random_pairs(ctx):
    rng = np.random.default_rng(5)
    px = rng.integers(0, 1920 * 1080, 65536, dtype=np.uint64)
    sm = rng.integers(0, 1 << 62, 65536, dtype=np.uint64)
    sm[:1024] = np.arange(1024, dtype=np.uint64)  # small sample indices, as renders use
    sm[1024:2048] |= np.uint64(1 << 63)            # training-path streams set the top bit
    got = ctx.pcg32_draws(px, sm, 16)
    assert np.array_equal(got, np_pcg32_draws(px, sm, 16))
