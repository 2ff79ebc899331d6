"""GPU adaptive sampling (WfSession::set_adaptive / average_spp,
renderer.rs:580-591, with the schedule of pt::render_adaptive,
pt/mod.rs:285-336).

Bars: the adaptive render is bit-identical to a plain render of the same
samples cut, pixel by pixel, where the stopping rule (evaluated here on
the plain render's own accumulators at 32, 48, ...) first holds; pixel
counts and means agree with the oracle's adaptive render within a paired
4-sigma test; render.rs's own test (metal_pt_tests.rs:507-525: savings,
warm-up floor, image mean within 2 % of a fixed-spp render) passes."""
import sys

import numpy as np
import pytest

import orc
import rrhip
import scenes

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def converged(acc, tol):
    """The device's stopping rule on accumulator sums (same operations, no
    fused multiply-adds)."""
    n = acc[:, 6]
    s = 0.2126 * acc[:, 0] + 0.7152 * acc[:, 1] + 0.0722 * acc[:, 2]
    with np.errstate(divide="ignore", invalid="ignore"):
        mean = s / n
        var = np.maximum((acc[:, 7] - s * mean) / (n - 1.0), 0.0)
        return (n >= 32) & (1.96 * np.sqrt(var / n) < tol * np.maximum(mean, 0.05))


def test_adaptive_render_is_the_plain_render_cut_by_the_rule(ctx):
    say("adaptive vs plain snapshots: sphere zoo 64x64, 128 spp max, tol 0.05, 8 spp per pass")
    s = scenes.sphere_zoo(64, 64)
    ps = rrhip.PtScene(ctx, s.desc())
    prm = rrhip.RenderParams.make(max_depth=8, samples_per_pass=8)
    max_spp, tol = 128, 0.05
    ad = ps.session(prm)
    avg = ad.render_adaptive(max_spp, tol)
    acc = ad.accum()
    plain = ps.session(prm)
    snaps = {}
    for k in range(0, max_spp, 16):
        plain.render(k, 16)
        snaps[k + 16] = plain.accum()
    want = np.full(acc.shape[0], max_spp)
    open_ = np.ones(acc.shape[0], bool)
    for b in range(32, max_spp, 16):
        stop = open_ & converged(snaps[b], tol)
        want[stop] = b
        open_ &= ~stop
    counts, freq = np.unique(want, return_counts=True)
    say(f"average spp {avg:.2f}; stopping points {dict(zip(counts.tolist(), freq.tolist()))}")
    assert np.array_equal(acc[:, 6], want.astype(np.float64))
    for b in counts:
        sel = want == b
        assert acc[sel].tobytes() == snaps[int(b)][sel].tobytes(), b
    assert abs(avg - want.mean()) < 1e-9
    assert len(counts) >= 3
    # the same passes split over several calls: identical
    ad2 = ps.session(prm)
    ad2.set_adaptive(tol)
    for a, b in ((0, 24), (24, 40), (40, 128)):
        ad2.render(a, b - a)
    assert ad2.accum().tobytes() == acc.tobytes()
    # a second adaptive render after clearing starts over (all pixels live)
    assert ad.render_adaptive(max_spp, tol) == avg and ad.accum().tobytes() == acc.tobytes()
    for se in (ad, ad2, plain):
        se.close()
    ps.close()


def test_adaptive_matches_the_oracle(ctx):
    say("adaptive vs oracle: sphere zoo 32x32, 128 spp max, tol 0.05")
    s = scenes.sphere_zoo(32, 32)
    d = s.desc()
    prm = rrhip.RenderParams.make(max_depth=8, samples_per_pass=16)
    ps = rrhip.PtScene(ctx, d)
    se = ps.session(prm)
    se.render_adaptive(128, 0.05)
    g = se.accum()
    c = orc.render_adaptive(d, prm, 128, 0.05)
    se.close()
    ps.close()
    lum = np.array([0.2126, 0.7152, 0.0722])
    for name, x, y in (("count", g[:, 6], c[:, 6]),
                       ("mean luminance", g[:, :3] @ lum / g[:, 6], c[:, :3] @ lum / c[:, 6])):
        dd = x - y
        sd = dd.std(ddof=1)
        z = abs(dd.mean()) / (sd / np.sqrt(dd.size)) if sd > 0 else 0.0
        say(f"{name}: gpu {x.mean():.5f} oracle {y.mean():.5f}; pixels equal {np.mean(dd == 0):.3f}; paired z {z:.2f}")
        assert z < 4, (name, z)
    assert np.mean(g[:, 6] == c[:, 6]) > 0.8


def test_adaptive_saves_samples(ctx):
    """metal_pt_tests.rs:507-525 on the sphere zoo at 96 x 96."""
    say("adaptive savings: sphere zoo 96x96, 256 spp max, tol 0.05")
    s = scenes.sphere_zoo(96, 96)
    ps = rrhip.PtScene(ctx, s.desc())
    prm = rrhip.RenderParams.make(max_depth=8)
    se = ps.session(prm)
    se.set_adaptive(0.05)
    se.render(0, 256)
    avg = se.average_spp()
    full = ps.session(prm)
    full.render(0, 256)
    a, f = se.accum(), full.accum()
    ma = (a[:, :3] / a[:, 6:7]).mean(0)
    mf = (f[:, :3] / f[:, 6:7]).mean(0)
    rel = np.abs(ma - mf).sum() / max(mf.sum(), 0.05)
    say(f"average spp {avg:.1f} (full {full.average_spp():.0f}); image mean rel diff {rel:.4f}; "
        f"device ms adaptive {se.stats()['last_render_ms']:.2f} full {full.stats()['last_render_ms']:.2f}")
    assert 32.0 <= avg < 200.0
    assert rel < 0.02
    assert full.average_spp() == 256.0
    se.close()
    full.close()
    ps.close()


def test_adaptive_modes_and_arguments(ctx):
    say("adaptive: tolerance 0, launch modes, cache, bad arguments")
    s = scenes.sphere_zoo(32, 24)
    ps = rrhip.PtScene(ctx, s.desc())
    prm = rrhip.RenderParams.make(max_depth=6, samples_per_pass=4)
    plain = ps.session(prm)
    plain.render(0, 64)
    off = ps.session(prm)
    off.set_adaptive(0.0)
    off.render(0, 64)
    assert off.accum().tobytes() == plain.accum().tobytes()
    with pytest.raises(rrhip.RrhipError):
        off.set_adaptive(-1.0)
    with pytest.raises(rrhip.RrhipError):
        off.set_adaptive(float("nan"))
    ref = ps.session(prm)
    ref.render_adaptive(96, 0.05)
    for mode in (1, 2, 3):
        se = ps.session(prm)
        se.set_launch_mode(mode)
        se.render_adaptive(96, 0.05)
        assert se.accum().tobytes() == ref.accum().tobytes(), mode
        se.close()
    # with a radiance cache: counts follow the schedule, values finite
    cache = rrhip.Nrc(ctx, rrhip.NrcConfig.make(log2_table=14, train_paths=2048, batch=1024))
    se = ps.session(prm)
    se.set_nrc(cache)
    se.render_adaptive(96, 0.05)
    acc = se.accum()
    n = acc[:, 6]
    assert np.isfinite(acc).all()
    assert np.all((n == 96) | ((n >= 32) & (n % 16 == 0)))
    for x in (se, plain, off, ref):
        x.close()
    cache.close()
    ps.close()
