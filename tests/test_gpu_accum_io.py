"""Checkpoint / resume and sample-range merge through the GPU path
(render_pt_checkpointed, renderer.rs:765-808; render merge, accum.rs:
86-119): a render stopped after 8 of 16 samples and resumed from its RCKD
checkpoint in a new session equals the uninterrupted render bit for bit
(per-(pixel, sample) seeding and in-order f64 accumulation); resumed from
render.rs's f32 RCKP it equals it to f32 rounding of the sums; GPU sample
ranges written as accumulation EXRs merge to the one-session mean."""
import sys

import numpy as np
import pytest

import accum_io
import rrhip
import scenes

pytestmark = pytest.mark.gpu

W = H = 32


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def session(ctx, s):
    ps = rrhip.PtScene(ctx, s.desc())
    return ps, ps.session(rrhip.RenderParams.make(max_depth=6, samples_per_pass=4))


def test_checkpoint_resume_is_exact(ctx, tmp_path):
    say("cornell 32x32: 16 spp straight vs 8 + resume 8 from RCKD and from RCKP")
    s = scenes.cornell_fixture(W, H)
    ps, se = session(ctx, s)
    se.render(0, 16)
    full = se.accum()
    se.close()
    ps.close()
    ck = str(tmp_path / "run.rckd")
    ps, se = session(ctx, s)
    assert accum_io.render_checkpointed(se, W, H, 8, ck) == 8
    se.close()
    ps.close()
    ps, se = session(ctx, s)
    assert accum_io.render_checkpointed(se, W, H, 16, ck) == 8  # resumed at 8
    res = se.accum()
    se.close()
    ps.close()
    assert np.array_equal(res.view(np.uint64), full.view(np.uint64))
    # from render.rs's f32 checkpoint: sums and counts to f32 rounding
    ps, se = session(ctx, s)
    se.render(0, 8)
    rk = str(tmp_path / "metal.ckpt")
    accum_io.write_rckp(rk, W, H, 8, se.accum())
    se.close()
    ps.close()
    ps, se = session(ctx, s)
    assert accum_io.render_checkpointed(se, W, H, 16, rk) == 8
    res = se.accum()
    se.close()
    ps.close()
    np.testing.assert_allclose(res[:, 0:3], full[:, 0:3], rtol=1e-6, atol=1e-6)
    assert np.array_equal(res[:, 6], full[:, 6])
    # a resumed RCKP stays an RCKP, so render.rs can still resume it
    done, _ = accum_io.read_rckp(rk, W, H)
    assert done == 16 and accum_io.read_rckd(rk, W, H) is None


def test_checkpointed_render_without_a_checkpoint_starts_from_zero(ctx, tmp_path):
    say("cornell 32x32: a checkpointed render in a used session starts from a clear accumulator")
    s = scenes.cornell_fixture(W, H)
    ps, se = session(ctx, s)
    se.render(0, 8)
    ref = se.accum()
    se.render(8, 4)  # leftovers a fresh checkpointed render must not keep
    ck = str(tmp_path / "fresh.rckd")
    assert accum_io.render_checkpointed(se, W, H, 8, ck) == 8
    assert np.array_equal(se.accum().view(np.uint64), ref.view(np.uint64))
    se.close()
    ps.close()


def test_sample_ranges_merge(ctx, tmp_path):
    say("cornell 32x32: [0, 5) and [5, 12) as accumulation EXRs merge to [0, 12)")
    s = scenes.cornell_fixture(W, H)
    paths = []
    for a, n in ((0, 5), (5, 7)):
        ps, se = session(ctx, s)
        se.render(a, n)
        p = str(tmp_path / f"r{a}.exr")
        accum_io.write_accum_exr(p, W, H, se.accum())
        paths.append(p)
        se.close()
        ps.close()
    ps, se = session(ctx, s)
    se.render(0, 12)
    full = se.accum()
    se.close()
    ps.close()
    _, _, mean = accum_io.merge(paths)
    np.testing.assert_allclose(mean, full[:, 0:3] / full[:, 6:7], rtol=2e-6, atol=1e-7)
