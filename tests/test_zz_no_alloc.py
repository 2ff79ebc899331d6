"""rrhip_trace_closest_noalloc (closest hit staged through the library's
fixed per-device array) against rrhip_trace_closest on the same rays: bit
for bit on two calls in a row (the second smaller, so stale staging from
the first would show), at exactly the staging limit, and the rejection of
one ray more."""
import ctypes
import sys

import numpy as np
import pytest

import rrhip

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def test_noalloc_equals_the_allocating_entry_point(ctx):
    pos, idx = rrhip.gen_heightfield(256, 128, 5)
    scene = ctx.upload_mesh(pos, idx)
    lo, hi = pos.min(0), pos.max(0)
    for n, seed in ((400_000, 3), (77_777, 4), (rrhip.NOALLOC_MAX_RAYS, 5)):
        say(f"noalloc closest hit: heightfield 256x128, {n} rays, seed {seed}")
        rays = rrhip.gen_query_rays(lo, hi, 0.1, seed, n)
        got = scene.trace_closest_noalloc(rays)
        want = scene.trace_closest(rays)
        assert got.tobytes() == want.tobytes(), f"{n} rays: hits differ from rrhip_trace_closest"
        assert (got["prim"] != rrhip.NO_HIT).any()
    say("noalloc: empty batch and one ray beyond the limit")
    assert scene.trace_closest_noalloc(rays[:0]).shape == (0,)
    big = np.zeros(rrhip.NOALLOC_MAX_RAYS + 1, rrhip.RAY_DTYPE)
    hits = np.empty(big.shape[0], rrhip.HIT_DTYPE)
    st = rrhip.load().rrhip_trace_closest_noalloc(scene.h, big.ctypes.data_as(ctypes.c_void_p), big.shape[0],
                                                  hits.ctypes.data_as(ctypes.c_void_p))
    assert st == 1 and b"staging limit" in rrhip.load().rrhip_last_error()
    scene.close()
