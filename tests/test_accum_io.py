"""Accumulation interchange (render.rs_amd/accum_io.py):

* RCKP checkpoints byte for byte as renderer.rs:222-235 lays them out
  (magic, three little-endian u32, f32 sums and counts), and refused on a
  wrong size, magic or resolution (renderer.rs:237-253);
* RCKD (f64) checkpoints round-trip every accumulator bit;
* accumulation EXRs: the writer's bytes against the OpenEXR layout (magic
  20000630, version 2, attribute list, line offset table, per-line
  channel planes in name order B, G, R, count), the reader against files
  assembled here independently with the format's RLE and ZIP coding
  (forward predictor and byte split written in this test), scanline and
  tiled, f16 and f32;
* merge: sample ranges [0, 3) and [3, 8) of the oracle's render merge to
  the mean of [0, 8) (deterministic per-(pixel, sample) seeding, accum.rs:
  1-7), up to the files' f32 rounding.
"""
import os
import struct
import zlib

import numpy as np
import pytest

import accum_io
import orc
import rrhip
import scenes


def rand_acc(npix, seed=1):
    rng = np.random.default_rng(seed)
    acc = rng.uniform(0, 10, (npix, rrhip.ACCUM_STRIDE))
    acc[:, 6] = rng.integers(0, 9, npix)
    return acc


def test_rckp_layout_and_refusals(tmp_path):
    p = str(tmp_path / "run.ckpt")
    acc = rand_acc(6)
    accum_io.write_rckp(p, 3, 2, 12, acc)
    b = open(p, "rb").read()
    assert b[:4] == b"RCKP" and b[4:16] == bytes([3, 0, 0, 0, 2, 0, 0, 0, 12, 0, 0, 0])
    assert len(b) == 16 + 6 * 16
    px1 = struct.unpack("<4f", b[16 + 16:16 + 32])
    np.testing.assert_array_equal(px1, np.float32([acc[1, 0], acc[1, 1], acc[1, 2], acc[1, 6]]))
    done, rgbn = accum_io.read_rckp(p, 3, 2)
    assert done == 12 and rgbn.dtype == np.float32
    assert not os.path.exists(str(tmp_path / "run.tmp"))
    assert accum_io.read_rckp(p, 2, 3) is None  # same size, other resolution
    assert accum_io.read_rckp(p, 3, 3) is None  # wrong size
    open(p, "r+b").write(b"XCKP")
    assert accum_io.read_rckp(p, 3, 2) is None
    assert accum_io.read_rckp(str(tmp_path / "missing"), 3, 2) is None


def test_rckd_round_trip_is_exact(tmp_path):
    p = str(tmp_path / "run.rckd")
    acc = rand_acc(20, 3)
    accum_io.write_rckd(p, 5, 4, 7, acc)
    done, back = accum_io.read_rckd(p, 5, 4)
    assert done == 7 and np.array_equal(back.view(np.uint64), acc.view(np.uint64))
    assert accum_io.read_rckd(p, 4, 5) is None
    accum_io.write_rckp(p, 5, 4, 7, acc)  # an RCKP under that name is not an RCKD
    assert accum_io.read_rckd(p, 5, 4) is None


def test_exr_writer_layout(tmp_path):
    p = str(tmp_path / "a.exr")
    w, h = 3, 2
    acc = rand_acc(w * h, 5)
    accum_io.write_accum_exr(p, w, h, acc)
    b = open(p, "rb").read()
    assert b[:4] == bytes([0x76, 0x2F, 0x31, 0x01]) and b[4:8] == bytes([2, 0, 0, 0])
    assert b"channels\0chlist\0" in b and b"compression\0compression\0" in b
    # the channel list: B, G, R, count, each FLOAT (2), linear 0, sampling 1 x 1
    i = b.index(b"chlist\0") + 7
    n = struct.unpack("<i", b[i:i + 4])[0]
    chl = b[i + 4:i + 4 + n]
    want = b"".join(c + b"\0" + struct.pack("<iiii", 2, 0, 1, 1) for c in (b"B", b"G", b"R", b"count")) + b"\0"
    assert chl == want
    end = b.index(b"screenWindowWidth\0float\0") + len(b"screenWindowWidth\0float\0") + 8 + 1
    offs = struct.unpack("<2Q", b[end:end + 16])
    assert offs[0] == end + 16
    y, size = struct.unpack("<ii", b[offs[1]:offs[1] + 8])
    assert (y, size) == (1, 4 * 4 * w)
    line = np.frombuffer(b, "<f4", 4 * w, offs[1] + 8).reshape(4, w)
    row = acc[w:2 * w]
    np.testing.assert_array_equal(line, np.float32([row[:, 2], row[:, 1], row[:, 0], row[:, 6]]))
    w2, h2, rgbn = accum_io.read_accum_exr(p)
    assert (w2, h2) == (w, h)
    np.testing.assert_array_equal(rgbn, np.float32(accum_io.to_rgbn(acc)).astype(np.float64))


# ---- files assembled independently from the OpenEXR layout -------------------------------

def _predict(raw):
    """Forward coding of ZIP / RLE data: even bytes then odd bytes, then
    each byte replaced by its difference from the previous one + 128."""
    a = np.frombuffer(raw, np.uint8)
    t = np.r_[a[0::2], a[1::2]].astype(np.int64)
    d = t.copy()
    d[1:] = (t[1:] - t[:-1] + 128) & 0xFF
    return d.astype(np.uint8).tobytes()


def _rle(raw):
    out, i = bytearray(), 0
    while i < len(raw):
        j = i
        while j < len(raw) and j - i < 127 and raw[j] == raw[i]:
            j += 1
        if j - i >= 3:
            out += struct.pack("b", j - i - 1) + raw[i:i + 1]
            i = j
        else:
            k = i
            while k < len(raw) and k - i < 127 and not (k + 2 < len(raw) and raw[k] == raw[k + 1] == raw[k + 2]):
                k += 1
            out += struct.pack("b", -(k - i)) + raw[i:k]
            i = k
    return bytes(out)


def _file(w, h, chans, comp, blocks, tiles=None):
    def attr(n, t, d):
        return n + b"\0" + t + b"\0" + struct.pack("<i", len(d)) + d
    chl = b"".join(n.encode() + b"\0" + struct.pack("<iB3xii", pt, 0, 1, 1) for n, pt in chans) + b"\0"
    head = struct.pack("<ii", 20000630, 2 | (0x200 if tiles else 0))
    head += attr(b"channels", b"chlist", chl) + attr(b"compression", b"compression", bytes([comp]))
    head += attr(b"dataWindow", b"box2i", struct.pack("<iiii", 0, 0, w - 1, h - 1))
    head += attr(b"displayWindow", b"box2i", struct.pack("<iiii", 0, 0, w - 1, h - 1))
    head += attr(b"lineOrder", b"lineOrder", b"\0")
    if tiles:
        head += attr(b"tiles", b"tiledesc", struct.pack("<IIB", tiles[0], tiles[1], 0))
    head += b"\0"
    base = len(head) + 8 * len(blocks)
    offs, body = [], b""
    for blk in blocks:
        offs.append(base + len(body))
        body += blk
    return head + struct.pack("<%dQ" % len(blocks), *offs) + body


def _plane_bytes(planes, chans, r0, r1, c0, c1):
    out = b""
    for r in range(r0, r1):
        for n, pt in chans:
            dt = "<f2" if pt == 1 else "<f4"
            out += planes[n][r, c0:c1].astype(dt).tobytes()
    return out


@pytest.mark.parametrize("comp,ptype", [(0, 2), (1, 2), (2, 1), (3, 2), (3, 1)])
def test_exr_reader_scanline_compressions(tmp_path, comp, ptype):
    w, h = 7, 19
    rng = np.random.default_rng(comp * 10 + ptype)
    chans = [("B", ptype), ("G", ptype), ("R", ptype), ("count", ptype)]
    planes = {n: np.round(rng.uniform(0, 4, (h, w)), 1) for n, _ in chans}
    planes["count"][:] = 16.0  # long runs for RLE
    lines = 16 if comp == 3 else 1
    blocks = []
    for y in range(0, h, lines):
        raw = _plane_bytes(planes, chans, y, min(y + lines, h), 0, w)
        data = raw if comp == 0 else (_rle(_predict(raw)) if comp == 1 else zlib.compress(_predict(raw)))
        if len(data) >= len(raw):
            data = raw  # the format stores a block raw when coding does not shrink it
        blocks.append(struct.pack("<ii", y, len(data)) + data)
    p = str(tmp_path / "s.exr")
    open(p, "wb").write(_file(w, h, chans, comp, blocks))
    w2, h2, rgbn = accum_io.read_accum_exr(p)
    dt = np.float16 if ptype == 1 else np.float32
    for k, n in enumerate(("R", "G", "B", "count")):
        np.testing.assert_array_equal(rgbn[:, k], planes[n].astype(dt).astype(np.float64).reshape(-1))


def test_exr_reader_tiled_zip(tmp_path):
    w, h, tw, th = 10, 7, 4, 4
    rng = np.random.default_rng(9)
    chans = [("B", 2), ("G", 2), ("R", 2), ("count", 2)]
    planes = {n: rng.uniform(0, 1, (h, w)) for n, _ in chans}
    blocks = []
    for ty in range((h + th - 1) // th):
        for tx in range((w + tw - 1) // tw):
            raw = _plane_bytes(planes, chans, ty * th, min((ty + 1) * th, h), tx * tw, min((tx + 1) * tw, w))
            data = zlib.compress(_predict(raw))
            data = raw if len(data) >= len(raw) else data
            blocks.append(struct.pack("<iiiii", tx, ty, 0, 0, len(data)) + data)
    p = str(tmp_path / "t.exr")
    open(p, "wb").write(_file(w, h, chans, 3, blocks, tiles=(tw, th)))
    _, _, rgbn = accum_io.read_accum_exr(p)
    for k, n in enumerate(("R", "G", "B", "count")):
        np.testing.assert_array_equal(rgbn[:, k], planes[n].astype(np.float32).astype(np.float64).reshape(-1))


def test_merge_of_sample_ranges_equals_one_render(tmp_path):
    s = scenes.cornell_fixture(24, 24)
    d = s.desc()
    prm = rrhip.RenderParams.make(max_depth=5)
    a, _ = orc.render(d, prm, 0, 3)
    b, _ = orc.render(d, prm, 3, 5)
    full, _ = orc.render(d, prm, 0, 8)
    pa, pb = str(tmp_path / "a.exr"), str(tmp_path / "b.exr")
    accum_io.write_accum_exr(pa, 24, 24, a)
    accum_io.write_accum_exr(pb, 24, 24, b)
    w, h, mean = accum_io.merge([pa, pb])
    assert (w, h) == (24, 24)
    want = full[:, 0:3] / full[:, 6:7]
    np.testing.assert_allclose(mean, want, rtol=2e-6, atol=1e-7)
    with pytest.raises(ValueError):
        accum_io.merge([])
    accum_io.write_accum_exr(pb, 12, 48, b)
    with pytest.raises(ValueError):
        accum_io.merge([pa, pb])


class _FakeScene:
    npix = 6


def test_load_accum_refuses_a_wrong_shape():
    """Session.load_accum hands the C call a pointer to npix * 8 doubles:
    any other shape is refused before the library is reached."""
    se = rrhip.Session.__new__(rrhip.Session)
    se.scene, se.h = _FakeScene(), None
    for bad in (np.zeros((5, rrhip.ACCUM_STRIDE)), np.zeros((6, 4)), np.zeros(6 * rrhip.ACCUM_STRIDE)):
        with pytest.raises(ValueError):
            se.load_accum(bad)


def test_render_checkpointed_refuses_another_resolution(tmp_path):
    class FakeSession:
        scene = _FakeScene()

        def load_accum(self, acc=None):
            raise AssertionError("must not be reached")

    p = str(tmp_path / "run.rckd")
    accum_io.write_rckd(p, 2, 3, 4, rand_acc(6))
    with pytest.raises(ValueError):
        accum_io.render_checkpointed(FakeSession(), 3, 3, 8, p)
