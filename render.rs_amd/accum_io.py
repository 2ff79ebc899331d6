"""Accumulation interchange: the files render.rs writes to resume a render
and to merge sample ranges rendered on different machines.

* RCKP checkpoints (src/raytracer/metal/renderer.rs:219-254): "RCKP",
  little-endian u32 width, height and samples done, then width x height x
  4 f32 (sum r, g, b and the sample count per pixel). `read_rckp` refuses
  (returns None) a file of the wrong size, magic or resolution, as
  `load_checkpoint` does; `write_rckp` writes a temporary file and renames
  it over the checkpoint.
* RCKD checkpoints: the same idea for this backend's f64 accumulator
  (all 8 moments of rrhip_read_accum), so a resumed render continues bit
  for bit: "RCKD", u32 version 1, width, height, samples done, values per
  pixel (8), then the f64 values.
* Accumulation EXRs (src/output/accum.rs:14-119): one part with channels
  B, G, R (the radiance sums) and count, sorted by name, f32; `merge`
  averages any number of them as `render merge` does (sum / count, 0
  where nothing was counted). The writer emits uncompressed scanlines;
  the reader takes single-part scanline or one-level tiled files with no,
  RLE, ZIPS or ZIP compression and f16 or f32 channels (the OpenEXR file
  layout), which covers what render.rs's exr writer produces.

Loading an RCKP or EXR accumulator into a session sets the sums and
counts; their second moments are unknown (zero), which only matters to
adaptive sampling's variance test.
"""
import os
import struct
import zlib

import numpy as np

import rrhip

RCKP_MAGIC = b"RCKP"
RCKD_MAGIC = b"RCKD"
EXR_MAGIC = 20000630


# ---- accumulator views ------------------------------------------------------------------

def to_rgbn(acc):
    """(npix, 8) f64 accumulator -> (npix, 4) sums r, g, b and count."""
    acc = np.asarray(acc, np.float64)
    return np.ascontiguousarray(np.c_[acc[:, 0:3], acc[:, 6]])


def from_rgbn(rgbn):
    """(npix, 4) sums and counts -> an (npix, 8) accumulator whose second
    moments are zero."""
    rgbn = np.asarray(rgbn, np.float64)
    acc = np.zeros((rgbn.shape[0], rrhip.ACCUM_STRIDE))
    acc[:, 0:3] = rgbn[:, 0:3]
    acc[:, 6] = rgbn[:, 3]
    return acc


def _atomic_write(path, data):
    tmp = os.path.splitext(path)[0] + ".tmp"
    with open(tmp, "wb") as f:
        f.write(data)
    os.replace(tmp, path)


# ---- RCKP / RCKD ------------------------------------------------------------------------

def write_rckp(path, width, height, done, acc):
    """render.rs's checkpoint (renderer.rs:222-235) from an (npix, 8) f64
    accumulator: the sums and counts rounded to f32."""
    rgbn = to_rgbn(acc)
    if rgbn.shape[0] != width * height:
        raise ValueError("accumulator size does not match the resolution")
    _atomic_write(path, RCKP_MAGIC + struct.pack("<III", width, height, done) + rgbn.astype("<f4").tobytes())


def read_rckp(path, width, height):
    """(done, (npix, 4) f32) or None (renderer.rs:237-253)."""
    try:
        with open(path, "rb") as f:
            b = f.read()
    except OSError:
        return None
    if len(b) != 16 + width * height * 16 or b[:4] != RCKP_MAGIC:
        return None
    w, h, done = struct.unpack("<III", b[4:16])
    if (w, h) != (width, height):
        return None
    return done, np.frombuffer(b, "<f4", offset=16).reshape(-1, 4).copy()


def write_rckd(path, width, height, done, acc):
    acc = np.ascontiguousarray(acc, "<f8")
    if acc.shape != (width * height, rrhip.ACCUM_STRIDE):
        raise ValueError("accumulator shape does not match the resolution")
    head = RCKD_MAGIC + struct.pack("<IIIII", 1, width, height, done, rrhip.ACCUM_STRIDE)
    _atomic_write(path, head + acc.tobytes())


def read_rckd(path, width, height):
    """(done, (npix, 8) f64) or None for a missing, foreign or mismatched file."""
    try:
        with open(path, "rb") as f:
            b = f.read()
    except OSError:
        return None
    if len(b) < 24 or b[:4] != RCKD_MAGIC:
        return None
    ver, w, h, done, stride = struct.unpack("<IIIII", b[4:24])
    if ver != 1 or (w, h) != (width, height) or stride != rrhip.ACCUM_STRIDE:
        return None
    if len(b) != 24 + width * height * stride * 8:
        return None
    return done, np.frombuffer(b, "<f8", offset=24).reshape(-1, stride).copy()


def render_checkpointed(session, width, height, spp, path, every=4):
    """render_pt_checkpointed (renderer.rs:765-808) over a session: resume
    from a matching checkpoint at `path` (RCKD, else render.rs's RCKP),
    render the remaining samples `every` at a time and write a checkpoint
    after each step, in the format the resumed one had (RCKP stays RCKP, so
    render.rs can still resume the file; RCKD otherwise). Without a
    checkpoint the session's accumulator is cleared first, as a fresh
    render_pt_checkpointed starts from zero. Returns the samples rendered
    by this call."""
    if width * height != session.scene.npix:
        raise ValueError(f"render_checkpointed: {width} x {height} does not match the session's "
                         f"{session.scene.npix} pixels")
    start = 0
    write = write_rckd
    got = read_rckd(path, width, height)
    if got is not None:
        start, acc = got
    else:
        got = read_rckp(path, width, height)
        if got is not None:
            start, rgbn = got
            acc = from_rgbn(rgbn)
            write = write_rckp
    if got is not None:
        start = min(start, spp)
        session.load_accum(acc)
    else:
        session.load_accum(None)
    s = start
    while s < spp:
        step = min(every, spp - s)
        session.render(s, step)
        s += step
        write(path, width, height, s, session.accum())
    return spp - start


# ---- OpenEXR ----------------------------------------------------------------------------

def _attr(name, typ, data):
    return name.encode() + b"\0" + typ.encode() + b"\0" + struct.pack("<i", len(data)) + data


def write_accum_exr(path, width, height, acc):
    """accum.rs:14-48: channels B, G, R (sums) and count, f32, one part
    named "accum"; uncompressed scanlines, increasing y."""
    rgbn = to_rgbn(acc).astype("<f4")
    if rgbn.shape[0] != width * height:
        raise ValueError("accumulator size does not match the resolution")
    names = ["B", "G", "R", "count"]
    cols = [2, 1, 0, 3]
    chl = b"".join(n.encode() + b"\0" + struct.pack("<iB3xii", 2, 0, 1, 1) for n in names) + b"\0"
    box = struct.pack("<iiii", 0, 0, width - 1, height - 1)
    head = struct.pack("<ii", EXR_MAGIC, 2)
    head += _attr("channels", "chlist", chl)
    head += _attr("compression", "compression", b"\0")
    head += _attr("dataWindow", "box2i", box)
    head += _attr("displayWindow", "box2i", box)
    head += _attr("lineOrder", "lineOrder", b"\0")
    head += _attr("name", "string", b"accum")
    head += _attr("pixelAspectRatio", "float", struct.pack("<f", 1.0))
    head += _attr("screenWindowCenter", "v2f", struct.pack("<ff", 0.0, 0.0))
    head += _attr("screenWindowWidth", "float", struct.pack("<f", 1.0))
    head += b"\0"
    img = rgbn.reshape(height, width, 4)
    line = 4 * 4 * width
    table = len(head) + 8 * height
    offsets = [table + y * (8 + line) for y in range(height)]
    body = b"".join(struct.pack("<ii", y, line) + np.ascontiguousarray(img[y][:, cols].T).tobytes()
                    for y in range(height))
    _atomic_write(path, head + struct.pack("<%dQ" % height, *offsets) + body)


def _undo_predictor(raw):
    """OpenEXR's ZIP / RLE post-processing, reversed: running sum of the
    deltas (offset 128), then the two byte halves re-interleaved."""
    d = np.frombuffer(raw, np.uint8).astype(np.int64)
    d[1:] -= 128
    t = (np.cumsum(d) & 0xFF).astype(np.uint8)
    half = (len(t) + 1) // 2
    out = np.empty_like(t)
    out[0::2] = t[:half]
    out[1::2] = t[half:]
    return out.tobytes()


def _rle_decode(data, expect):
    out = bytearray()
    i = 0
    while i < len(data) and len(out) < expect:
        c = struct.unpack("b", data[i:i + 1])[0]
        i += 1
        if c < 0:
            out += data[i:i - c]
            i -= c
        else:
            out += data[i:i + 1] * (c + 1)
            i += 1
    if len(out) != expect:
        raise ValueError("EXR: RLE block decodes to the wrong size")
    return bytes(out)


def _parse_header(b):
    if len(b) < 8 or struct.unpack("<i", b[:4])[0] != EXR_MAGIC:
        raise ValueError("not an OpenEXR file")
    ver = struct.unpack("<i", b[4:8])[0]
    if ver & 0xFF != 2 or ver & 0x1000 or ver & 0x800:
        raise ValueError("EXR: only single-part, non-deep files are read")
    pos, attrs = 8, {}
    while b[pos] != 0:
        e = b.index(b"\0", pos)
        name = b[pos:e].decode()
        e2 = b.index(b"\0", e + 1)
        typ = b[e + 1:e2].decode()
        n = struct.unpack("<i", b[e2 + 1:e2 + 5])[0]
        attrs[name] = (typ, b[e2 + 5:e2 + 5 + n])
        pos = e2 + 5 + n
    return attrs, pos + 1


def _channels(data):
    out, pos = [], 0
    while data[pos] != 0:
        e = data.index(b"\0", pos)
        ptype, _, xs, ys = struct.unpack("<iB3xii", data[e + 1:e + 17])
        if (xs, ys) != (1, 1):
            raise ValueError("EXR: subsampled channels are not read")
        out.append((data[pos:e].decode(), ptype))
        pos = e + 17
    return out


_PTYPE = {1: ("<f2", 2), 2: ("<f4", 4)}


def read_accum_exr(path):
    """accum.rs:50-84: (width, height, (npix, 4) f64 sums r, g, b and count)."""
    with open(path, "rb") as f:
        b = f.read()
    attrs, pos = _parse_header(b)
    chans = _channels(attrs["channels"][1])
    for _, pt in chans:
        if pt not in _PTYPE:
            raise ValueError("EXR: unexpected u32 channel")
    comp = attrs["compression"][1][0]
    if comp not in (0, 1, 2, 3):
        raise ValueError(f"EXR: compression {comp} is not read")
    x0, y0, x1, y1 = struct.unpack("<iiii", attrs["dataWindow"][1])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    planes = {n: np.zeros((h, w)) for n, _ in chans}
    bpp = sum(_PTYPE[pt][1] for _, pt in chans)

    def decode(data, nbytes):
        if comp == 0 or len(data) == nbytes:
            return data
        if comp == 1:
            return _undo_predictor(_rle_decode(data, nbytes))
        return _undo_predictor(zlib.decompress(data))

    def scatter(raw, rows, cols, ry, cx):
        p = 0
        for r in range(rows):
            for n, pt in chans:
                dt, sz = _PTYPE[pt]
                planes[n][ry + r, cx:cx + cols] = np.frombuffer(raw, dt, cols, p)
                p += sz * cols

    if "tiles" in attrs:
        tw, th, mode = struct.unpack("<IIB", attrs["tiles"][1])
        if mode & 0xF != 0:
            raise ValueError("EXR: only one-level tiled files are read")
        nt = ((w + tw - 1) // tw) * ((h + th - 1) // th)
        offs = struct.unpack("<%dQ" % nt, b[pos:pos + 8 * nt])
        for o in offs:
            tx, ty, _, _, size = struct.unpack("<iiiii", b[o:o + 20])
            cols = min(tw, w - tx * tw)
            rows = min(th, h - ty * th)
            raw = decode(b[o + 20:o + 20 + size], rows * cols * bpp)
            scatter(raw, rows, cols, ty * th, tx * tw)
    else:
        lines = {0: 1, 1: 1, 2: 1, 3: 16}[comp]
        nb = (h + lines - 1) // lines
        offs = struct.unpack("<%dQ" % nb, b[pos:pos + 8 * nb])
        for o in offs:
            y, size = struct.unpack("<ii", b[o:o + 8])
            rows = min(lines, y1 - y + 1)
            raw = decode(b[o + 8:o + 8 + size], rows * w * bpp)
            scatter(raw, rows, w, y - y0, 0)
    for n in ("R", "G", "B", "count"):
        if n not in planes:
            raise ValueError(f"accum EXR missing channel {n}")
    rgbn = np.stack([planes[n].reshape(-1) for n in ("R", "G", "B", "count")], 1)
    return w, h, rgbn


def merge(paths):
    """render merge (accum.rs:86-119): the mean image (npix, 3) of the summed
    accumulation files, 0 where no sample was counted."""
    if not paths:
        raise ValueError("merge needs at least one accum file")
    w, h, tot = read_accum_exr(paths[0])
    for p in paths[1:]:
        w2, h2, a = read_accum_exr(p)
        if (w2, h2) != (w, h):
            raise ValueError(f"{p}: resolution mismatch")
        tot = tot + a
    n = tot[:, 3:4]
    return w, h, np.where(n > 0, tot[:, 0:3] / np.where(n > 0, n, 1.0), 0.0)
