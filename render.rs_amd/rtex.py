"""Reader of render.rs's tiled, mip-mapped texture files (.tex, "RTEX").

Layout (render.rs src/texture/tex.rs:1-15): magic "RTEX"; u32 version (1),
width, height, channels (3), tile size, mip count; per mip u32 width,
height, tiles_x, tiles_y and a u64 byte offset; then each mip's tiles,
row-major, tile_size^2 * channels little-endian f32 each, edge tiles
padded by replication. The library builds its own mip chain from level 0
(the same 2x2 box filter), so scenes hand over level 0 only.
"""
import struct

import numpy as np


class TexError(ValueError):
    pass


def read_header(buf):
    if len(buf) < 28 or buf[:4] != b"RTEX":
        raise TexError("not a .tex file (bad magic)")
    version, width, height, channels, tile, nmips = struct.unpack_from("<6I", buf, 4)
    if version != 1:
        raise TexError(f"unsupported .tex version {version}")
    if channels != 3 or tile == 0 or nmips == 0:
        raise TexError(f"unsupported .tex layout: {channels} channels, tile {tile}, {nmips} mips")
    mips = []
    for k in range(nmips):
        w, h, tx, ty, off = struct.unpack_from("<4IQ", buf, 28 + 24 * k)
        mips.append((w, h, tx, ty, off))
    return width, height, tile, mips


def read_level(buf, level=0):
    """One mip level as an (h, w, 3) float32 array."""
    _, _, tile, mips = read_header(buf)
    if level >= len(mips):
        raise TexError(f"level {level} past the chain ({len(mips)} levels)")
    w, h, tx, ty, off = mips[level]
    n = tx * ty * tile * tile * 3
    if off + 4 * n > len(buf):
        raise TexError("truncated .tex file")
    tiles = np.frombuffer(buf, "<f4", n, off).reshape(ty, tx, tile, tile, 3)
    img = tiles.transpose(0, 2, 1, 3, 4).reshape(ty * tile, tx * tile, 3)
    return np.ascontiguousarray(img[:h, :w], np.float32)


def read_tex(path, levels=False):
    """Level 0 of a .tex file (all levels with levels=True)."""
    with open(path, "rb") as f:
        buf = f.read()
    if not levels:
        return read_level(buf, 0)
    _, _, _, mips = read_header(buf)
    return [read_level(buf, k) for k in range(len(mips))]


def _downsample(img):
    """One 2x2 box-filter level (odd sizes clamp the second tap), f32 sums
    in row-major tap order."""
    h, w = img.shape[:2]
    nh, nw = max(h // 2, 1), max(w // 2, 1)
    y0 = np.minimum(2 * np.arange(nh), h - 1)
    y1 = np.minimum(2 * np.arange(nh) + 1, h - 1)
    x0 = np.minimum(2 * np.arange(nw), w - 1)
    x1 = np.minimum(2 * np.arange(nw) + 1, w - 1)
    s = img[y0][:, x0] + img[y0][:, x1]
    s = s + img[y1][:, x0]
    s = s + img[y1][:, x1]
    return (s * np.float32(0.25)).astype(np.float32)


def write_tex(path, rgb, tile=32):
    """Writes level 0 (h x w x 3 linear rgb) and its box-filtered chain in
    the layout read_header reads (edge tiles padded by replication)."""
    img = np.ascontiguousarray(rgb, np.float32)
    if img.ndim != 3 or img.shape[2] != 3 or img.shape[0] == 0 or img.shape[1] == 0:
        raise TexError("write_tex needs an h x w x 3 image")
    levels = [img]
    while levels[-1].shape[0] > 1 or levels[-1].shape[1] > 1:
        levels.append(_downsample(levels[-1]))
    header = 28 + 24 * len(levels)
    parts, infos, off = [], [], header
    for lv in levels:
        h, w = lv.shape[:2]
        tx, ty = -(-w // tile), -(-h // tile)
        ys = np.minimum(np.arange(ty * tile), h - 1)
        xs = np.minimum(np.arange(tx * tile), w - 1)
        padded = lv[ys][:, xs]
        tiles = padded.reshape(ty, tile, tx, tile, 3).transpose(0, 2, 1, 3, 4)
        data = np.ascontiguousarray(tiles, "<f4").tobytes()
        infos.append(struct.pack("<4IQ", w, h, tx, ty, off))
        parts.append(data)
        off += len(data)
    with open(path, "wb") as f:
        f.write(b"RTEX" + struct.pack("<6I", 1, img.shape[1], img.shape[0], 3, tile, len(levels)))
        for i in infos:
            f.write(i)
        for p in parts:
            f.write(p)
