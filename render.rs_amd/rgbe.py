"""Radiance RGBE (.hdr) reader for environment maps (the lat-long maps of
PxrDomeLight "lightColorMap", which render.rs opens with the image crate,
src/scene/envmap.rs:28-35): flat or new-style run-length encoded
scanlines, "-Y h +X w" orientation (top row first)."""
import numpy as np


def read_hdr(path):
    with open(path, "rb") as f:
        data = f.read()
    pos = 0
    while True:  # header lines up to the blank line
        end = data.index(b"\n", pos)
        line = data[pos:end]
        pos = end + 1
        if line.strip() == b"":
            break
    end = data.index(b"\n", pos)
    res = data[pos:end].split()
    pos = end + 1
    if res[0] != b"-Y" or res[2] != b"+X":
        raise ValueError("only -Y h +X w images are read")
    h, w = int(res[1]), int(res[3])
    buf = np.frombuffer(data, np.uint8, offset=pos)
    rgbe = np.empty((h, w, 4), np.uint8)
    i = 0
    for y in range(h):
        if w >= 8 and w < 32768 and buf[i] == 2 and buf[i + 1] == 2 and ((int(buf[i + 2]) << 8) | buf[i + 3]) == w:
            i += 4
            for c in range(4):
                x = 0
                while x < w:
                    n = int(buf[i])
                    i += 1
                    if n > 128:
                        n -= 128
                        rgbe[y, x:x + n, c] = buf[i]
                        i += 1
                    else:
                        rgbe[y, x:x + n, c] = buf[i:i + n]
                        i += n
                    x += n
        else:
            rgbe[y] = buf[i:i + 4 * w].reshape(w, 4)
            i += 4 * w
    e = rgbe[..., 3].astype(np.int32)
    scale = np.where(e > 0, np.ldexp(1.0, e - 136), 0.0)
    return (rgbe[..., :3].astype(np.float64) * scale[..., None]).astype(np.float32)
