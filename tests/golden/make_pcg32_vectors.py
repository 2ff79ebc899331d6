"""Writes pcg32_kat.json from an independent pure-Python-integer PCG32.

Restates render.rs src/raytracer/pt/sampler.rs:12-42 (Pcg32::new,
for_pixel_sample, next_u32) with Python ints masked to 64 bits, so the C++
oracle and the HIP sampler are checked against arithmetic that shares no
code with them. The seed-42/stream-54 row is PCG's published demo output.
"""
import json
import os

M64 = (1 << 64) - 1


class Pcg32:
    def __init__(self, seed, stream):
        self.state = 0
        self.inc = ((stream << 1) | 1) & M64
        self.next_u32()
        self.state = (self.state + seed) & M64
        self.next_u32()

    def next_u32(self):
        old = self.state
        self.state = (old * 6364136223846793005 + self.inc) & M64
        xs = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        return ((xs >> rot) | (xs << ((32 - rot) & 31))) & 0xFFFFFFFF


def mix(z):
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def for_pixel_sample(pixel, sample):
    return Pcg32(mix(((pixel * 0x9E3779B97F4A7C15) & M64) ^ sample), mix(sample) | 1)


def main():
    out = {
        "published_demo": {"seed": 42, "stream": 54,
                           "u32": [0xA15C02B7, 0x7B47F409, 0xBA1D3330, 0x83D2F293, 0xBFA4784B, 0xCBED606E]},
        "pixel_sample": [],
    }
    r = Pcg32(42, 54)
    assert [r.next_u32() for _ in range(6)] == out["published_demo"]["u32"]
    for pixel, sample in [(0, 0), (7, 3), (7, 4), (65535, 15), (2073599, 63), (123456789, 4095)]:
        g = for_pixel_sample(pixel, sample)
        out["pixel_sample"].append({"pixel": pixel, "sample": sample, "u32": [g.next_u32() for _ in range(8)]})
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pcg32_kat.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
