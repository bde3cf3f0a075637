"""Writes rng_known_answers.json: known answers of the per-pixel RNG of raytrace.rs:108-149.

The values are computed here by a pure-Python restatement (independent of the C++
oracle) of murmur3_combine / murmur3_finalize / floatConstruct; the same numbers are
listed in the build survey's RNG appendix. Run: python tests/golden/make_rng_known_answers.py
"""
import json
import os
import struct

M = 0xFFFFFFFF


def combine(h, k):
    k = (k * 0x1B873593) & M
    h ^= k
    h = ((h << 13) | (h >> 19)) & M
    return (h * 5 + 0xE6546B64) & M


def finalize(h):
    h ^= h >> 16
    h = (h * 0x85EBCA6B) & M
    h ^= h >> 13
    h = (h * 0xC2B2AE35) & M
    h ^= h >> 16
    return h


def fconstruct(m):
    f = struct.unpack("<f", struct.pack("<I", (m & 0x7FFFFF) | 0x3F800000))[0]
    return struct.unpack("<f", struct.pack("<f", f - 1.0))[0]


def main():
    out = {"combine": [], "finalize": [], "streams": []}
    for h, k in [(0, 0), (0, 1), (0x3039, 0x2A6), (0xFFFFFFFF, 7), (123456789, 987654321)]:
        out["combine"].append([h, k, combine(h, k)])
    for h in [0, 1, combine(0, 0), 0xDEADBEEF, 0xFFFFFFFF]:
        out["finalize"].append([h, finalize(h)])
    for seed_inv, pixel in [(0, 0), (7, 12345), (3 * 16 + 2, 2073599), (1, 1)]:
        s = combine(seed_inv, pixel)
        out["streams"].append({"seed_inv": seed_inv, "pixel": pixel, "seed": s,
                               "u": [fconstruct(finalize(combine(s, k))) for k in range(6)]})
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "rng_known_answers.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(path)


if __name__ == "__main__":
    main()
