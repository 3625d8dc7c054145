"""Writes tests/golden/w4_tiny.json: a tiny packed-int4 tensor with scales,
zero-points and its expected fp16 dequantisation.

The expected values are computed here with numpy's IEEE binary16 arithmetic
straight from the format definition (quantize.cpp:42-65 with per-(group,
column) zero-points): W[k][n] = fp16((q - z) * s), one rounding.  No library
or oracle code is used, so the fixture is an independent pin of both.
Run: python tests/golden/make_golden.py
"""
import json
import os

import numpy as np

K, N, GROUP = 16, 8, 8
# packed[k/8][n]: row k of column n in bits 4*(k%8)
words = [
    [0x76543210, 0xFEDCBA98, 0x00000000, 0xFFFFFFFF, 0x0F0F0F0F, 0x88888888, 0x12345678, 0x9ABCDEF0],
    [0x13579BDF, 0x02468ACE, 0x80000001, 0x7FFFFFFE, 0xA5A5A5A5, 0x5A5A5A5A, 0x01234567, 0xFEDCBA98],
]
# scales [K/GROUP][N] as binary16 bit patterns: 1.0, ~0.1, min subnormal, max
# normal, -2.0, 0.005-ish, 2^-14, 0.0199-ish
scale_bits = [
    [0x3C00, 0x2E66, 0x0001, 0x7BFF, 0xC000, 0x1D1F, 0x0400, 0x2518],
    [0x3800, 0x211F, 0x03FF, 0x5BFF, 0x8001, 0x2000, 0x3555, 0x1C00],
]
# zero-points [K/GROUP][N]
zeros = [[0, 15, 8, 1, 7, 8, 3, 12], [15, 0, 4, 9, 8, 2, 11, 6]]


def main():
    packed = np.array(words, dtype=np.uint32)
    scales = np.array(scale_bits, dtype=np.uint16).view(np.float16)
    z = np.array(zeros, dtype=np.uint32)
    qz = np.zeros((K // GROUP, N // 8), dtype=np.uint32)
    for n in range(N):
        qz[:, n // 8] |= z[:, n] << np.uint32(4 * (n % 8))
    q = np.zeros((K, N), dtype=np.int64)
    for k in range(K):
        q[k] = (packed[k // 8] >> np.uint32(4 * (k % 8))) & 15
    g = np.arange(K) // GROUP
    prod = (q - z[g].astype(np.int64)).astype(np.float32) * scales[g].astype(np.float32)   # exact in fp32
    with np.errstate(over="ignore"):
        deq = prod.astype(np.float16)     # 15 * 65504 overflows to inf, as IEEE says
    out = {
        "doc": "W[k][n] = fp16((q - z) * s); packed[k/8][n] bits 4*(k%8); qzeros[g][n/8] bits 4*(n%8)",
        "K": K, "N": N, "group": GROUP,
        "packed": packed.tolist(),
        "scales_bits": np.array(scale_bits, dtype=np.uint16).tolist(),
        "qzeros": qz.tolist(),
        "dequant_bits": deq.view(np.uint16).tolist(),
    }
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "w4_tiny.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
