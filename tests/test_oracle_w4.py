"""Pins the oracle's int4 weight format (quantize.cpp:8-65, generalised with
zero-points) and its w4a16 product (w4a16.cpp:178-184 restated on the CPU).

Independent checks: hand-built packed words, numpy's binary16 arithmetic for
the one-rounding dequantisation over every (q, z, scale bit pattern), the
quantiser's defining properties, and the committed golden fixture.
"""
import json
import os

import numpy as np
import pytest

from libthom_amd import inputs

RNG = np.random.default_rng(7)
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "w4_tiny.json")


def test_pack_bit_layout(orc):
    # q[k][n] = (k + n) % 16 -> packed word of column n holds k at bits 4k
    K, N = 16, 8
    q = np.fromfunction(lambda k, n: (k + n) % 16, (K, N), dtype=np.int64).astype(np.uint8)
    p = orc.w4_pack(q)
    assert p.shape == (2, N)
    assert p[0, 0] == 0x76543210 and p[1, 0] == 0xFEDCBA98
    assert p[0, 1] == 0x87654321
    assert np.array_equal(orc.w4_unpack(p), q)
    assert np.array_equal(inputs.pack_w4(q), p)


def test_pack_zeros_layout(orc):
    z = np.tile(np.arange(16, dtype=np.uint8), (2, 1))    # 2 groups x 16 columns
    qz = orc.w4_pack_zeros(z)
    assert qz.shape == (2, 2) and qz[0, 0] == 0x76543210 and qz[0, 1] == 0xFEDCBA98
    assert np.array_equal(inputs.pack_zeros(z), qz)


def test_dequant_exhaustive_one_rounding(orc):
    # K = 16 rows carry q = 0..15; columns enumerate (scale bit pattern, z)
    group = 16
    s_bits = np.arange(65536, dtype=np.uint16)
    s_bits = s_bits[(s_bits & 0x7C00) != 0x7C00]          # finite scales only
    S = s_bits.size
    N = S * 16
    z = np.tile(np.arange(16, dtype=np.uint8), S)[None, :]
    scales = np.repeat(s_bits, 16)[None, :].view(np.float16)
    q = np.repeat(np.arange(16, dtype=np.uint8)[:, None], N, axis=1)
    W = orc.w4_dequant(orc.w4_pack(q), scales, orc.w4_pack_zeros(z), group)
    want = (np.arange(16, dtype=np.float32)[:, None] - z.astype(np.float32)) * scales.astype(np.float32)
    want = want.astype(np.float16)          # the exact product rounded once to binary16
    assert np.array_equal(W.view(np.uint16), want.view(np.uint16))


def test_dequant_symmetric_default(orc):
    q = RNG.integers(0, 16, (256, 24)).astype(np.uint8)
    s = RNG.uniform(0.005, 0.02, (2, 24)).astype(np.float16)
    W = orc.w4_dequant(orc.w4_pack(q), s, None, 128)
    g = np.repeat(np.arange(2), 128)
    want = ((q.astype(np.float32) - 8) * s[g].astype(np.float32)).astype(np.float16)
    assert np.array_equal(W, want)


def test_quantize_properties(orc):
    K, N, group = 256, 40, 128
    W = RNG.uniform(-1, 1, (K, N)).astype(np.float16)
    W[:128, 5] = 0                                          # an all-zero group
    W[7, 9] = np.float16(0.0)
    q, s = orc.w4_quantize_sym(W, group)
    # scale = fp16(2 max|W| / 15), or 1 for an all-zero group (quantize.cpp:22-25)
    for g in range(K // group):
        mx = np.abs(W[g * group:(g + 1) * group].astype(np.float32)).max(axis=0)
        want = np.where(mx > 0, (np.float32(2.0) * mx / np.float32(15.0)).astype(np.float16), np.float16(1.0))
        assert np.array_equal(s[g], want)
    assert np.all(q[:128, 5] == 8)
    assert q.max() <= 15
    Wd = orc.w4_dequant(orc.w4_pack(q), s, None, group).astype(np.float32)
    sc = np.repeat(s.astype(np.float32), group, axis=0)
    # round-to-nearest: |W - dequant| <= scale/2 (plus the fp16 rounding) unless clipped at q=0/15
    inner = (q > 0) & (q < 15)
    assert np.all(np.abs(W.astype(np.float32) - Wd)[inner] <= sc[inner] * 0.5 + 1e-3)


def test_quantize_round_half_even(orc):
    # one group of 8 rows, scale exactly 2*7.5/15 = 1.0: values +-0.5, 1.5, 2.5 are ties
    W = np.array([7.5, 0.5, 1.5, 2.5, -0.5, -1.5, -2.5, 0.0], np.float16)[:, None]
    q, s = orc.w4_quantize_sym(W, 8)
    assert float(s[0, 0]) == 1.0
    # nearbyint(v) + 8 with ties to even: 7.5->8(+8=16->15 clipped), 0.5->0, 1.5->2, 2.5->2
    assert list(q[:, 0]) == [15, 8, 10, 10, 8, 6, 6, 8]


def test_golden_fixture(orc):
    with open(GOLDEN) as f:
        g = json.load(f)
    packed = np.array(g["packed"], dtype=np.uint32)
    scales = np.array(g["scales_bits"], dtype=np.uint16).view(np.float16)
    qz = np.array(g["qzeros"], dtype=np.uint32)
    W = orc.w4_dequant(packed, scales, qz, g["group"])
    assert W.view(np.uint16).tolist() == g["dequant_bits"]


def test_w4a16_rows_small_integers_exact(orc):
    # scale 1, z = 8: W = q - 8 in [-8, 7]; A small integers: everything exact
    K, N, M = 128, 16, 3
    q = RNG.integers(0, 16, (K, N)).astype(np.uint8)
    Wd = orc.w4_dequant(orc.w4_pack(q), np.ones((1, N), np.float16), None, 128)
    A = RNG.integers(-3, 4, (M, K)).astype(np.float16)
    C = orc.w4a16_rows(A, Wd)
    ref = (A.astype(np.int64) @ (q.astype(np.int64) - 8)).astype(np.float16)
    assert np.array_equal(C, ref)
    assert np.array_equal(orc.w4a16_rows(A, Wd, 1, 2), ref[1:3])


def test_quantize_asym_exact_round_trip(orc):
    # W = (q - z) s with s a power of two and each (group, column) holding both
    # q = 0 and q = 15: the quantizer must give back q, z and s exactly
    K, N, group = 256, 24, 64
    rng = np.random.default_rng(31)
    q = rng.integers(0, 16, (K, N)).astype(np.uint8)
    for g in range(K // group):
        q[g * group, :] = 0
        q[g * group + 1, :] = 15
    z = rng.integers(0, 16, (K // group, N)).astype(np.uint8)
    s = (2.0 ** rng.integers(-12, 2, (K // group, N))).astype(np.float16)
    gi = np.repeat(np.arange(K // group), group)
    W = ((q.astype(np.float32) - z[gi]) * s[gi].astype(np.float32)).astype(np.float16)
    q2, z2, s2 = orc.w4_quantize_asym(W, group)
    assert np.array_equal(s2, s) and np.array_equal(z2, z) and np.array_equal(q2, q)


def test_quantize_asym_properties(orc):
    K, N, group = 256, 40, 128
    rng = np.random.default_rng(32)
    W = rng.uniform(-1, 1, (K, N)).astype(np.float16)
    W[:, 3] = np.abs(W[:, 3])                               # all positive: z = 0
    W[:, 4] = -np.abs(W[:, 4])                              # all negative: z = 15
    W[:128, 5] = 0                                          # an all-zero group: s = 1, z = 0, q = 0
    q, z, s = orc.w4_quantize_asym(W, group)
    assert np.all(z[:, 3] == 0) and np.all(z[:, 4] == 15)
    assert s[0, 5] == 1.0 and z[0, 5] == 0 and np.all(q[:128, 5] == 0)
    assert q.max() <= 15 and z.max() <= 15
    Wd = orc.w4_dequant(orc.w4_pack(q), s, orc.w4_pack_zeros(z), group).astype(np.float32)
    sc = np.repeat(s.astype(np.float32), group, axis=0)
    # every value lies inside [lo, hi], so nothing clips: error <= scale / 2
    # (the fp16 scale may round down by half an ulp, and dequant rounds once)
    assert np.all(np.abs(W.astype(np.float32) - Wd) <= sc * 0.5 * (1 + 2 ** -10) + np.abs(Wd) * 2 ** -11 + 1e-7)
    # zero is representable: W = 0 dequantises to exactly 0
    zero = W == 0
    assert np.all(Wd[zero] == 0)


def test_quantize_asym_round_half_even(orc):
    # lo = -7, hi = 8: scale (8 + 7) / 15 = 1, z = 7; ties round to even
    W = np.array([-7.0, 8.0, 0.5, 1.5, 2.5, -0.5, -1.5, 0.0], np.float16)[:, None]
    q, z, s = orc.w4_quantize_asym(W, 8)
    assert float(s[0, 0]) == 1.0 and int(z[0, 0]) == 7
    assert list(q[:, 0]) == [0, 15, 7, 9, 9, 7, 5, 7]


def test_exact_inputs_expose_shortcut(orc):
    # the bit-exact GPU test's inputs: the oracle equals an exact float64
    # evaluation of sum A * fp16((q - z) s), and the scale-after-the-sum order
    # of w4a16_naive.cu:12-18 differs from it in many outputs
    from _w4_exact import exact_problem, shortcut
    for M, group in ((1, 32), (16, 32), (64, 64)):
        A, q, z, s = exact_problem(M, 288, 64, group, seed=1000 + 7 * M + group)
        packed, qz = orc.w4_pack(q), orc.w4_pack_zeros(z)
        W = orc.w4_dequant(packed, s, qz, group)
        want = orc.w4a16_rows(A, W)
        exact = (A.astype(np.float64) @ W.astype(np.float64)).astype(np.float32).astype(np.float16)
        assert np.array_equal(want.view(np.uint16), exact.view(np.uint16))
        differ = (shortcut(A, q, z, s, group).view(np.uint16) != want.view(np.uint16)).mean()
        assert differ > 0.1, differ
