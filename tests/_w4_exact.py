"""Inputs on which every w4a16 FP32 sum is exact, so the library must match the
oracle bit for bit, and the scale-after-the-sum shortcut (w4a16_naive.cu:12-18,
fp16(s * sum A (q - z)) instead of fp16(sum A * fp16((q - z) s))) must not.

A in {-2..2}, K <= 64, scales uniform in [0.005, 0.02] as fp16 (not powers of
two).  Each dequantised weight fp16((q - z) s) has |w| <= 0.3 and ulp >= 2^-18,
so every product A w is an exact FP32 number with ulp >= 2^-18, and every
partial sum of at most 64 of them is below 2^6: 24 significant bits cover the
range, so the FP32 sums are exact in any order.
"""
import numpy as np


def exact_problem(M, N, K, group, seed):
    assert K <= 64 and K % group == 0
    rng = np.random.default_rng(seed)
    q = rng.integers(0, 16, (K, N)).astype(np.uint8)
    z = rng.integers(0, 16, (K // group, N)).astype(np.uint8)
    s = rng.uniform(0.005, 0.02, (K // group, N)).astype(np.float16)
    A = rng.integers(-2, 3, (M, K)).astype(np.float16)
    return A, q, z, s


def shortcut(A, q, z, s, group):
    """The non-bit-exact order: per group, s * sum_k A (q - z), then fp16."""
    K = q.shape[0]
    acc = np.zeros((A.shape[0], q.shape[1]), np.float64)
    for g in range(K // group):
        d = q[g * group:(g + 1) * group].astype(np.float64) - z[g].astype(np.float64)
        acc += s[g].astype(np.float64) * (A[:, g * group:(g + 1) * group].astype(np.float64) @ d)
    return acc.astype(np.float32).astype(np.float16)
