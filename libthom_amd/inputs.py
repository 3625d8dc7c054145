"""Seeded synthetic inputs (host side, numpy).

The generator replaces the reference's cuRAND (every handler seeds it with
1234ULL, e.g. reduction.cpp:140-141).  Element i of stream s under seed is
output i of a SplitMix64 run started at seed + s * 0xD1B54A32D192ED03;
u = (x >> 40) * 2^-24 in [0, 1).  Uniform values are bit-identical to
lthip_fill_uniform_* on the GPU.  Normal values use Box-Muller in float64 and
may differ from another libm in the last bit.
"""
from __future__ import annotations

import math

import numpy as np

SEED = 1234
GAMMA = 0x9E3779B97F4A7C15
STREAM_MUL = 0xD1B54A32D192ED03
MASK = (1 << 64) - 1

# stream ids of the benchmark configurations
S_REDUCE, S_SGEMV_A, S_SGEMV_X, S_DGEMM_A, S_DGEMM_B, S_DGEMM_RS, S_DGEMM_CS = 1, 2, 3, 4, 5, 6, 7
S_BATCH_DIMS, S_W4_A, S_W4_Q, S_W4_S, S_W4_Z, S_SDOT_X, S_SDOT_Y = 8, 9, 10, 11, 12, 13, 14
S_BATCH_BASE = 1000


def rng_u64(n: int, seed: int = SEED, stream: int = 0, start: int = 0) -> np.ndarray:
    base = (seed + stream * STREAM_MUL) & MASK
    i = np.arange(start + 1, start + n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(base) + i * np.uint64(GAMMA)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def unit(n: int, seed: int = SEED, stream: int = 0) -> np.ndarray:
    return (rng_u64(n, seed, stream) >> np.uint64(40)).astype(np.float64) * 2.0 ** -24


def uniform_f64(n, lo, hi, seed=SEED, stream=0):
    return lo + (hi - lo) * unit(n, seed, stream)


def uniform_f32(n, lo, hi, seed=SEED, stream=0):
    return uniform_f64(n, lo, hi, seed, stream).astype(np.float32)


def uniform_f16(n, lo, hi, seed=SEED, stream=0):
    return uniform_f64(n, lo, hi, seed, stream).astype(np.float32).astype(np.float16)


def normal_f64(n, seed=SEED, stream=0):
    m = n + (n & 1)
    x = rng_u64(m, seed, stream)
    a, b = x[0::2], x[1::2]
    u1 = ((a >> np.uint64(11)) + np.uint64(1)).astype(np.float64) * 2.0 ** -53
    u2 = (b >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    r = np.sqrt(-2.0 * np.log(u1))
    out = np.empty(m, dtype=np.float64)
    out[0::2] = r * np.cos(2.0 * math.pi * u2)
    out[1::2] = r * np.sin(2.0 * math.pi * u2)
    return out[:n]


def int_uniform(n, lo, hi, seed=SEED, stream=0):
    span = np.uint64(hi - lo + 1)
    x = rng_u64(n, seed, stream) >> np.uint64(32)
    return (lo + ((x * span) >> np.uint64(32)).astype(np.int64)).astype(np.int64)


# ---- int4 weight format helpers (quantize.cpp:42-53 layout) ----------------
def pack_w4(q: np.ndarray) -> np.ndarray:
    """q (K, N) ints in [0, 15] -> (K/8, N) uint32, row k in bits 4*(k%8)."""
    K, N = q.shape
    q = q.astype(np.uint32).reshape(K // 8, 8, N)
    out = np.zeros((K // 8, N), dtype=np.uint32)
    for i in range(8):
        out |= q[:, i, :] << np.uint32(4 * i)
    return out


def pack_zeros(z: np.ndarray) -> np.ndarray:
    """z (G, N) ints in [0, 15] -> (G, N/8) uint32, column n in bits 4*(n%8)."""
    G, N = z.shape
    z = z.astype(np.uint32).reshape(G, N // 8, 8)
    out = np.zeros((G, N // 8), dtype=np.uint32)
    for i in range(8):
        out |= z[:, :, i] << np.uint32(4 * i)
    return out


def log_uniform_dims(count: int, lo: int = 64, hi: int = 4096, step: int = 16, seed=SEED, stream=S_BATCH_DIMS):
    """d = step * round(exp(U[ln lo, ln hi)) / step), clamped to [lo, hi]."""
    u = unit(count, seed, stream)
    v = np.exp(math.log(lo) + (math.log(hi) - math.log(lo)) * u)
    d = (np.rint(v / step) * step).astype(np.int64)
    return np.clip(d, lo, hi)


def w4a16_problem(M: int, N: int = 11008, K: int = 4096, group: int = 128, seed: int = SEED):
    """The w4a16 configuration: A ~ N(0,1) fp16; uniform nibbles; scales
    U[0.005, 0.02] fp16; zero-points U{0..15}."""
    A = normal_f64(M * K, seed, S_W4_A).astype(np.float32).astype(np.float16).reshape(M, K)
    q = int_uniform(K * N, 0, 15, seed, S_W4_Q).reshape(K, N)
    scales = uniform_f16((K // group) * N, 0.005, 0.02, seed, S_W4_S).reshape(K // group, N)
    z = int_uniform((K // group) * N, 0, 15, seed, S_W4_Z).reshape(K // group, N)
    return A, q, scales, z
