"""numpy-facing wrapper of the CPU oracle (oracle/_build/liboracle.so).

Test infrastructure: used only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SO = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
try:
    _NCPU = len(os.sched_getaffinity(0))
except AttributeError:
    _NCPU = os.cpu_count() or 1
THREADS = max(1, min(16, _NCPU))   # all usable CPUs, at most a GPU box's share of 16

if not os.path.exists(SO):
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True, capture_output=True)
_L = ctypes.CDLL(SO)

P = ctypes.c_void_p
I = ctypes.c_int
Z = ctypes.c_size_t
D = ctypes.c_double
U64 = ctypes.c_uint64

_sigs = {
    "orc_f32_to_f16": (ctypes.c_uint16, [ctypes.c_float]),
    "orc_f16_to_f32": (ctypes.c_float, [ctypes.c_uint16]),
    "orc_int_to_e4m3": (ctypes.c_uint8, [I]),
    "orc_e4m3_to_f32": (ctypes.c_float, [ctypes.c_uint8]),
    "orc_splitmix64": (U64, [ctypes.POINTER(U64)]),
    "orc_rng_u64": (U64, [U64, U64, U64]),
    "orc_fill_uniform_f32": (None, [P, Z, D, D, U64, U64]),
    "orc_fill_uniform_f16": (None, [P, Z, D, D, U64, U64]),
    "orc_fill_uniform_f64": (None, [P, Z, D, D, U64, U64]),
    "orc_fill_normal_f64": (None, [P, Z, U64, U64]),
    "orc_fill_int_u8": (None, [P, Z, I, I, U64, U64]),
    "orc_sum_f64": (D, [P, Z]),
    "orc_dot_f64": (D, [P, P, Z]),
    "orc_sgemv": (None, [I, I, I, ctypes.c_float, P, I, P, I, ctypes.c_float, P, I]),
    "orc_dgemm": (None, [P, P, P, I, I, I, D, D, I]),
    "orc_dgemm_dot2_entries": (None, [P, P, I, I, I, P, P, Z, P, P, I]),
    "orc_ozaki_split_rows": (None, [P, I, I, I, I, P, P]),
    "orc_ozaki_dgemm": (None, [P, P, P, I, I, I, D, D, I, I]),
    "orc_hgemm": (None, [P, P, P, I, I, I, ctypes.c_uint16, ctypes.c_uint16, I]),
    "orc_w4_quantize_sym": (None, [P, P, P, I, I, I]),
    "orc_w4_quantize_asym": (None, [P, P, P, P, I, I, I]),
    "orc_w4_pack": (None, [P, P, I, I]),
    "orc_w4_unpack": (None, [P, P, I, I]),
    "orc_w4_pack_zeros": (None, [P, P, I, I]),
    "orc_w4_dequant": (None, [P, P, P, P, I, I, I]),
    "orc_w4a16_gemm_rows": (None, [P, P, P, I, I, I, I, I, I]),
    "orc_mlp_forward": (D, [P, P, P, P, P, I, I, I, I, P, P]),
    "orc_mlp_loss_grad": (D, [P, P, P, P, P, P, P, I, I, I, I, P, P, P, P]),
}
for _n, (_r, _a) in _sigs.items():
    _f = getattr(_L, _n)
    _f.restype = _r
    _f.argtypes = _a


def _p(a):
    return None if a is None else a.ctypes.data


def f32_to_f16(v: float) -> int:
    return _L.orc_f32_to_f16(v)


def f16_to_f32(h: int) -> float:
    return _L.orc_f16_to_f32(h)


def int_to_e4m3(d: int) -> int:
    return _L.orc_int_to_e4m3(d)


def e4m3_to_f32(b: int) -> float:
    return _L.orc_e4m3_to_f32(b)


def splitmix64_seq(seed: int, n: int):
    s = U64(seed)
    return [_L.orc_splitmix64(ctypes.byref(s)) for _ in range(n)]


def rng_u64(seed, stream, i):
    return _L.orc_rng_u64(seed, stream, i)


def fill_uniform(n, lo, hi, seed, stream, dtype=np.float32):
    out = np.empty(n, dtype=dtype)
    name = {np.float32: "orc_fill_uniform_f32", np.float16: "orc_fill_uniform_f16",
            np.float64: "orc_fill_uniform_f64"}[dtype]
    getattr(_L, name)(_p(out), n, lo, hi, seed, stream)
    return out


def fill_normal(n, seed, stream):
    out = np.empty(n, dtype=np.float64)
    _L.orc_fill_normal_f64(_p(out), n, seed, stream)
    return out


def fill_int_u8(n, lo, hi, seed, stream):
    out = np.empty(n, dtype=np.uint8)
    _L.orc_fill_int_u8(_p(out), n, lo, hi, seed, stream)
    return out


def sum_f64(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    return _L.orc_sum_f64(_p(x), x.size)


def dot_f64(x, y):
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.ascontiguousarray(y, dtype=np.float32)
    return _L.orc_dot_f64(_p(x), _p(y), x.size)


def sgemv(row_major, m, n, alpha, A, lda, x, incx, beta, y, incy):
    A = np.ascontiguousarray(A, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.array(y, dtype=np.float32, copy=True)
    _L.orc_sgemv(int(row_major), m, n, alpha, _p(A), lda, _p(x), incx, beta, _p(y), incy)
    return y


def dgemm(A, B, C=None, alpha=1.0, beta=0.0, threads=THREADS):
    A = np.ascontiguousarray(A, dtype=np.float64)
    B = np.ascontiguousarray(B, dtype=np.float64)
    M, K = A.shape
    N = B.shape[1]
    C = np.zeros((M, N)) if C is None else np.array(C, dtype=np.float64, copy=True)
    _L.orc_dgemm(_p(A), _p(B), _p(C), M, N, K, alpha, beta, threads)
    return C


def dgemm_dot2(A, B, rows, cols, threads=THREADS):
    A = np.ascontiguousarray(A, dtype=np.float64)
    B = np.ascontiguousarray(B, dtype=np.float64)
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    cols = np.ascontiguousarray(cols, dtype=np.int32)
    hi = np.empty(rows.size)
    lo = np.empty(rows.size)
    _L.orc_dgemm_dot2_entries(_p(A), _p(B), A.shape[0], B.shape[1], A.shape[1], _p(rows), _p(cols), rows.size,
                              _p(hi), _p(lo), threads)
    return hi, lo


def ozaki_split_rows(X, slices):
    X = np.ascontiguousarray(X, dtype=np.float64)
    r, c = X.shape
    d = np.empty((slices, r, c), dtype=np.int8)
    e = np.empty(r, dtype=np.int32)
    _L.orc_ozaki_split_rows(_p(X), r, c, c, slices, _p(d), _p(e))
    return d, e


def ozaki_dgemm(A, B, slices, C=None, alpha=1.0, beta=0.0, threads=THREADS):
    A = np.ascontiguousarray(A, dtype=np.float64)
    B = np.ascontiguousarray(B, dtype=np.float64)
    M, K = A.shape
    N = B.shape[1]
    C = np.zeros((M, N)) if C is None else np.array(C, dtype=np.float64, copy=True)
    _L.orc_ozaki_dgemm(_p(A), _p(B), _p(C), M, N, K, alpha, beta, slices, threads)
    return C


def hgemm(A, B, C=None, alpha=1.0, beta=0.0, threads=THREADS):
    """A, B float16 arrays; returns float16 C."""
    A = np.ascontiguousarray(A, dtype=np.float16)
    B = np.ascontiguousarray(B, dtype=np.float16)
    M, K = A.shape
    N = B.shape[1]
    C = np.zeros((M, N), dtype=np.float16) if C is None else np.array(C, dtype=np.float16, copy=True)
    ab = int(np.array([alpha], np.float16).view(np.uint16)[0])
    bb = int(np.array([beta], np.float16).view(np.uint16)[0])
    _L.orc_hgemm(_p(A.view(np.uint16)), _p(B.view(np.uint16)), _p(C.view(np.uint16)), M, N, K, ab, bb, threads)
    return C


def w4_quantize_sym(W, group=128):
    W = np.ascontiguousarray(W, dtype=np.float16)
    K, N = W.shape
    q = np.empty((K, N), dtype=np.uint8)
    s = np.empty((K // group, N), dtype=np.float16)
    _L.orc_w4_quantize_sym(_p(W.view(np.uint16)), _p(q), _p(s.view(np.uint16)), K, N, group)
    return q, s


def w4_quantize_asym(W, group=128):
    W = np.ascontiguousarray(W, dtype=np.float16)
    K, N = W.shape
    q = np.empty((K, N), dtype=np.uint8)
    z = np.empty((K // group, N), dtype=np.uint8)
    s = np.empty((K // group, N), dtype=np.float16)
    _L.orc_w4_quantize_asym(_p(W.view(np.uint16)), _p(q), _p(z), _p(s.view(np.uint16)), K, N, group)
    return q, z, s


def w4_pack(q):
    q = np.ascontiguousarray(q, dtype=np.uint8)
    K, N = q.shape
    out = np.empty((K // 8, N), dtype=np.uint32)
    _L.orc_w4_pack(_p(q), _p(out), K, N)
    return out


def w4_unpack(packed):
    packed = np.ascontiguousarray(packed, dtype=np.uint32)
    K, N = packed.shape[0] * 8, packed.shape[1]
    q = np.empty((K, N), dtype=np.uint8)
    _L.orc_w4_unpack(_p(packed), _p(q), K, N)
    return q


def w4_pack_zeros(z):
    z = np.ascontiguousarray(z, dtype=np.uint8)
    G, N = z.shape
    out = np.empty((G, N // 8), dtype=np.uint32)
    _L.orc_w4_pack_zeros(_p(z), _p(out), G, N)
    return out


def w4_dequant(packed, scales, qzeros, group=128):
    packed = np.ascontiguousarray(packed, dtype=np.uint32)
    scales = np.ascontiguousarray(scales, dtype=np.float16)
    K, N = packed.shape[0] * 8, packed.shape[1]
    W = np.empty((K, N), dtype=np.float16)
    qz = None if qzeros is None else np.ascontiguousarray(qzeros, dtype=np.uint32)
    _L.orc_w4_dequant(_p(packed), _p(scales.view(np.uint16)), _p(qz), _p(W.view(np.uint16)), K, N, group)
    return W


def w4a16_rows(A, Wdq, row0=0, rows=None, threads=THREADS):
    """C rows [row0, row0+rows) of fp16(A @ Wdq) with FP32 k-ascending accumulation."""
    A = np.ascontiguousarray(A, dtype=np.float16)
    Wdq = np.ascontiguousarray(Wdq, dtype=np.float16)
    M, K = A.shape
    N = Wdq.shape[1]
    rows = M - row0 if rows is None else rows
    C = np.zeros((M, N), dtype=np.float16)
    _L.orc_w4a16_gemm_rows(_p(A.view(np.uint16)), _p(Wdq.view(np.uint16)), _p(C.view(np.uint16)), M, N, K, row0,
                           rows, threads)
    return C[row0:row0 + rows]


def mlp_forward(x, W1, b1, W2, b2):
    x, W1, b1, W2, b2 = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, W1, b1, W2, b2))
    B, I_ = x.shape
    H = W1.shape[0]
    O = W2.shape[0]
    hid = np.empty((B, H))
    y = np.empty((B, O))
    _L.orc_mlp_forward(_p(x), _p(W1), _p(b1), _p(W2), _p(b2), B, I_, H, O, _p(hid), _p(y))
    return hid, y


def mlp_loss_grad(x, W1, b1, W2, b2, t, mask):
    x, W1, b1, W2, b2, t = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, W1, b1, W2, b2, t))
    mask = np.ascontiguousarray(mask, dtype=np.uint8)
    B, I_ = x.shape
    H = W1.shape[0]
    O = W2.shape[0]
    g = [np.empty_like(W1), np.empty_like(b1), np.empty_like(W2), np.empty_like(b2)]
    loss = _L.orc_mlp_loss_grad(_p(x), _p(W1), _p(b1), _p(W2), _p(b2), _p(t), _p(mask), B, I_, H, O,
                                *[_p(a) for a in g])
    return loss, g
