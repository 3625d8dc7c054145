"""Torch-facing calls of the gfx950 routines.

torch supplies device memory and the stream (the current stream of the tensor's
device); every computation runs in liblthip.so.  Shapes follow the reference's
row-major conventions (routines/*/*.cpp).
"""
from __future__ import annotations

import ctypes
from typing import Optional

import torch

from ._lib import HgemmConfig, W4A16Config, call, lib

_workspaces: dict = {}


def _stream(dev: torch.device) -> int:
    return torch.cuda.current_stream(dev).cuda_stream


def _ptr(t: Optional[torch.Tensor]) -> Optional[int]:
    return None if t is None else t.data_ptr()


def workspace(nbytes: int, device: torch.device, name: str) -> torch.Tensor:
    """A zero-initialised workspace of at least nbytes for entry point `name`.

    Each entry point keeps its own: the counters an entry point leaves at zero
    would be overwritten by another entry point's partial sums."""
    key = (name, device.type, device.index if device.index is not None else torch.cuda.current_device())
    ws = _workspaces.get(key)
    if ws is None or ws.numel() < nbytes:
        size = max(nbytes, 1 << 20)
        ws = torch.empty(size, dtype=torch.uint8, device=device)
        call("lthip_workspace_clear", ws.data_ptr(), size, _stream(device))
        _workspaces[key] = ws
    return ws


def _need_cuda(*ts):
    for t in ts:
        if t is not None and not t.is_cuda:
            raise ValueError("tensor must be on a GPU")


def fill_uniform(out: torch.Tensor, lo: float, hi: float, seed: int, stream_id: int = 0) -> torch.Tensor:
    _need_cuda(out)
    name = {torch.float32: "lthip_fill_uniform_f32", torch.float16: "lthip_fill_uniform_f16",
            torch.float64: "lthip_fill_uniform_f64"}[out.dtype]
    call(name, out.data_ptr(), out.numel(), float(lo), float(hi), seed, stream_id, _stream(out.device))
    return out


def copy_f32(dst: torch.Tensor, src: torch.Tensor) -> torch.Tensor:
    _need_cuda(dst, src)
    call("lthip_copy_f32", dst.data_ptr(), src.data_ptr(), src.numel(), _stream(src.device))
    return dst


def reduce_sum(x: torch.Tensor, out: Optional[torch.Tensor] = None) -> torch.Tensor:
    """FP32 sum of x (reduction.cpp:20 ReductionImpl)."""
    _need_cuda(x)
    assert x.dtype == torch.float32 and x.is_contiguous()
    if out is None:
        out = torch.empty(1, dtype=torch.float32, device=x.device)
    need = lib().lthip_reduce_sum_f32_workspace_size(x.numel())
    ws = workspace(need, x.device, "reduce")
    call("lthip_reduce_sum_f32", x.data_ptr(), x.numel(), out.data_ptr(), ws.data_ptr(), ws.numel(),
         _stream(x.device))
    return out


def sdot(x: torch.Tensor, y: torch.Tensor, out: Optional[torch.Tensor] = None) -> torch.Tensor:
    """FP32 dot product (sdot.cpp:32 SdotImpl)."""
    _need_cuda(x, y)
    assert x.dtype == torch.float32 and y.dtype == torch.float32 and x.numel() == y.numel()
    if out is None:
        out = torch.empty(1, dtype=torch.float32, device=x.device)
    need = lib().lthip_sdot_f32_workspace_size(x.numel())
    ws = workspace(need, x.device, "sdot")
    call("lthip_sdot_f32", x.numel(), x.data_ptr(), y.data_ptr(), out.data_ptr(), ws.data_ptr(), ws.numel(),
         _stream(x.device))
    return out


def sgemv(A: torch.Tensor, x: torch.Tensor, y: Optional[torch.Tensor] = None, alpha: float = 1.0,
          beta: float = 0.0, row_major: bool = True, m: Optional[int] = None, n: Optional[int] = None,
          lda: Optional[int] = None, incx: int = 1, incy: int = 1,
          alpha_dev: Optional[torch.Tensor] = None, beta_dev: Optional[torch.Tensor] = None) -> torch.Tensor:
    """y = alpha*op(A) x + beta*y (sgemv.cpp:38 SgemvImpl).

    Row-major: A is (m, lda) storage with n <= lda used columns.  Column-major:
    A is (n, lda) storage (column j at A[j]) with m <= lda used rows."""
    _need_cuda(A, x)
    if row_major:
        m = A.shape[0] if m is None else m
        n = A.shape[1] if n is None else n
        lda = A.stride(0) if lda is None else lda
    else:
        n = A.shape[0] if n is None else n
        m = A.shape[1] if m is None else m
        lda = A.stride(0) if lda is None else lda
    dev = A.device
    if y is None:
        y = torch.zeros(m * incy, dtype=torch.float32, device=dev)
    if alpha_dev is None:
        alpha_dev = torch.tensor([alpha], dtype=torch.float32, device=dev)
    if beta_dev is None:
        beta_dev = torch.tensor([beta], dtype=torch.float32, device=dev)
    need = lib().lthip_sgemv_f32_workspace_size(int(row_major), m, n)
    ws = workspace(need, dev, "sgemv") if need else None
    call("lthip_sgemv_f32", int(row_major), m, n, alpha_dev.data_ptr(), A.data_ptr(), lda, x.data_ptr(), incx,
         beta_dev.data_ptr(), y.data_ptr(), incy, _ptr(ws), 0 if ws is None else ws.numel(), _stream(dev))
    return y


def dgemm(A: torch.Tensor, B: torch.Tensor, C: Optional[torch.Tensor] = None, alpha: float = 1.0,
          beta: float = 0.0) -> torch.Tensor:
    """C = alpha*A@B + beta*C on FP64 MFMA (dgemm.cpp:36 DgemmImpl)."""
    _need_cuda(A, B)
    M, K = A.shape
    K2, N = B.shape
    assert K == K2 and A.dtype == torch.float64 and B.dtype == torch.float64
    if C is None:
        C = torch.zeros(M, N, dtype=torch.float64, device=A.device)
    call("lthip_dgemm_f64", A.data_ptr(), B.data_ptr(), C.data_ptr(), M, N, K, float(alpha), float(beta),
         _stream(A.device))
    return C


def _f16_bits(v: float) -> int:
    import numpy as np
    return int(np.array([v], dtype=np.float16).view(np.uint16)[0])


def hgemm_configs():
    n = lib().lthip_hgemm_config_count()
    out = []
    for i in range(n):
        c = HgemmConfig()
        call("lthip_hgemm_config_get", i, ctypes.byref(c))
        out.append(c)
    return out


def hgemm_default_config(M: int, N: int, K: int) -> HgemmConfig:
    c = HgemmConfig()
    call("lthip_hgemm_default_config", M, N, K, ctypes.byref(c))
    return c


def hgemm(A: torch.Tensor, B: torch.Tensor, C: Optional[torch.Tensor] = None, alpha: float = 1.0,
          beta: float = 0.0, config: Optional[HgemmConfig] = None) -> torch.Tensor:
    """C = fp16(alpha*A@B + beta*C), FP32 accumulate (hgemm.h:6 HgemmFunction)."""
    _need_cuda(A, B)
    M, K = A.shape
    K2, N = B.shape
    assert K == K2 and A.dtype == torch.float16 and B.dtype == torch.float16
    if C is None:
        C = torch.zeros(M, N, dtype=torch.float16, device=A.device)
    cfg = ctypes.byref(config) if config is not None else None
    need = lib().lthip_hgemm_workspace_size(M, N, K, cfg)
    ws = workspace(need, A.device, "hgemm") if need else None
    call("lthip_hgemm_f16", A.data_ptr(), B.data_ptr(), C.data_ptr(), M, N, K, _f16_bits(alpha), _f16_bits(beta),
         cfg, _ptr(ws), 0 if ws is None else ws.numel(), _stream(A.device))
    return C


def w4a16(A: torch.Tensor, W_packed: torch.Tensor, scales: torch.Tensor, qzeros: Optional[torch.Tensor],
          group: int = 128, C: Optional[torch.Tensor] = None, config: Optional[W4A16Config] = None) -> torch.Tensor:
    """C = A @ dequant(W) (w4a16.h:34 W4A16Function, plus zero-points).

    A (M, K) fp16; W_packed (K/8, N) int32; scales (K/group, N) fp16;
    qzeros (K/group, N/8) int32 or None (zero-point 8)."""
    _need_cuda(A, W_packed, scales, qzeros)
    M, K = A.shape
    N = W_packed.shape[1]
    assert W_packed.shape[0] * 8 == K
    if C is None:
        C = torch.empty(M, N, dtype=torch.float16, device=A.device)
    cfg = ctypes.byref(config) if config is not None else None
    need = lib().lthip_w4a16_workspace_size(M, N, K, cfg)
    ws = workspace(need, A.device, "w4a16") if need else None
    call("lthip_w4a16_gemm", A.data_ptr(), W_packed.data_ptr(), scales.data_ptr(), _ptr(qzeros), C.data_ptr(),
         M, N, K, group, cfg, _ptr(ws), 0 if ws is None else ws.numel(), _stream(A.device))
    return C


def w4_dequant(W_packed: torch.Tensor, scales: torch.Tensor, qzeros: Optional[torch.Tensor],
               group: int = 128) -> torch.Tensor:
    _need_cuda(W_packed, scales, qzeros)
    K = W_packed.shape[0] * 8
    N = W_packed.shape[1]
    out = torch.empty(K, N, dtype=torch.float16, device=W_packed.device)
    call("lthip_w4_dequant", W_packed.data_ptr(), scales.data_ptr(), _ptr(qzeros), out.data_ptr(), K, N, group,
         _stream(W_packed.device))
    return out
