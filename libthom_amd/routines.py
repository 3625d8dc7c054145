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


def dgemm_ozaki(A: torch.Tensor, B: torch.Tensor, slices: int, C: Optional[torch.Tensor] = None,
                alpha: float = 1.0, beta: float = 0.0) -> torch.Tensor:
    """Split-precision dgemm on FP8 MFMA with `slices` digits per operand."""
    _need_cuda(A, B)
    M, K = A.shape
    K2, N = B.shape
    assert K == K2 and A.dtype == torch.float64 and B.dtype == torch.float64
    if C is None:
        C = torch.zeros(M, N, dtype=torch.float64, device=A.device)
    need = lib().lthip_dgemm_ozaki_workspace_size(M, N, K, slices)
    ws = workspace(need, A.device, "ozaki")
    call("lthip_dgemm_ozaki", A.data_ptr(), B.data_ptr(), C.data_ptr(), M, N, K, float(alpha), float(beta), slices,
         ws.data_ptr(), ws.numel(), _stream(A.device))
    return C


def ozaki_split_rows(X: torch.Tensor, slices: int):
    """(digits[slices][rows][cols] E4M3 bytes, exps[rows]) of a row-major FP64 matrix."""
    _need_cuda(X)
    rows, cols = X.shape
    d = torch.empty(slices, rows, cols, dtype=torch.uint8, device=X.device)
    e = torch.empty(rows, dtype=torch.int32, device=X.device)
    call("lthip_ozaki_split_rows", X.data_ptr(), rows, cols, X.stride(0), slices, d.data_ptr(), e.data_ptr(),
         _stream(X.device))
    return d, e


def dgemm_compensated(A: torch.Tensor, B: torch.Tensor):
    """(hi, lo) with hi + lo ~ A@B to about twice FP64 precision (Dot2)."""
    _need_cuda(A, B)
    M, K = A.shape
    N = B.shape[1]
    hi = torch.empty(M, N, dtype=torch.float64, device=A.device)
    lo = torch.empty(M, N, dtype=torch.float64, device=A.device)
    call("lthip_dgemm_compensated", A.data_ptr(), B.data_ptr(), hi.data_ptr(), lo.data_ptr(), M, N, K,
         _stream(A.device))
    return hi, lo


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


def w4_quantize(W: torch.Tensor, group: int = 128, asymmetric: bool = False):
    """quantize_w4 + pack_w4 (quantize.cpp:8-53) on the GPU: W (K, N) fp16 ->
    (W_packed (K/8, N) int32, scales (K/group, N) fp16, qzeros (K/group, N/8)
    int32 or None).  asymmetric: the min/max variant with zero-points."""
    _need_cuda(W)
    K, N = W.shape
    assert W.dtype == torch.float16 and W.is_contiguous()
    packed = torch.empty(K // 8, N, dtype=torch.int32, device=W.device)
    scales = torch.empty(K // group, N, dtype=torch.float16, device=W.device)
    qzeros = torch.empty(K // group, N // 8, dtype=torch.int32, device=W.device) if asymmetric else None
    call("lthip_w4_quantize", W.data_ptr(), packed.data_ptr(), scales.data_ptr(), _ptr(qzeros), K, N, group,
         1 if asymmetric else 0, _stream(W.device))
    return packed, scales, qzeros


class GroupedHgemm:
    """A plan for many independent hgemms in one persistent launch.

    problems: list of (A, B, C, split_k) fp16 tensors (row-major, C written in
    place); config: one tile shape for the whole batch.  run() launches on the
    current stream; capture(stream) records the launch into a hipGraph on a
    side stream and replay(stream) relaunches it."""

    def __init__(self, problems, config: HgemmConfig, alpha: float = 1.0, beta: float = 0.0):
        from ._lib import HgemmProblem
        self._keep = problems
        n = len(problems)
        arr = (HgemmProblem * n)()
        self.device = problems[0][0].device
        for i, (A, B, C, split) in enumerate(problems):
            M, K = A.shape
            N = B.shape[1]
            arr[i] = HgemmProblem(A.data_ptr(), B.data_ptr(), C.data_ptr(), M, N, K, A.stride(0), B.stride(0),
                                  C.stride(0), int(split))
        self._arr = arr
        self.config = config
        need = lib().lthip_hgemm_grouped_workspace_size(ctypes.byref(arr), n, ctypes.byref(config))
        self.ws = torch.empty(max(need, 256), dtype=torch.uint8, device=self.device)
        call("lthip_workspace_clear", self.ws.data_ptr(), self.ws.numel(), _stream(self.device))
        self._plan = ctypes.c_void_p()
        call("lthip_hgemm_grouped_plan_create", ctypes.byref(arr), n, ctypes.byref(config), _f16_bits(alpha),
             _f16_bits(beta), self.ws.data_ptr(), self.ws.numel(), _stream(self.device), ctypes.byref(self._plan))

    def info(self):
        w, g = ctypes.c_int(), ctypes.c_int()
        call("lthip_hgemm_grouped_plan_info", self._plan, ctypes.byref(w), ctypes.byref(g))
        return {"work_items": w.value, "grid": g.value}

    def run(self, stream: Optional[torch.cuda.Stream] = None):
        s = (stream or torch.cuda.current_stream(self.device)).cuda_stream
        call("lthip_hgemm_grouped_run", self._plan, s)

    def capture(self, stream: torch.cuda.Stream):
        call("lthip_hgemm_grouped_capture", self._plan, stream.cuda_stream)

    def replay(self, stream: torch.cuda.Stream):
        call("lthip_hgemm_grouped_graph_launch", self._plan, stream.cuda_stream)

    def close(self):
        if self._plan:
            lib().lthip_hgemm_grouped_plan_destroy(self._plan)
            self._plan = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batched_problem_dims(count: int = 256, seed: int = 1234):
    """The batched configuration: count (M, N, K), each log-uniform in
    [64, 4096] rounded to a multiple of 16."""
    from . import inputs
    d = inputs.log_uniform_dims(3 * count, seed=seed)
    return [(int(d[3 * i]), int(d[3 * i + 1]), int(d[3 * i + 2])) for i in range(count)]


class GroupedW4A16:
    """A plan for many independent w4a16 problems in one persistent launch.

    problems: list of (A, W_packed, scales, qzeros or None, C, group) tensors
    in the layout of w4a16() (C written in place); config: None (each problem
    takes the 256/128/64-row tile by its M) or a kind-3 W4A16Config with
    tile_n 256 forcing one tile.  run() launches on the current stream;
    capture(stream) records the launch into a hipGraph on a side stream and
    replay(stream) relaunches it."""

    def __init__(self, problems, config: Optional[W4A16Config] = None):
        from ._lib import W4A16Problem
        _need_cuda(*[t for prob in problems for t in prob[:5]])
        self._keep = problems
        n = len(problems)
        arr = (W4A16Problem * n)()
        self.device = problems[0][0].device
        for i, (A, W, S, Z, C, group) in enumerate(problems):
            M, K = A.shape
            N = W.shape[1]
            assert W.shape[0] * 8 == K and C.shape == (M, N) and A.is_contiguous() and C.is_contiguous()
            arr[i] = W4A16Problem(A.data_ptr(), W.data_ptr(), S.data_ptr(), _ptr(Z), C.data_ptr(), M, N, K, group)
        self._arr = arr
        self.config = config
        cfg = ctypes.byref(config) if config is not None else None
        need = lib().lthip_w4a16_grouped_workspace_size(ctypes.byref(arr), n, cfg)
        self.ws = torch.empty(max(need, 256), dtype=torch.uint8, device=self.device)
        call("lthip_workspace_clear", self.ws.data_ptr(), self.ws.numel(), _stream(self.device))
        self._plan = ctypes.c_void_p()
        call("lthip_w4a16_grouped_plan_create", ctypes.byref(arr), n, cfg, self.ws.data_ptr(), self.ws.numel(),
             _stream(self.device), ctypes.byref(self._plan))

    def info(self):
        w, g = ctypes.c_int(), ctypes.c_int()
        call("lthip_w4a16_grouped_plan_info", self._plan, ctypes.byref(w), ctypes.byref(g))
        return {"work_items": w.value, "grid": g.value}

    def run(self, stream: Optional[torch.cuda.Stream] = None):
        s = (stream or torch.cuda.current_stream(self.device)).cuda_stream
        call("lthip_w4a16_grouped_run", self._plan, s)

    def capture(self, stream: torch.cuda.Stream):
        call("lthip_w4a16_grouped_capture", self._plan, stream.cuda_stream)

    def replay(self, stream: torch.cuda.Stream):
        call("lthip_w4a16_grouped_graph_launch", self._plan, stream.cuda_stream)

    def close(self):
        if self._plan:
            lib().lthip_w4a16_grouped_plan_destroy(self._plan)
            self._plan = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def measured_peaks(device=None, read_bytes: int = 2 << 30) -> dict:
    """What this device delivers: dense MFMA TFLOP/s (fp16, fp8, fp64; best
    of a few launch widths, lthip_probe_mfma) and HBM read and copy GB/s
    (lthip_probe_read_f32, lthip_copy_f32 over buffers far larger than the
    caches)."""
    device = torch.device(device or "cuda:0")
    cus = torch.cuda.get_device_properties(device).multi_processor_count
    sink = torch.empty(cus * 16 * 512, dtype=torch.float32, device=device)   # 16 blocks of <= 512 threads per CU
    stream = _stream(device)

    def timed(fn, reps=3):
        fn()
        best = float("inf")
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            e1.synchronize()
            best = min(best, e0.elapsed_time(e1))
        return best

    out = {}
    for kind, name, iters in ((0, "fp16_tflops", 2000), (1, "fp8_tflops", 2000), (2, "fp64_tflops", 2000)):
        best = 0.0
        for mult in (2, 4, 8, 16):
            blocks = cus * mult
            ms = timed(lambda: call("lthip_probe_mfma", kind, blocks, iters, sink.data_ptr(), stream))
            best = max(best, lib().lthip_probe_mfma_flops(kind, blocks, iters) / ms / 1e9)
        out[name] = round(best, 1)
    n = read_bytes // 4
    x = torch.empty(n, dtype=torch.float32, device=device)
    fill_uniform(x, 0.0, 1.0, 1, 99)
    best = 0.0
    # grid-stride widths: the stride between a thread's loads (blocks x 4 KiB)
    # decides how the concurrent requests spread over HBM banks, so the best
    # of several widths is the device's read rate
    for blocks in (cus * 3 // 2, cus * 2, cus * 3, cus * 4, cus * 8):
        ms = timed(lambda: call("lthip_probe_read_f32", x.data_ptr(), n, sink.data_ptr(), blocks, stream))
        best = max(best, n * 4 / ms / 1e6)
    out["hbm_read_gbs"] = round(best, 1)
    y = torch.empty_like(x)
    ms = timed(lambda: copy_f32(y, x))
    out["hbm_copy_gbs"] = round(2 * n * 4 / ms / 1e6, 1)
    del x, y
    return out
