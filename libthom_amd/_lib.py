"""ctypes binding of lib/liblthip.so (the C ABI declared in include/lthip.h).

Loading fails loudly when the library has not been built: there is no CPU
fallback anywhere in this package.
"""
from __future__ import annotations

import ctypes
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "liblthip.so")
HEADER_PATH = os.path.join(os.path.dirname(HERE), "include", "lthip.h")

c_int, c_size, c_u64, c_u16, c_dbl, c_vp = (ctypes.c_int, ctypes.c_size_t, ctypes.c_uint64,
                                             ctypes.c_uint16, ctypes.c_double, ctypes.c_void_p)

STATUS = {0: "ok", 1: "invalid argument", 2: "unsupported", 3: "workspace", 4: "hip error"}


class LthipError(RuntimeError):
    def __init__(self, func: str, status: int, message: str):
        super().__init__(f"{func}: status {status} ({STATUS.get(status, '?')}): {message}")
        self.status = status


class HgemmConfig(ctypes.Structure):
    _fields_ = [("bm", c_int), ("bn", c_int), ("bk", c_int), ("split_k", c_int), ("stages", c_int)]

    def key(self):
        return (self.bm, self.bn, self.bk, self.split_k, self.stages)

    def __repr__(self):
        return f"HgemmConfig(bm={self.bm}, bn={self.bn}, bk={self.bk}, split_k={self.split_k}, stages={self.stages})"


class HgemmProblem(ctypes.Structure):
    _fields_ = [("A", c_vp), ("B", c_vp), ("C", c_vp), ("M", c_int), ("N", c_int), ("K", c_int),
                ("lda", c_int), ("ldb", c_int), ("ldc", c_int), ("split_k", c_int)]


class W4A16Config(ctypes.Structure):
    _fields_ = [("kind", c_int), ("tile_n", c_int), ("tile_m", c_int), ("split_k", c_int)]

    def key(self):
        return (self.kind, self.tile_n, self.tile_m, self.split_k)

    def __repr__(self):
        return f"W4A16Config(kind={self.kind}, tile_n={self.tile_n}, tile_m={self.tile_m}, split_k={self.split_k})"


class W4A16Problem(ctypes.Structure):
    _fields_ = [("A", c_vp), ("W_packed", c_vp), ("scales", c_vp), ("qzeros", c_vp), ("C", c_vp), ("M", c_int),
                ("N", c_int), ("K", c_int), ("group", c_int)]


# name -> (restype, argtypes)
_SIGS = {
    "lthip_status_string": (ctypes.c_char_p, [c_int]),
    "lthip_last_error": (ctypes.c_char_p, []),
    "lthip_version": (ctypes.c_char_p, []),
    "lthip_workspace_clear": (c_int, [c_vp, c_size, c_vp]),
    "lthip_fill_uniform_f32": (c_int, [c_vp, c_size, c_dbl, c_dbl, c_u64, c_u64, c_vp]),
    "lthip_fill_uniform_f16": (c_int, [c_vp, c_size, c_dbl, c_dbl, c_u64, c_u64, c_vp]),
    "lthip_fill_uniform_f64": (c_int, [c_vp, c_size, c_dbl, c_dbl, c_u64, c_u64, c_vp]),
    "lthip_copy_f32": (c_int, [c_vp, c_vp, c_size, c_vp]),
    "lthip_probe_mfma_flops": (ctypes.c_double, [c_int, c_int, c_int]),
    "lthip_probe_mfma": (c_int, [c_int, c_int, c_int, c_vp, c_vp]),
    "lthip_probe_read_f32": (c_int, [c_vp, c_size, c_vp, c_int, c_vp]),
    "lthip_reduce_sum_f32_workspace_size": (c_size, [c_size]),
    "lthip_reduce_sum_f32": (c_int, [c_vp, c_size, c_vp, c_vp, c_size, c_vp]),
    "lthip_reduce_sum_f32_noalloc": (c_int, [c_vp, c_size, c_vp, c_vp]),
    "lthip_sdot_f32_workspace_size": (c_size, [c_size]),
    "lthip_sdot_f32": (c_int, [c_size, c_vp, c_vp, c_vp, c_vp, c_size, c_vp]),
    "lthip_sgemv_f32_workspace_size": (c_size, [c_int, c_int, c_int]),
    "lthip_sgemv_f32": (c_int, [c_int, c_int, c_int, c_vp, c_vp, c_int, c_vp, c_int, c_vp, c_vp, c_int,
                                c_vp, c_size, c_vp]),
    "lthip_dgemm_f64": (c_int, [c_vp, c_vp, c_vp, c_size, c_size, c_size, c_dbl, c_dbl, c_vp]),
    "lthip_ozaki_max_slices": (c_int, [c_size]),
    "lthip_dgemm_ozaki_workspace_size": (c_size, [c_size, c_size, c_size, c_int]),
    "lthip_dgemm_ozaki": (c_int, [c_vp, c_vp, c_vp, c_size, c_size, c_size, c_dbl, c_dbl, c_int, c_vp, c_size,
                                  c_vp]),
    "lthip_ozaki_split_rows": (c_int, [c_vp, c_size, c_size, c_size, c_int, c_vp, c_vp, c_vp]),
    "lthip_dgemm_compensated": (c_int, [c_vp, c_vp, c_vp, c_vp, c_size, c_size, c_size, c_vp]),
    "lthip_hgemm_config_count": (c_int, []),
    "lthip_hgemm_config_get": (c_int, [c_int, ctypes.POINTER(HgemmConfig)]),
    "lthip_hgemm_default_config": (c_int, [c_size, c_size, c_size, ctypes.POINTER(HgemmConfig)]),
    "lthip_hgemm_workspace_size": (c_size, [c_size, c_size, c_size, ctypes.POINTER(HgemmConfig)]),
    "lthip_hgemm_f16": (c_int, [c_vp, c_vp, c_vp, c_size, c_size, c_size, c_u16, c_u16,
                                ctypes.POINTER(HgemmConfig), c_vp, c_size, c_vp]),
    "lthip_hgemm_grouped_workspace_size": (c_size, [c_vp, c_int, ctypes.POINTER(HgemmConfig)]),
    "lthip_hgemm_grouped_plan_create": (c_int, [c_vp, c_int, ctypes.POINTER(HgemmConfig), c_u16, c_u16, c_vp,
                                                c_size, c_vp, ctypes.POINTER(c_vp)]),
    "lthip_hgemm_grouped_run": (c_int, [c_vp, c_vp]),
    "lthip_hgemm_grouped_capture": (c_int, [c_vp, c_vp]),
    "lthip_hgemm_grouped_graph_launch": (c_int, [c_vp, c_vp]),
    "lthip_hgemm_grouped_plan_info": (c_int, [c_vp, ctypes.POINTER(c_int), ctypes.POINTER(c_int)]),
    "lthip_hgemm_grouped_plan_destroy": (c_int, [c_vp]),
    "lthip_w4a16_grouped_workspace_size": (c_size, [c_vp, c_int, ctypes.POINTER(W4A16Config)]),
    "lthip_w4a16_grouped_plan_create": (c_int, [c_vp, c_int, ctypes.POINTER(W4A16Config), c_vp, c_size, c_vp,
                                                ctypes.POINTER(c_vp)]),
    "lthip_w4a16_grouped_run": (c_int, [c_vp, c_vp]),
    "lthip_w4a16_grouped_capture": (c_int, [c_vp, c_vp]),
    "lthip_w4a16_grouped_graph_launch": (c_int, [c_vp, c_vp]),
    "lthip_w4a16_grouped_plan_info": (c_int, [c_vp, ctypes.POINTER(c_int), ctypes.POINTER(c_int)]),
    "lthip_w4a16_grouped_plan_destroy": (c_int, [c_vp]),
    "lthip_mlp_forward": (c_int, [c_vp] * 7 + [c_int] * 4 + [c_vp]),
    "lthip_mlp_workspace_size": (c_size, [c_int, c_int, c_int]),
    "lthip_mlp_loss_grad": (c_int, [c_vp] * 7 + [c_int] * 4 + [c_vp] * 5 + [c_vp, c_size, c_vp]),
    "lthip_adam_step": (c_int, [c_vp, c_vp, c_vp, c_vp, c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                ctypes.c_float, c_int, c_vp]),
    "lthip_w4a16_workspace_size": (c_size, [c_size, c_size, c_size, ctypes.POINTER(W4A16Config)]),
    "lthip_w4a16_gemm": (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_size, c_size, c_size, c_int,
                                 ctypes.POINTER(W4A16Config), c_vp, c_size, c_vp]),
    "lthip_w4_dequant": (c_int, [c_vp, c_vp, c_vp, c_vp, c_size, c_size, c_int, c_vp]),
    "lthip_w4_quantize": (c_int, [c_vp, c_vp, c_vp, c_vp, c_size, c_size, c_int, c_int, c_vp]),
}

_lock = threading.Lock()
_lib = None


def available() -> bool:
    return os.path.exists(LIB_PATH)


def lib() -> ctypes.CDLL:
    """The loaded library; raises if lib/liblthip.so was not built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(f"{LIB_PATH} is not built; run `make -C libthom_amd` "
                                   "(or __graft_entry__.build()) first")
            handle = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in _SIGS.items():
                fn = getattr(handle, name, None)
                if fn is None:
                    continue
                fn.restype = res
                fn.argtypes = args
            _lib = handle
    return _lib


def check(name: str, status: int) -> None:
    if status != 0:
        msg = lib().lthip_last_error()
        raise LthipError(name, status, msg.decode() if msg else "")


def call(name: str, *args) -> None:
    check(name, getattr(lib(), name)(*args))
