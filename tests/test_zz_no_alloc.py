"""The allocation-free reduction (lthip_reduce_sum_f32_noalloc) against the
workspace entry point lthip_reduce_sum_f32 and the oracle, through the C ABI.

Two calls in a row share the library's per-GPU scratch array, the second on a
smaller input: each must equal the workspace entry point bit for bit (same
kernel, same grid) and the oracle's FP64 sum within 1e-5.
"""
import ctypes

import numpy as np
import pytest
import torch

from libthom_amd import inputs
from libthom_amd._lib import lib

pytestmark = pytest.mark.gpu


def _noalloc():
    f = lib().lthip_reduce_sum_f32_noalloc
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p]
    return f


def _with_workspace(x, out, stream):
    L = lib()
    need = L.lthip_reduce_sum_f32_workspace_size(x.numel())
    ws = torch.zeros(need, dtype=torch.uint8, device=x.device)
    st = L.lthip_reduce_sum_f32(x.data_ptr(), x.numel(), out.data_ptr(), ws.data_ptr(), ws.numel(), stream)
    assert st == 0, L.lthip_last_error()


def test_noalloc_matches_workspace_entry_point(dev, orc):
    f = _noalloc()
    stream = torch.cuda.current_stream(dev).cuda_stream
    for n, seed in ((5_000_003, 11), (1_000_000, 12)):   # the second call smaller
        x = inputs.uniform_f32(n, 0.0, 1.0, seed, inputs.S_REDUCE)
        xt = torch.from_numpy(x).to(dev)
        got = torch.full((1,), -1.0, dtype=torch.float32, device=dev)
        want = torch.full((1,), -2.0, dtype=torch.float32, device=dev)
        print(f"noalloc reduction n={n}", flush=True)
        st = f(xt.data_ptr(), n, got.data_ptr(), stream)
        assert st == 0, lib().lthip_last_error()
        _with_workspace(xt, want, stream)
        torch.cuda.synchronize()
        g, w = got.item(), want.item()
        assert np.float32(g).tobytes() == np.float32(w).tobytes(), (n, g, w)
        ref = orc.sum_f64(x)
        assert abs(g - ref) <= 1e-5 * ref, (n, g, ref)


def test_noalloc_edge_sizes(dev):
    # empty, one element, a ragged float4 tail
    f = _noalloc()
    stream = torch.cuda.current_stream(dev).cuda_stream
    for n in (0, 1, 7, 4099):
        xt = torch.ones(max(n, 1), dtype=torch.float32, device=dev)
        got = torch.full((1,), -1.0, dtype=torch.float32, device=dev)
        st = f(xt.data_ptr(), n, got.data_ptr(), stream)
        assert st == 0, lib().lthip_last_error()
        torch.cuda.synchronize()
        assert got.item() == float(n), (n, got.item())
