"""The C-ABI library loads on a machine without a GPU, exports every entry
point include/lthip.h declares, and rejects bad arguments before touching
the HIP runtime."""
import ctypes
import re

import pytest

import libthom_amd
from libthom_amd import _lib


def header_functions():
    text = open(_lib.HEADER_PATH).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lthip_[a-z0-9_]+)\s*\(", text)))


def test_exports_every_declared_symbol():
    lib = libthom_amd.lib()
    names = header_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding knows each signature
    assert not [n for n in names if n not in _lib._SIGS]


def test_status_strings_and_version():
    lib = libthom_amd.lib()
    assert lib.lthip_status_string(0) == b"ok"
    assert lib.lthip_status_string(3).startswith(b"workspace")
    assert libthom_amd.version().endswith("gfx950")


def test_invalid_arguments_fail_without_gpu():
    lib = libthom_amd.lib()
    P = ctypes.c_void_p
    fake = P(0x1000)
    # sgemv: incx must be positive
    assert lib.lthip_sgemv_f32(1, 4, 4, fake, fake, 4, fake, 0, fake, fake, 1, None, 0, None) == 1
    # sgemv: lda < n for row-major
    assert lib.lthip_sgemv_f32(1, 4, 8, fake, fake, 4, fake, 1, fake, fake, 1, None, 0, None) == 1
    # hgemm: K not a multiple of 8
    assert lib.lthip_hgemm_f16(fake, fake, fake, 16, 16, 12, 0x3C00, 0, None, None, 0, None) == 1
    # w4a16: K not a multiple of the group
    assert lib.lthip_w4a16_gemm(fake, fake, fake, None, fake, 1, 64, 200, 128, None, None, 0, None) == 1
    # reduction: missing workspace
    assert lib.lthip_reduce_sum_f32(fake, 16, fake, None, 0, None) == 3
    assert b"workspace" in lib.lthip_last_error()
    # zero-size problems are no-ops
    assert lib.lthip_dgemm_f64(None, None, None, 0, 5, 5, 1.0, 0.0, None) == 0


def test_workspace_queries():
    lib = libthom_amd.lib()
    assert lib.lthip_reduce_sum_f32_workspace_size(10 ** 9) >= 256
    assert lib.lthip_sgemv_f32_workspace_size(1, 100, 100) == 0
    # column-major: 16 row blocks x 12 column slices (about 192 blocks)
    assert lib.lthip_sgemv_f32_workspace_size(0, 16384, 8192) == 12 * 16384 * 4
    # few columns cap the slices at n / 64; n = 0 still needs one slice of partials
    assert lib.lthip_sgemv_f32_workspace_size(0, 1024, 1024) == 16 * 1024 * 4
    assert lib.lthip_sgemv_f32_workspace_size(0, 5, 0) == 5 * 4
    cfg = _lib.HgemmConfig(128, 128, 64, 4, 2)
    assert lib.lthip_hgemm_workspace_size(512, 512, 4096, ctypes.byref(cfg)) >= 4 * 512 * 512 * 4
    n = lib.lthip_hgemm_config_count()
    assert n >= 4
    c = _lib.HgemmConfig()
    assert lib.lthip_hgemm_config_get(0, ctypes.byref(c)) == 0 and c.bm > 0
    assert lib.lthip_hgemm_config_get(n, ctypes.byref(c)) == 1


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "nope.so"))
    with pytest.raises(RuntimeError, match="not built"):
        _lib.lib()


def test_hgemm_default_config_rule():
    # without a GPU the library assumes 256 CUs (MI355X): the largest ring
    # tile that gives at least 144 CUs a tile, no split-K; 64 x 64 with
    # split-K below 512 rows or columns
    lib = libthom_amd.lib()
    want = {(8192, 8192, 8192): (256, 256, 32, 1, 4), (4096, 4096, 4096): (256, 256, 32, 1, 4),
            (3072, 3072, 3072): (256, 256, 32, 1, 4), (2560, 2560, 2560): (256, 128, 64, 1, 3),
            (2048, 3072, 2048): (128, 256, 64, 1, 3), (2048, 2048, 2048): (128, 128, 64, 1, 3),
            (1024, 1024, 1024): (128, 128, 64, 1, 3), (512, 512, 512): (128, 128, 64, 1, 3),
            (2560, 8192, 1024): (256, 256, 32, 1, 4), (128, 4096, 4096): (64, 64, 64, 2, 2)}
    for (M, N, K), cfg in want.items():
        c = _lib.HgemmConfig()
        assert lib.lthip_hgemm_default_config(M, N, K, ctypes.byref(c)) == 0
        assert c.key() == cfg, ((M, N, K), c)
