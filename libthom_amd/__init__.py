"""libthom for MI355X (gfx950): hand-written HIP kernels behind a C ABI
(include/lthip.h, lib/liblthip.so) plus the libthom CLI (bin/libthom).

This package is the Python binding used by tests and bench.py; it has no CPU
fallback: every routine raises if the HIP library is not built.
"""
from ._lib import LIB_PATH, HgemmConfig, LthipError, W4A16Config, available, lib  # noqa: F401

__version__ = "0.1.0"


def version() -> str:
    return lib().lthip_version().decode()
