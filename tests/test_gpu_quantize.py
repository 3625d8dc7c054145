"""GPU parity of the int4 quantizer (quantize.cpp:8-53 and its asymmetric
min/max variant): bit-identical packed weights, scales and zero-points against
the oracle's host quantizer, on random and edge-case weights."""
import numpy as np
import pytest
import torch

from libthom_amd import routines as R
from libthom_amd._lib import LthipError

pytestmark = pytest.mark.gpu


def gpu_quantize(W, group, asym, dev):
    packed, s, qz = R.w4_quantize(torch.from_numpy(W).to(dev), group, asym)
    return packed.cpu().numpy().view(np.uint32), s.cpu().numpy(), None if qz is None else qz.cpu().numpy().view(np.uint32)


def check(orc, W, group, asym, dev):
    packed, s, qz = gpu_quantize(W, group, asym, dev)
    if asym:
        q, z, s_ref = orc.w4_quantize_asym(W, group)
        assert np.array_equal(qz, orc.w4_pack_zeros(z))
    else:
        q, s_ref = orc.w4_quantize_sym(W, group)
        assert qz is None
    assert np.array_equal(s.view(np.uint16), s_ref.view(np.uint16))
    assert np.array_equal(packed, orc.w4_pack(q))


@pytest.mark.parametrize("asym", [False, True])
@pytest.mark.parametrize("group", [8, 32, 64, 128, 256])
@pytest.mark.parametrize("N", [520, 64, 8])
def test_quantize_random(dev, orc, asym, group, N):
    K = 1024
    rng = np.random.default_rng(group + N + asym)
    W = (rng.standard_normal((K, N)) * rng.uniform(0.01, 3.0, (1, N))).astype(np.float16)
    check(orc, W, group, asym, dev)


@pytest.mark.parametrize("asym", [False, True])
def test_quantize_edges(dev, orc, asym):
    K, N, group = 512, 72, 64
    rng = np.random.default_rng(40)
    W = rng.uniform(-1, 1, (K, N)).astype(np.float16)
    W[:64, 0] = 0                                               # all-zero group
    W[:, 1] = np.abs(W[:, 1])                                   # all positive
    W[:, 2] = -np.abs(W[:, 2])                                  # all negative
    W[:, 3] = (W[:, 3].astype(np.float32) * 1e-6).astype(np.float16)   # subnormal scales
    W[:, 4] = np.float16(65504.0) * np.sign(W[:, 4])            # the largest finite fp16
    W[64:128, 5] = np.float16(7.5)                              # ties after scaling
    W[64:72, 5] = np.array([7.5, 0.5, 1.5, 2.5, -0.5, -1.5, -2.5, -7.5], np.float16)
    W[:, 6] = np.float16(6e-7)                                  # scales round to the smallest subnormal
    check(orc, W, group, asym, dev)


@pytest.mark.parametrize("asym", [False, True])
def test_quantize_ties_every_column(dev, orc, asym):
    # group 128 (the register-resident kernel): every column holds +-7.5 f and
    # the half-integers (k + 0.5) f, so max -> scale f exactly and W / scale
    # lands on exact ties (round half to even) in every column, with scales
    # that are not powers of two
    K, N, group = 256, 136, 128
    f = np.array([0.25, 1.0, 1.5, 3.0, 0.75, 1.25, 0.375, 2.5], np.float32)[np.arange(N) % 8]
    k = np.arange(K) % 16 - 8 + 0.5                               # -7.5 .. 7.5
    W = (k[:, None] * f[None, :]).astype(np.float16)
    assert np.array_equal(W.astype(np.float32), (k[:, None] * f[None, :]).astype(np.float32))
    check(orc, W, group, asym, dev)


def test_quantize_large_group(dev, orc):
    # group 1024: the staged tile needs more than 64 KiB of LDS
    W = np.random.default_rng(41).standard_normal((2048, 64)).astype(np.float16)
    for asym in (False, True):
        check(orc, W, 1024, asym, dev)


def test_quantize_then_w4a16(dev, orc):
    # the GPU producer feeding the GEMM: same result as the oracle pipeline
    M, K, N, group = 16, 1024, 512, 128
    rng = np.random.default_rng(42)
    W = (rng.standard_normal((K, N)) * 0.05).astype(np.float16)
    A = rng.standard_normal((M, K)).astype(np.float16)
    At = torch.from_numpy(A).to(dev)
    for asym in (False, True):
        packed, s, qz = R.w4_quantize(torch.from_numpy(W).to(dev), group, asym)
        C = R.w4a16(At, packed, s, qz, group).cpu().numpy().astype(np.float32)
        Wd = orc.w4_dequant(packed.cpu().numpy().view(np.uint32), s.cpu().numpy(),
                            None if qz is None else qz.cpu().numpy().view(np.uint32), group)
        ref = orc.w4a16_rows(A, Wd).astype(np.float32)
        rms = float(np.sqrt(np.mean(ref.astype(np.float64) ** 2)))
        assert np.all(np.abs(C - ref) <= 1e-2 * np.abs(ref) + 1e-3 * rms), asym
        # and the quantisation error itself is about half a scale step (the
        # symmetric extremes clip at 7 s against 7.5 s with a scale rounded
        # down by up to half an fp16 ulp)
        sc = np.repeat(s.cpu().numpy().astype(np.float32), group, axis=0)
        assert np.all(np.abs(Wd.astype(np.float32) - W.astype(np.float32)) <= sc * 0.51 +
                      np.abs(Wd.astype(np.float32)) * 2 ** -11 + 1e-7), asym


def test_quantize_rejects(dev):
    W = torch.zeros(64, 24, dtype=torch.float16, device=dev)
    for group, asym in ((12, False), (2048, True)):
        with pytest.raises(LthipError):
            R.w4_quantize(W, group, asym)
    with pytest.raises(LthipError):
        R.w4_quantize(torch.zeros(64, 12, dtype=torch.float16, device=dev), 32, False)
