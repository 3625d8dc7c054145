"""GPU parity: dgemm (FP64 MFMA) and hgemm (FP16 MFMA, FP32 accumulate).

dgemm: exact on small-integer inputs; on N(0,1) inputs within 1e-12 of the
oracle normwise-per-element (|C - C_ref| <= 1e-12 (|A||B| + |beta C0|)); at the
4096^3 configuration, sampled entries against the compensated Dot2 oracle.
hgemm: exact on small integers; otherwise the reference's criterion
(hgemm.cu:165-168: fail iff rel > 1e-2 with denominator max(|ref|, 1e-3) and
abs > 1e-3), with rel scaled by (|A||B|)_ij for signed inputs.
"""
import numpy as np
import pytest
import torch

from libthom_amd import inputs, routines as R

pytestmark = pytest.mark.gpu


def t(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


@pytest.mark.parametrize("M,N,K", [(128, 128, 16), (256, 384, 512), (1, 1, 1), (37, 53, 71), (130, 129, 33),
                                   (200, 17, 1000), (64, 300, 0),
                                   (2048, 2304, 64), (2560, 2048, 16)])   # 256 x 128 tiles (>= 144 of them)
def test_dgemm_small_integers_exact(dev, orc, M, N, K):
    rng = np.random.default_rng(M + N + K)
    A = rng.integers(-2, 3, (M, K)).astype(np.float64)
    B = rng.integers(-2, 3, (K, N)).astype(np.float64)
    C0 = rng.integers(-4, 5, (M, N)).astype(np.float64)
    got = R.dgemm(t(A, dev), t(B, dev), t(C0, dev), 2.0, -1.0).cpu().numpy()
    ref = orc.dgemm(A, B, C0, 2.0, -1.0)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("M,N,K", [(512, 512, 512), (300, 200, 1000), (1024, 256, 4096)])
def test_dgemm_random(dev, orc, M, N, K):
    rng = np.random.default_rng(K)
    A = rng.standard_normal((M, K))
    B = rng.standard_normal((K, N))
    got = R.dgemm(t(A, dev), t(B, dev)).cpu().numpy()
    ref = orc.dgemm(A, B)
    scale = np.abs(A) @ np.abs(B)
    assert np.max(np.abs(got - ref) / scale) < 1e-14


def test_dgemm_config_sampled_vs_dot2(dev, orc):
    n = 4096
    A = inputs.normal_f64(n * n, 1234, inputs.S_DGEMM_A).reshape(n, n)
    B = inputs.normal_f64(n * n, 1234, inputs.S_DGEMM_B).reshape(n, n)
    C = R.dgemm(t(A, dev), t(B, dev)).cpu().numpy()
    rng = np.random.default_rng(0)
    rows = rng.integers(0, n, 2000).astype(np.int32)
    cols = rng.integers(0, n, 2000).astype(np.int32)
    hi, lo = orc.dgemm_dot2(A, B, rows, cols)
    scale = np.einsum("ij,ji->i", np.abs(A[rows]), np.abs(B[:, cols]))
    err = np.abs(C[rows, cols] - (hi + lo)) / scale
    assert err.max() < 1e-14
    # row/column power-of-two scaling commutes exactly with the product
    u = inputs.int_uniform(n, -20, 20, 1234, inputs.S_DGEMM_RS)
    As = A * (2.0 ** u)[:, None]
    Cs = R.dgemm(t(As, dev), t(B, dev)).cpu().numpy()
    assert np.array_equal(Cs, C * (2.0 ** u)[:, None])
    # input set (b): rows of A and columns of B scaled, against the CPU Dot2 oracle
    v = inputs.int_uniform(n, -20, 20, 1234, inputs.S_DGEMM_CS)
    Bs = B * (2.0 ** v)[None, :]
    Cb = R.dgemm(t(As, dev), t(Bs, dev)).cpu().numpy()
    assert np.array_equal(Cb, C * (2.0 ** u)[:, None] * (2.0 ** v)[None, :])
    hi, lo = orc.dgemm_dot2(As, Bs, rows, cols)
    scale_b = np.einsum("ij,ji->i", np.abs(As[rows]), np.abs(Bs[:, cols]))
    assert (np.abs(Cb[rows, cols] - (hi + lo)) / scale_b).max() < 1e-14


def hgemm_check(got, ref, A, B):
    got = got.astype(np.float32)
    ref = ref.astype(np.float32)
    scale = np.abs(A.astype(np.float32)) @ np.abs(B.astype(np.float32))
    diff = np.abs(got - ref)
    rel = diff / np.maximum(np.maximum(np.abs(ref), 1e-3), 1e-2 * scale)
    bad = (rel > 1e-2) & (diff > 1e-3)
    assert not bad.any(), f"{bad.sum()} bad of {bad.size}; max rel {rel.max()}"


def hgemm_cfgs():
    """Every compiled tile: the 2-stage register-staged ones and the DMA-ring ones."""
    from libthom_amd._lib import HgemmConfig
    tiles = [(64, 64, 64, 2), (64, 128, 64, 2), (128, 64, 64, 2), (128, 128, 32, 2), (128, 128, 64, 2),
             (128, 256, 64, 2), (256, 128, 64, 2), (256, 256, 64, 2),
             (256, 256, 32, 4), (256, 256, 32, 3), (256, 256, 32, 5), (256, 128, 64, 3), (128, 256, 64, 3),
             (256, 128, 32, 4), (128, 256, 32, 4), (128, 128, 64, 3), (128, 128, 32, 4)]
    return [HgemmConfig(bm, bn, bk, 1, st) for bm, bn, bk, st in tiles]


def test_hgemm_cfgs_cover_library():
    got = sorted((c.bm, c.bn, c.bk, c.stages) for c in R.hgemm_configs())
    assert got == sorted((c.bm, c.bn, c.bk, c.stages) for c in hgemm_cfgs())


@pytest.mark.parametrize("ci", range(17))
@pytest.mark.parametrize("M,N,K", [(256, 256, 256), (96, 200, 72), (520, 264, 1032), (96, 200, 64), (33, 8, 32),
                                   (1000, 1032, 1024), (40, 136, 8), (300, 264, 1000)])
def test_hgemm_tiles_small_integers_exact(dev, orc, ci, M, N, K):
    cfg = hgemm_cfgs()[ci]
    rng = np.random.default_rng(ci * 100 + M)
    A = rng.integers(-2, 3, (M, K)).astype(np.float16)
    B = rng.integers(-2, 3, (K, N)).astype(np.float16)
    got = R.hgemm(t(A, dev), t(B, dev), config=cfg).cpu().numpy()
    assert np.array_equal(got, orc.hgemm(A, B)), cfg


@pytest.mark.parametrize("tile", [(128, 128, 64, 2), (256, 256, 32, 4), (128, 128, 32, 4), (256, 128, 64, 3)])
@pytest.mark.parametrize("split", [1, 2, 3, 8])
@pytest.mark.parametrize("M,N,K", [(512, 512, 512), (64, 128, 4096), (1024, 1024, 1024)])
def test_hgemm_random_split_k(dev, orc, tile, split, M, N, K):
    from libthom_amd._lib import HgemmConfig
    rng = np.random.default_rng(split + M)
    A = rng.uniform(-1, 1, (M, K)).astype(np.float16)
    B = rng.uniform(-1, 1, (K, N)).astype(np.float16)
    C0 = rng.uniform(-1, 1, (M, N)).astype(np.float16)
    bm, bn, bk, st = tile
    cfg = HgemmConfig(bm, bn, bk, split, st)
    got = R.hgemm(t(A, dev), t(B, dev), t(C0, dev), 1.0, 0.5, config=cfg).cpu().numpy()
    hgemm_check(got, orc.hgemm(A, B, C0, 1.0, 0.5), A, B)


@pytest.mark.parametrize("stages", [2, 4])
def test_hgemm_empty_k_scales_c(dev, orc, stages):
    from libthom_amd._lib import HgemmConfig
    M, N = 520, 512
    A = np.zeros((M, 0), np.float16)
    B = np.zeros((0, N), np.float16)
    C0 = np.random.default_rng(9).uniform(-1, 1, (M, N)).astype(np.float16)
    cfg = HgemmConfig(128, 128, 32, 1, stages)
    got = R.hgemm(t(A, dev), t(B, dev), t(C0, dev), 1.0, 0.5, config=cfg).cpu().numpy()
    assert np.array_equal(got, orc.hgemm(A, B, C0, 1.0, 0.5))


def test_hgemm_default_config_large(dev, orc):
    M = N = K = 2048
    A = inputs.uniform_f16(M * K, -1, 1, 1234, 31).reshape(M, K)
    B = inputs.uniform_f16(K * N, -1, 1, 1234, 32).reshape(K, N)
    got = R.hgemm(t(A, dev), t(B, dev)).cpu().numpy()
    rows = slice(0, 96)
    hgemm_check(got[rows], orc.hgemm(A[rows], B), A[rows], B)
