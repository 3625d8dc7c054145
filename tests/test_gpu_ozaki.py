"""GPU parity: split-precision dgemm (FP8 MFMA) and the compensated FP64 GEMM.

Both are bit-exact against the oracle: the FP8 level sums are exact integers
and the FP64 recombination follows the oracle's order; Dot2 runs the same
sequence of error-free transformations.  At the 4096^3 configuration the
emulated error against the Dot2 result is bounded by the digit count.
"""
import numpy as np
import pytest
import torch

from libthom_amd import inputs, routines as R

pytestmark = pytest.mark.gpu


def t(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


@pytest.fixture(scope="module")
def e4m3_table(orc):
    return np.array([orc.e4m3_to_f32(b) for b in range(256)], dtype=np.float64)


@pytest.mark.parametrize("S", [2, 5, 8, 10])
def test_split_bit_exact(dev, orc, e4m3_table, S):
    rng = np.random.default_rng(S)
    X = rng.standard_normal((37, 96)) * (2.0 ** rng.integers(-30, 31, (37, 1)))
    X[5] = 0.0
    X[6, 3] = 2.0 ** -1074   # subnormal element
    d, e = R.ozaki_split_rows(t(X, dev), S)
    dg = e4m3_table[d.cpu().numpy()]
    od, oe = orc.ozaki_split_rows(X, S)
    assert np.array_equal(e.cpu().numpy(), oe)
    assert np.array_equal(dg, od.astype(np.float64))


@pytest.mark.parametrize("S", [2, 4, 7, 10])
@pytest.mark.parametrize("M,N,K", [(64, 64, 128), (100, 70, 256), (1, 130, 16), (257, 65, 400),
                                   (256, 128, 256), (512, 256, 512)])
def test_ozaki_bit_exact(dev, orc, S, M, N, K):
    # the last two shapes take the level-GEMM path (M % 256, N % 128, K % 256)
    _bit_exact(dev, orc, S, M, N, K)


@pytest.mark.parametrize("S", [11, 13, 15])
def test_ozaki_level_path_beyond_ten_slices(dev, orc, S):
    _bit_exact(dev, orc, S, 256, 256, 512)


def test_ozaki_slice_limits(dev):
    from libthom_amd._lib import LthipError
    A = torch.zeros(100, 64, dtype=torch.float64, device=dev)
    B = torch.zeros(64, 100, dtype=torch.float64, device=dev)
    with pytest.raises(LthipError):
        R.dgemm_ozaki(A, B, 11)          # only the level-GEMM shapes go past 10


def _bit_exact(dev, orc, S, M, N, K):
    rng = np.random.default_rng(M * S + N)
    A = rng.standard_normal((M, K)) * (2.0 ** rng.integers(-20, 21, (M, 1)))
    B = rng.standard_normal((K, N)) * (2.0 ** rng.integers(-20, 21, (1, N)))
    C0 = rng.standard_normal((M, N))
    got = R.dgemm_ozaki(t(A, dev), t(B, dev), S, t(C0, dev), 1.5, -0.5).cpu().numpy()
    ref = orc.ozaki_dgemm(A, B, S, C0, 1.5, -0.5)
    assert np.array_equal(got, ref)


def test_compensated_bit_exact(dev, orc):
    rng = np.random.default_rng(9)
    M, N, K = 45, 77, 333
    A = rng.standard_normal((M, K)) * (2.0 ** rng.integers(-25, 25, (M, K)))
    B = rng.standard_normal((K, N))
    hi, lo = R.dgemm_compensated(t(A, dev), t(B, dev))
    rows, cols = np.meshgrid(np.arange(M, dtype=np.int32), np.arange(N, dtype=np.int32), indexing="ij")
    oh, ol = orc.dgemm_dot2(A, B, rows.ravel(), cols.ravel())
    assert np.array_equal(hi.cpu().numpy().ravel(), oh)
    assert np.array_equal(lo.cpu().numpy().ravel(), ol)


@pytest.mark.parametrize("scaled", [False, True])
def test_config_error_vs_slices(dev, scaled):
    """4096^3, inputs (a) N(0,1) and (b) rows of A / columns of B scaled by 2^u,
    u in [-20, 20]: the emulated error against Dot2 falls ~5 bits per slice and
    the native FP64 MFMA error stays at the FP64 level."""
    n = 4096
    A = inputs.normal_f64(n * n, 1234, inputs.S_DGEMM_A).reshape(n, n)
    B = inputs.normal_f64(n * n, 1234, inputs.S_DGEMM_B).reshape(n, n)
    if scaled:
        A = A * (2.0 ** inputs.int_uniform(n, -20, 20, 1234, inputs.S_DGEMM_RS))[:, None]
        B = B * (2.0 ** inputs.int_uniform(n, -20, 20, 1234, inputs.S_DGEMM_CS))[None, :]
    At, Bt = t(A, dev), t(B, dev)
    hi, lo = R.dgemm_compensated(At, Bt)
    scale = R.dgemm(At.abs(), Bt.abs())
    native = R.dgemm(At, Bt)
    err_native = ((native - hi - lo).abs() / scale).max().item()
    assert err_native < 1e-14
    errs = {}
    for S in (4, 6, 8, 10):
        C = R.dgemm_ozaki(At, Bt, S)
        errs[S] = ((C - hi - lo).abs() / scale).max().item()
    print("native", err_native, "ozaki", errs)
    assert errs[6] < errs[4] / 100 and errs[8] < errs[6] / 100
    # 4 + 5 (S - 1) bits of each operand relative to its row/column maximum
    assert errs[8] < 2.0 ** -34


@pytest.mark.parametrize("M,N,K", [(100, 70, 256), (256, 128, 256)])
def test_ozaki_nonfinite_inputs(dev, orc, M, N, K):
    # NaN in a row of A, +inf and -inf in columns of B, a 0 * inf entry: the
    # marked rows/columns follow the FP64 product, bit for bit with the oracle
    # (NaN in the same places), on both the all-levels and the level-GEMM path
    rng = np.random.default_rng(M + N)
    A = rng.standard_normal((M, K))
    B = rng.standard_normal((K, N))
    A[3, 5] = np.nan
    B[7, 2] = np.inf
    B[1, 4] = -np.inf
    A[6, 1] = 0.0
    C0 = rng.standard_normal((M, N))
    got = R.dgemm_ozaki(t(A, dev), t(B, dev), 6, t(C0, dev), 1.5, -0.5).cpu().numpy()
    ref = orc.ozaki_dgemm(A, B, 6, C0, 1.5, -0.5)
    assert np.array_equal(got, ref, equal_nan=True)
    assert np.isnan(got[3]).all() and np.isnan(got[6, 4]) and np.isinf(got[0, 2])
    keep = np.ones((M, N), bool)
    keep[3, :] = False
    keep[:, [2, 4]] = False
    assert np.isfinite(got[keep]).all()
