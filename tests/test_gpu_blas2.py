"""GPU parity: sgemv (row- and column-major) against the oracle.

Inputs are signed (U[-1,1)), so the per-element criterion is componentwise:
|y - y_ref| <= tol * (|alpha| |A| |x| + |beta| |y0|)_i with tol = 1e-4 (the
configuration's tolerance), plus the max relative error is reported.
"""
import numpy as np
import pytest
import torch

from libthom_amd import inputs, routines as R

pytestmark = pytest.mark.gpu


def check(y, ref, scale, tol=1e-4):
    err = np.abs(y.astype(np.float64) - ref.astype(np.float64))
    bad = np.nonzero(err > tol * scale)[0]
    assert bad.size == 0, f"{bad.size} bad, first {bad[:4]}: got {y[bad[:4]]} want {ref[bad[:4]]}"


def test_sgemv_config_shape(dev, orc):
    m, n = 16384, 8192
    A = torch.empty(m * n, dtype=torch.float32, device=dev)
    R.fill_uniform(A, -1.0, 1.0, 1234, inputs.S_SGEMV_A)
    x = torch.empty(n, dtype=torch.float32, device=dev)
    R.fill_uniform(x, -1.0, 1.0, 1234, inputs.S_SGEMV_X)
    y = R.sgemv(A.view(m, n), x)
    An, xn = A.cpu().numpy(), x.cpu().numpy()
    ref = orc.sgemv(True, m, n, 1.0, An, n, xn, 1, 0.0, np.zeros(m, np.float32), 1)
    scale = np.abs(An.reshape(m, n)) @ np.abs(xn)
    check(y.cpu().numpy(), ref, scale)


@pytest.mark.parametrize("m,n,lda,incx,incy,beta", [
    (1, 1, 1, 1, 1, 0.0), (3, 5, 5, 1, 1, 0.0), (100, 257, 260, 1, 1, 0.5), (513, 1000, 1001, 2, 3, -1.0),
    (64, 20000, 20000, 1, 1, 0.0), (64, 20001, 20004, 3, 1, 2.0), (4000, 4096, 4096, 1, 1, 1.0),
    (5, 0, 8, 1, 1, 0.5), (1030, 777, 1032, 1, 2, 1.0)])
@pytest.mark.parametrize("row_major", [True, False])
def test_sgemv_shapes(dev, orc, m, n, lda, incx, incy, beta, row_major):
    rng = np.random.default_rng(m * 7 + n)
    if not row_major:
        lda = max(lda, m)
    outer = m if row_major else n
    A = rng.uniform(-1, 1, outer * lda).astype(np.float32)
    x = rng.uniform(-1, 1, n * incx).astype(np.float32)
    y0 = rng.uniform(-1, 1, m * incy).astype(np.float32)
    alpha = 1.5
    yt = torch.from_numpy(y0.copy()).to(dev)
    At = torch.from_numpy(A).to(dev).view(outer, lda)
    R.sgemv(At, torch.from_numpy(x).to(dev), yt, alpha, beta, row_major, m=m, n=n, lda=lda, incx=incx, incy=incy)
    ref = orc.sgemv(row_major, m, n, alpha, A, lda, x, incx, beta, y0, incy)
    Am = A.reshape(outer, lda)[:, :n] if row_major else A.reshape(outer, lda)[:, :m].T
    scale = abs(alpha) * (np.abs(Am) @ np.abs(x[::incx][:n])) + abs(beta) * np.abs(y0[::incy][:m]) + 1e-30
    got = yt.cpu().numpy()
    check(got[::incy][:m], ref[::incy][:m], scale)
    # entries between the strided outputs are untouched
    if incy > 1:
        mask = np.ones(m * incy, bool)
        mask[::incy] = False
        assert np.array_equal(got[mask], y0[mask])


def test_sgemv_config_shape_col_major(dev, orc):
    # the configured 16384 x 8192 matrix stored column-major: 16 row blocks
    # x 12 column slices
    m, n = 16384, 8192
    A = torch.empty(n * m, dtype=torch.float32, device=dev)
    R.fill_uniform(A, -1.0, 1.0, 1234, inputs.S_SGEMV_A)
    x = torch.empty(n, dtype=torch.float32, device=dev)
    R.fill_uniform(x, -1.0, 1.0, 1234, inputs.S_SGEMV_X)
    y = R.sgemv(A.view(n, m), x, row_major=False)
    An, xn = A.cpu().numpy(), x.cpu().numpy()
    ref = orc.sgemv(False, m, n, 1.0, An, m, xn, 1, 0.0, np.zeros(m, np.float32), 1)
    scale = np.abs(An.reshape(n, m)).T @ np.abs(xn)
    check(y.cpu().numpy(), ref, scale)


def test_sgemv_col_major_many_slices(dev, orc):
    # one row block, 192 column slices of 104-105 columns (the eight-column
    # loop plus a remainder), the last lane's rows partly past m
    m, n = 1030, 20000
    rng = np.random.default_rng(11)
    A = rng.uniform(-1, 1, n * 1032).astype(np.float32)
    x = rng.uniform(-1, 1, n).astype(np.float32)
    y = R.sgemv(torch.from_numpy(A).to(dev).view(n, 1032), torch.from_numpy(x).to(dev), row_major=False, m=m)
    ref = orc.sgemv(False, m, n, 1.0, A, 1032, x, 1, 0.0, np.zeros(m, np.float32), 1)
    check(y.cpu().numpy(), ref, np.abs(A.reshape(n, 1032)[:, :m]).T @ np.abs(x))


def test_sgemv_unaligned_matrix(dev, orc):
    m, n = 77, 301
    A = np.random.default_rng(3).uniform(-1, 1, m * n + 1).astype(np.float32)
    x = np.random.default_rng(4).uniform(-1, 1, n).astype(np.float32)
    At = torch.from_numpy(A).to(dev)[1:].view(m, n)      # 4-byte offset: scalar path
    y = R.sgemv(At, torch.from_numpy(x).to(dev))
    ref = orc.sgemv(True, m, n, 1.0, A[1:], n, x, 1, 0.0, np.zeros(m, np.float32), 1)
    check(y.cpu().numpy(), ref, np.abs(A[1:].reshape(m, n)) @ np.abs(x))
    # column-major from the same misaligned storage (n columns of m rows)
    y = R.sgemv(torch.from_numpy(A).to(dev)[1:].view(n, m), torch.from_numpy(x).to(dev), row_major=False)
    ref = orc.sgemv(False, m, n, 1.0, A[1:], m, x, 1, 0.0, np.zeros(m, np.float32), 1)
    check(y.cpu().numpy(), ref, np.abs(A[1:].reshape(n, m)).T @ np.abs(x))


def test_sgemv_deterministic(dev):
    A = torch.rand(4096, 4096, device=dev)
    x = torch.rand(4096, device=dev)
    assert torch.equal(R.sgemv(A, x), R.sgemv(A, x))
    assert torch.equal(R.sgemv(A.t().contiguous(), x, row_major=False),
                       R.sgemv(A.t().contiguous(), x, row_major=False))
