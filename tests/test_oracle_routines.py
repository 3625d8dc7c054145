"""Pins the oracle's routines (reduction, sdot, sgemv, dgemm, Dot2, Ozaki,
hgemm, MLP) with checks that do not restate its loops: exact rational
arithmetic, closed forms, small-integer exactness, independent numpy
evaluations and finite differences.  A dropped term, a wrong index or a
transposed operand in the oracle fails at least one of these.
"""
import math
from fractions import Fraction

import numpy as np
import pytest

RNG = np.random.default_rng(1234)


# ---- reduction / sdot (reduction.cpp:61-67, sdot.cpp:96-102) ---------------
def test_sum_closed_forms(orc):
    assert orc.sum_f64(np.ones(1_000_003, np.float32)) == 1_000_003.0
    x = np.arange(1, 10001, dtype=np.float32)
    assert orc.sum_f64(x) == 10000 * 10001 / 2
    assert orc.sum_f64(np.zeros(0, np.float32)) == 0.0


def test_sum_matches_exact_sum(orc):
    x = RNG.uniform(-3, 3, 100_001).astype(np.float32)
    exact = math.fsum(float(v) for v in x)
    assert abs(orc.sum_f64(x) - exact) <= 1e-12 * np.abs(x).sum()
    # permutation invariance within the same bound
    assert abs(orc.sum_f64(x[::-1].copy()) - exact) <= 1e-12 * np.abs(x).sum()


def test_dot_basis_and_exact(orc):
    x = RNG.uniform(-1, 1, 4097).astype(np.float32)
    e = np.zeros_like(x)
    e[1234] = 1.0
    assert orc.dot_f64(x, e) == float(x[1234])
    y = RNG.uniform(-1, 1, 4097).astype(np.float32)
    exact = math.fsum(float(a) * float(b) for a, b in zip(x, y))   # float*float is exact in double
    assert abs(orc.dot_f64(x, y) - exact) <= 1e-13 * np.abs(x * y).sum()


# ---- sgemv (sgemv.cpp:132-147) ---------------------------------------------
@pytest.mark.parametrize("row_major", [True, False])
def test_sgemv_vs_float64(orc, row_major):
    m, n = 37, 53
    A = RNG.uniform(-1, 1, (m, n)).astype(np.float32)
    x = RNG.uniform(-1, 1, n).astype(np.float32)
    y0 = RNG.uniform(-1, 1, m).astype(np.float32)
    store = A if row_major else np.ascontiguousarray(A.T)
    lda = n if row_major else m
    got = orc.sgemv(row_major, m, n, 2.0, store, lda, x, 1, 0.5, y0, 1)
    ref = (2.0 * (A.astype(np.float64) @ x.astype(np.float64)) + 0.5 * y0.astype(np.float64)).astype(np.float32)
    assert np.array_equal(got, ref) or np.max(np.abs(got - ref) / np.abs(ref)) < 1e-6


def test_sgemv_strides_and_lda(orc):
    m, n, lda = 9, 7, 11
    A = RNG.uniform(-1, 1, (m, lda)).astype(np.float32)
    x = RNG.uniform(-1, 1, n * 3).astype(np.float32)
    y = np.zeros(m * 2, np.float32)
    got = orc.sgemv(True, m, n, 1.0, A, lda, x, 3, 0.0, y, 2)
    ref = A[:, :n].astype(np.float64) @ x[::3].astype(np.float64)
    assert np.allclose(got[::2], ref, rtol=1e-6)
    assert np.all(got[1::2] == 0)


# ---- dgemm (dgemm.cpp:122-133), Dot2 -----------------------------------------
def test_dgemm_small_integers_exact(orc):
    A = RNG.integers(-2, 3, (33, 2048)).astype(np.float64)
    B = RNG.integers(-2, 3, (2048, 29)).astype(np.float64)
    C0 = RNG.integers(-5, 6, (33, 29)).astype(np.float64)
    ref = A.astype(np.int64) @ B.astype(np.int64) * 2 + C0.astype(np.int64) * 3
    assert np.array_equal(orc.dgemm(A, B, C0, 2.0, 3.0), ref.astype(np.float64))


def test_dgemm_random(orc):
    A = RNG.standard_normal((40, 70))
    B = RNG.standard_normal((70, 50))
    ref = A @ B
    assert np.allclose(orc.dgemm(A, B), ref, rtol=1e-13, atol=1e-13)
    # orientation: C[i,j] uses row i of A and column j of B
    got = orc.dgemm(A, B)
    assert got[3, 5] == pytest.approx(float(np.dot(A[3], B[:, 5])), rel=1e-14)


def test_dot2_against_rationals(orc):
    K = 96
    # ill-conditioned: huge cancelling terms plus small ones
    A = RNG.standard_normal((4, K)) * (2.0 ** RNG.integers(-30, 30, (4, K)))
    B = RNG.standard_normal((K, 3))
    rows = np.array([0, 1, 2, 3, 0, 3], np.int32)
    cols = np.array([0, 1, 2, 0, 2, 1], np.int32)
    hi, lo = orc.dgemm_dot2(A, B, rows, cols)
    for t in range(rows.size):
        exact = sum(Fraction(float(A[rows[t], k])) * Fraction(float(B[k, cols[t]])) for k in range(K))
        err = abs(Fraction(float(hi[t])) + Fraction(float(lo[t])) - exact)
        scale = sum(abs(Fraction(float(A[rows[t], k])) * Fraction(float(B[k, cols[t]]))) for k in range(K))
        # Dot2 bound: |err| <= u^2 (K-ish) sum|a b| + u |exact|, u = 2^-53
        assert err <= Fraction(2) ** -100 * K * scale + Fraction(2) ** -52 * abs(exact)
        # and hi is the correctly rounded value up to one ulp
        assert abs(Fraction(float(hi[t])) - exact) <= Fraction(math.ulp(float(hi[t])))


# ---- Ozaki split and emulated dgemm ---------------------------------------
def test_ozaki_split_reconstructs(orc):
    X = RNG.standard_normal((6, 40)) * (2.0 ** RNG.integers(-20, 21, (6, 1)))
    X[2] = 0.0
    d, e = orc.ozaki_split_rows(X, 11)
    assert d.min() >= -16 and d.max() <= 16
    for r in range(6):
        mx = np.abs(X[r]).max()
        if mx > 0:
            assert 0.5 <= mx * 2.0 ** -int(e[r]) < 1.0
    # 11 digits resolve 2^(e - 54): the residual is at most half the last
    # digit's unit, and elements in the row maximum's binade (53 significant
    # bits at or above 2^(e-53)) come back exactly.  Horner from the last
    # digit keeps every partial value an exact remainder.
    rec = d[10].astype(np.float64)
    for p in range(9, -1, -1):
        rec = d[p].astype(np.float64) + rec / 32.0
    rec = rec / 16.0 * (2.0 ** e.astype(np.float64))[:, None]
    unit = 2.0 ** (e.astype(np.float64) - 54)[:, None]
    assert np.all(np.abs(rec - X) <= 0.5 * unit)
    top = np.abs(X) >= 2.0 ** (e.astype(np.float64) - 1)[:, None]
    assert np.array_equal(rec[top], X[top]) and top.sum() >= 5


def test_ozaki_exact_on_representable_inputs(orc):
    # entries k * 2^-9 with |k| <= 2^9 are exact in 2 digits (16t then 32r)
    A = RNG.integers(-511, 512, (24, 64)) / 512.0
    B = RNG.integers(-511, 512, (64, 16)) / 512.0
    ref = A @ B   # exact: small dyadic rationals, K=64
    assert np.array_equal(orc.ozaki_dgemm(A, B, 3), ref)


def test_ozaki_error_falls_with_slices(orc):
    A = RNG.standard_normal((32, 128))
    B = RNG.standard_normal((128, 24))
    rows, cols = np.meshgrid(np.arange(32, dtype=np.int32), np.arange(24, dtype=np.int32), indexing="ij")
    hi, lo = orc.dgemm_dot2(A, B, rows.ravel(), cols.ravel())
    ref = (hi + lo).reshape(32, 24)
    scale = np.abs(A) @ np.abs(B)
    errs = []
    for s in (3, 5, 7, 9, 11):
        C = orc.ozaki_dgemm(A, B, s)
        errs.append(float(np.max(np.abs(C - ref) / scale)))
    for a, b in zip(errs, errs[1:]):
        assert b < a / 100 or b < 1e-15
    # 5s - 1 bits relative to the row/column maxima
    assert errs[2] < 2.0 ** -30


def test_ozaki_alpha_beta_and_scaling(orc):
    A = RNG.standard_normal((8, 32)) * 2.0 ** 20
    B = RNG.standard_normal((32, 8)) * 2.0 ** -20
    C0 = RNG.standard_normal((8, 8))
    got = orc.ozaki_dgemm(A, B, 11, C0, 2.0, -1.0)
    assert np.allclose(got, 2.0 * (A @ B) - C0, rtol=1e-13, atol=1e-13)


# ---- hgemm (hgemm.cu:137-152) -----------------------------------------------
def test_hgemm_small_integers_exact(orc):
    A = RNG.integers(-2, 3, (20, 256)).astype(np.float16)
    B = RNG.integers(-2, 3, (256, 24)).astype(np.float16)
    ref = (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float16)
    assert np.array_equal(orc.hgemm(A, B), ref)


def test_hgemm_fp32_sequential_accumulation(orc):
    A = RNG.uniform(-1, 1, (5, 40)).astype(np.float16)
    B = RNG.uniform(-1, 1, (40, 6)).astype(np.float16)
    C0 = RNG.uniform(-1, 1, (5, 6)).astype(np.float16)
    ref = np.empty((5, 6), np.float16)
    for i in range(5):
        for j in range(6):
            s = np.float32(0)
            for k in range(40):
                s = np.float32(s + np.float32(A[i, k]) * np.float32(B[k, j]))
            ref[i, j] = np.float16(np.float32(np.float32(1.5) * s + np.float32(0.5) * np.float32(C0[i, j])))
    assert np.array_equal(orc.hgemm(A, B, C0, 1.5, 0.5), ref)


# ---- MLP cost model ---------------------------------------------------------
def test_mlp_gradient_finite_differences(orc):
    B, I_, H, O = 7, 4, 6, 5
    x = RNG.standard_normal((B, I_))
    W1 = RNG.standard_normal((H, I_)) * 0.5
    b1 = RNG.standard_normal(H) * 0.1
    W2 = RNG.standard_normal((O, H)) * 0.5
    b2 = RNG.standard_normal(O) * 0.1
    t = RNG.standard_normal((B, O))
    mask = (RNG.uniform(size=(B, O)) > 0.3).astype(np.uint8)
    params = [W1, b1, W2, b2]
    loss, grads = orc.mlp_loss_grad(x, *params, t, mask)
    # loss against an independent numpy forward
    h = np.maximum(x @ W1.T + b1, 0)
    y = h @ W2.T + b2
    assert loss == pytest.approx(float(((y - t) ** 2)[mask.astype(bool)].mean()), rel=1e-12)
    eps = 1e-6
    for p, g in zip(params, grads):
        for idx in np.ndindex(p.shape):
            old = p[idx]
            p[idx] = old + eps
            lp, _ = orc.mlp_loss_grad(x, *params, t, mask)
            p[idx] = old - eps
            lm, _ = orc.mlp_loss_grad(x, *params, t, mask)
            p[idx] = old
            assert g[idx] == pytest.approx((lp - lm) / (2 * eps), rel=1e-5, abs=1e-7)


def _nonfinite_inputs(M=9, N=7, K=32, seed=4):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((M, K))
    B = rng.standard_normal((K, N))
    A[3, 5] = np.nan            # row 3 of A
    B[7, 2] = np.inf            # column 2 of B
    B[1, 4] = -np.inf           # column 4 of B
    A[6, 1] = 0.0               # 0 * (-inf) in entry (6, 4): NaN
    return A, B


def test_ozaki_nonfinite_rows_and_columns(orc):
    # a row or column holding a NaN or an infinity: split exponent marked, digits
    # zero; its entries follow the FP64 product (NaN and infinities propagate,
    # compared with an exact float64 evaluation); every other entry equals the
    # Ozaki product of the same inputs with the bad rows/columns made finite
    A, B = _nonfinite_inputs()
    d, e = orc.ozaki_split_rows(A, 5)
    assert e[3] == -2 ** 31 and not d[:, 3].any() and (e[np.arange(9) != 3] != -2 ** 31).all()
    C0 = np.random.default_rng(1).standard_normal((9, 7))
    got = orc.ozaki_dgemm(A, B, 5, C0, 1.5, -0.5)
    with np.errstate(invalid="ignore"):
        plain = 1.5 * np.array([[sum(A[i, k] * B[k, j] for k in range(32)) for j in range(7)] for i in range(9)]) \
            - 0.5 * C0
    bad = np.zeros((9, 7), bool)
    bad[3, :] = True
    bad[:, [2, 4]] = True
    assert np.isnan(got[3]).all()
    assert np.array_equal(np.isnan(got[bad]), np.isnan(plain[bad]))
    assert np.array_equal(got[bad & ~np.isnan(got)], plain[bad & ~np.isnan(plain)])   # the infinities
    assert np.isnan(got[6, 4]) and np.isinf(got[0, 2])   # 0 * -inf; a finite row against +inf
    Af, Bf = A.copy(), B.copy()
    Af[3] = 1.0
    Bf[:, [2, 4]] = 1.0
    fin = orc.ozaki_dgemm(Af, Bf, 5, C0, 1.5, -0.5)
    assert np.array_equal(got[~bad], fin[~bad]) and np.isfinite(got[~bad]).all()
