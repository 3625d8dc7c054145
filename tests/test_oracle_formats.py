"""Pins the oracle's number formats and input generator.

Checks come from outside the oracle: the published SplitMix64 test vector,
numpy's IEEE binary16 conversion, and the OCP FP8 E4M3 table values.
"""
import numpy as np
import pytest

from libthom_amd import inputs


def test_splitmix64_published_vector(orc):
    # Vigna's splitmix64.c, seed 1234567: the first five outputs
    want = [6457827717110365317, 3203168211198807973, 9817491932198370423, 4593380528125082431,
            16408922859458223821]
    assert orc.splitmix64_seq(1234567, 5) == want
    assert [orc.rng_u64(1234567, 0, i) for i in range(5)] == want
    assert list(inputs.rng_u64(5, 1234567, 0)) == want


def test_stream_offset_is_a_sequential_run(orc):
    # stream s starts at seed + s*0xD1B54A32D192ED03
    start = (99 + 3 * 0xD1B54A32D192ED03) & ((1 << 64) - 1)
    assert [orc.rng_u64(99, 3, i) for i in range(4)] == orc.splitmix64_seq(start, 4)


@pytest.mark.parametrize("dtype", [np.float32, np.float16, np.float64])
def test_numpy_generator_matches_oracle(orc, dtype):
    n = 50_000
    o = orc.fill_uniform(n, -1.0, 1.0, 1234, 7, dtype)
    g = {np.float32: inputs.uniform_f32, np.float16: inputs.uniform_f16, np.float64: inputs.uniform_f64}[dtype]
    p = g(n, -1.0, 1.0, 1234, 7)
    assert p.dtype == dtype
    assert np.array_equal(o.view(np.uint8), p.view(np.uint8))


def test_uniform_range_and_moments(orc):
    x = orc.fill_uniform(200_000, 0.0, 1.0, 1234, 1, np.float64)
    assert x.min() >= 0.0 and x.max() < 1.0
    assert abs(x.mean() - 0.5) < 0.005 and abs(x.var() - 1 / 12) < 0.002


def test_int_generator(orc):
    o = orc.fill_int_u8(100_000, 0, 15, 1234, 10)
    p = inputs.int_uniform(100_000, 0, 15, 1234, 10)
    assert np.array_equal(o.astype(np.int64), p)
    counts = np.bincount(p, minlength=16)
    assert counts.min() > 5800 and counts.max() < 6700


def test_normal_generator(orc):
    o = orc.fill_normal(100_001, 1234, 4)
    p = inputs.normal_f64(100_001, 1234, 4)
    assert np.allclose(o, p, rtol=1e-14, atol=1e-14)
    assert abs(o.mean()) < 0.02 and abs(o.std() - 1.0) < 0.02


def test_f16_to_f32_exhaustive(orc):
    bits = np.arange(65536, dtype=np.uint16)
    want = bits.view(np.float16).astype(np.float32)
    got = np.array([orc.f16_to_f32(int(b)) for b in bits], dtype=np.float32)
    nan = np.isnan(want)
    assert np.array_equal(np.isnan(got), nan)
    assert np.array_equal(got[~nan].view(np.uint32), want[~nan].view(np.uint32))


def test_f32_to_f16_against_numpy(orc):
    # every binary16 value, every midpoint between neighbours (ties to even),
    # and points just off the midpoints, plus random floats across the range
    h = np.arange(0, 0x7C00, dtype=np.uint16).view(np.float16).astype(np.float64)
    mids = (h[:-1] + h[1:]) / 2
    cases = np.concatenate([h, mids, np.nextafter(mids, 0), np.nextafter(mids, 1e9),
                            [65504.0, 65519.99, 65520.0, 1e6, 2.0 ** -25, 2.0 ** -24 * 1.5, 2.0 ** -26, 0.0]])
    cases = cases.astype(np.float32)
    rng = np.random.default_rng(0)
    cases = np.concatenate([cases, -cases, (rng.standard_normal(20000) * 10.0 ** rng.uniform(-9, 5, 20000))
                           .astype(np.float32)])
    with np.errstate(over="ignore"):
        want = cases.astype(np.float16).view(np.uint16)
    got = np.array([orc.f32_to_f16(float(c)) for c in cases], dtype=np.uint16)
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, [(float(cases[i]), hex(got[i]), hex(want[i])) for i in bad[:5]]


def test_e4m3_known_values(orc):
    # OCP 8-bit floating point spec, E4M3: bias 7, no infinities, S.1111.111 = NaN
    known = {0x38: 1.0, 0x58: 16.0, 0xD8: -16.0, 0x7E: 448.0, 0x08: 2.0 ** -6, 0x01: 2.0 ** -9,
             0x07: 7 * 2.0 ** -9, 0x40: 2.0, 0x3C: 1.5}
    for b, v in known.items():
        assert orc.e4m3_to_f32(b) == v
    assert np.isnan(orc.e4m3_to_f32(0x7F)) and np.isnan(orc.e4m3_to_f32(0xFF))
    for d in range(-16, 17):
        assert orc.e4m3_to_f32(orc.int_to_e4m3(d)) == d
    # every byte decodes to a distinct value (except +-0 and the two NaNs)
    vals = [orc.e4m3_to_f32(b) for b in range(256)]
    finite = [v for v in vals if not np.isnan(v)]
    assert len(set(finite)) == 253
