"""GPU parity: input generation, reduction and sdot against the oracle.

Bars: generators bit-exact; reduction/sdot within rel 1e-5 of the oracle's
FP64 sum (the configuration's tolerance), exact on integer inputs, bitwise
deterministic across runs.
"""
import numpy as np
import pytest
import torch

from libthom_amd import inputs, routines as R

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype,npt", [(torch.float32, np.float32), (torch.float16, np.float16),
                                       (torch.float64, np.float64)])
def test_fill_uniform_bit_exact(dev, orc, dtype, npt):
    n = 1_000_003
    t = torch.empty(n, dtype=dtype, device=dev)
    R.fill_uniform(t, -1.0, 1.0, 1234, 5)
    got = t.cpu().numpy()
    want = orc.fill_uniform(n, -1.0, 1.0, 1234, 5, npt)
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8))


def test_copy(dev):
    src = torch.arange(1_000_001, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    R.copy_f32(dst, src)
    assert torch.equal(dst, src)


SIZES = [0, 1, 3, 4, 5, 255, 1000, 65537, 1_000_000, 10_000_019]


@pytest.mark.parametrize("n", SIZES)
def test_reduction_vs_oracle(dev, orc, n):
    x = inputs.uniform_f32(n, 0.0, 1.0, 1234, inputs.S_REDUCE)
    got = float(R.reduce_sum(torch.from_numpy(x).to(dev)).item())
    ref = orc.sum_f64(x)
    if n == 0:
        assert got == 0.0
    else:
        assert abs(got - ref) <= 1e-5 * abs(ref), (got, ref)


def test_reduction_signed_and_unaligned(dev, orc):
    x = inputs.uniform_f32(3_000_001, -3.0, 3.0, 1234, 21)
    t = torch.from_numpy(x).to(dev)
    for off in (1, 2, 3):   # misaligned start -> scalar path
        got = float(R.reduce_sum(t[off:]).item())
        ref = orc.sum_f64(x[off:])
        assert abs(got - ref) <= 1e-5 * np.abs(x[off:]).sum()


def test_reduction_exact_integers_and_deterministic(dev):
    n = 1 << 24
    t = torch.ones(n, dtype=torch.float32, device=dev)
    assert R.reduce_sum(t).item() == float(n)
    x = torch.rand(5_000_000, device=dev)
    a = R.reduce_sum(x).clone()
    b = R.reduce_sum(x).clone()
    assert torch.equal(a, b)     # fixed order, and the ticket counter was reset


def test_reduction_resets_counter(dev):
    # the last block resets the ticket counter (the workspace's first 256
    # bytes hold it): zero again after each call, small and large grids; a NaN
    # or an infinity in the input reaches the result
    from libthom_amd._lib import lib
    L = lib()
    stream = torch.cuda.current_stream(dev).cuda_stream
    for n, bad in ((1_000_000, None), (60_000_000, None), (3, None), (1_000_000, float("nan")),
                   (2_000_001, float("inf"))):
        x = torch.rand(n, device=dev)
        if bad is not None:
            x[n // 2] = bad
        need = L.lthip_reduce_sum_f32_workspace_size(n)
        ws = torch.zeros(need, dtype=torch.uint8, device=dev)
        out = torch.empty(1, dtype=torch.float32, device=dev)
        print(f"reduction workspace n={n} bad={bad}", flush=True)
        for _ in range(2):
            assert L.lthip_reduce_sum_f32(x.data_ptr(), n, out.data_ptr(), ws.data_ptr(), need, stream) == 0
            torch.cuda.synchronize()
            assert int(ws[:256].count_nonzero()) == 0, n
        got = out.item()
        if bad is None:
            ref = float(x.double().sum())
            assert abs(got - ref) <= 1e-5 * ref, (n, got, ref)
        elif bad != bad:
            assert got != got
        else:
            assert got == bad


@pytest.mark.parametrize("n", [1, 7, 4096, 1_000_000, 9_999_999])
def test_sdot_vs_oracle(dev, orc, n):
    x = inputs.uniform_f32(n, -1.0, 1.0, 1234, inputs.S_SDOT_X)
    y = inputs.uniform_f32(n, -1.0, 1.0, 1234, inputs.S_SDOT_Y)
    got = float(R.sdot(torch.from_numpy(x).to(dev), torch.from_numpy(y).to(dev)).item())
    ref = orc.dot_f64(x, y)
    # signed inputs cancel: scale the tolerance by sum |x y|
    assert abs(got - ref) <= 1e-5 * float(np.abs(x.astype(np.float64) * y).sum()), (got, ref)


def test_sdot_basis_vector(dev):
    x = torch.rand(100_003, device=dev)
    e = torch.zeros_like(x)
    e[77777] = 1.0
    assert R.sdot(x, e).item() == x[77777].item()


def test_read_probe_reads_every_word(dev):
    # the XOR of the per-thread XORs is the XOR of every input word: each
    # element is read exactly once
    from libthom_amd._lib import call
    n = (1 << 20) + 4 * 37
    x = torch.empty(n, dtype=torch.float32, device=dev)
    R.fill_uniform(x, -1, 1, 5, 6)
    blocks = 64
    sink = torch.empty(blocks * 256, dtype=torch.float32, device=dev)
    call("lthip_probe_read_f32", x.data_ptr(), n, sink.data_ptr(), blocks, torch.cuda.current_stream(dev).cuda_stream)
    got = np.bitwise_xor.reduce(sink.cpu().numpy().view(np.uint32))
    want = np.bitwise_xor.reduce(x.cpu().numpy().view(np.uint32))
    assert got == want


def test_measured_peaks_are_plausible(dev):
    p = R.measured_peaks(dev, read_bytes=256 << 20)
    assert p["fp8_tflops"] > p["fp16_tflops"] > 100 * p["fp64_tflops"] / 10 > 1
    assert 500 < p["hbm_read_gbs"] and 500 < p["hbm_copy_gbs"]
