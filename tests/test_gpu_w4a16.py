"""GPU parity: int4 dequantisation (bit-exact) and w4a16 against the oracle.

The dequantisation is checked bit for bit over every (q, z, finite scale bit
pattern).  The product uses the reference's criterion (w4a16.cpp:194-197):
|r - ref| <= 1e-2 |ref| + 1e-3 RMS(C_ref), against the oracle's FP32
k-ascending sum over the oracle-dequantised weights.
"""
import numpy as np
import pytest
import torch

from libthom_amd import W4A16Config, inputs, routines as R

pytestmark = pytest.mark.gpu


def t(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def i32(a, dev):
    return t(np.ascontiguousarray(a).view(np.int32), dev)


def test_dequant_exhaustive_bit_exact(dev, orc):
    group = 16
    s_bits = np.arange(65536, dtype=np.uint16)
    s_bits = s_bits[(s_bits & 0x7C00) != 0x7C00]
    N = s_bits.size * 16
    N -= N % 8
    z = np.tile(np.arange(16, dtype=np.uint8), s_bits.size)[None, :N]
    scales = np.repeat(s_bits, 16)[None, :N].view(np.float16)
    q = np.repeat(np.arange(16, dtype=np.uint8)[:, None], N, axis=1)
    packed = orc.w4_pack(q)
    qz = orc.w4_pack_zeros(z)
    want = orc.w4_dequant(packed, scales, qz, group)
    got = R.w4_dequant(i32(packed, dev), t(scales, dev), i32(qz, dev), group).cpu().numpy()
    assert np.array_equal(got.view(np.uint16), want.view(np.uint16))


def test_dequant_golden_and_symmetric(dev, orc):
    import json
    import os
    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "w4_tiny.json")))
    packed = np.array(g["packed"], dtype=np.uint32)
    scales = np.array(g["scales_bits"], dtype=np.uint16).view(np.float16)
    qz = np.array(g["qzeros"], dtype=np.uint32)
    got = R.w4_dequant(i32(packed, dev), t(scales, dev), i32(qz, dev), g["group"]).cpu().numpy()
    assert got.view(np.uint16).tolist() == g["dequant_bits"]
    q = np.random.default_rng(1).integers(0, 16, (256, 64)).astype(np.uint8)
    s = np.random.default_rng(2).uniform(0.005, 0.02, (2, 64)).astype(np.float16)
    p = orc.w4_pack(q)
    got = R.w4_dequant(i32(p, dev), t(s, dev), None, 128).cpu().numpy()
    assert np.array_equal(got, orc.w4_dequant(p, s, None, 128))


def w4_close(got, ref):
    got = got.astype(np.float32)
    ref = ref.astype(np.float32)
    rms = float(np.sqrt(np.mean(ref.astype(np.float64) ** 2)))
    bad = np.abs(got - ref) > 1e-2 * np.abs(ref) + 1e-3 * rms
    assert not bad.any(), f"{bad.sum()} of {bad.size} outside 1% + 1e-3 RMS; max abs err {np.abs(got - ref).max()}"


_PROBLEMS = {}


def problem(M, N=11008, K=4096, group=128):
    key = (M, N, K, group)
    if key not in _PROBLEMS:
        A, q, s, z = inputs.w4a16_problem(M, N, K, group)
        _PROBLEMS[key] = (A, inputs.pack_w4(q), s, inputs.pack_zeros(z))
    return _PROBLEMS[key]


@pytest.mark.parametrize("M", [1, 16, 128, 2048])
def test_w4a16_config_shapes(dev, orc, M):
    N, K, group = 11008, 4096, 128
    A, packed, s, qz = problem(M)
    C = R.w4a16(t(A, dev), i32(packed, dev), t(s, dev), i32(qz, dev), group).cpu().numpy()
    Wdq = orc.w4_dequant(packed, s, qz, group)
    # M = 2048: every column against the oracle on row ranges spread over all
    # eight 256-row tiles (two of them across a tile edge) and the last 32 rows
    ranges = [(0, M)] if M <= 128 else [(0, 8), (250, 12), (517, 8), (1020, 8), (1279, 6), (1540, 8),
                                        (1790, 4), (2016, 32)]
    for row0, rows in ranges:
        ref = orc.w4a16_rows(A, Wdq, row0, rows)
        w4_close(C[row0:row0 + rows], ref)


@pytest.mark.parametrize("cfg", [W4A16Config(1, 64, 16, 1), W4A16Config(1, 64, 16, 4), W4A16Config(1, 64, 16, 32),
                                 W4A16Config(2, 64, 64, 1), W4A16Config(2, 128, 128, 1),
                                 W4A16Config(2, 256, 256, 1), W4A16Config(2, 128, 256, 1),
                                 W4A16Config(2, 128, 64, 4), W4A16Config(2, 64, 64, 8)])
@pytest.mark.parametrize("M", [1, 5, 16, 33, 64])
def test_w4a16_kinds(dev, orc, cfg, M):
    N, K, group = 520, 1024, 128
    A, q, s, z = inputs.w4a16_problem(M, N, K, group, seed=77)
    packed, qz = inputs.pack_w4(q), inputs.pack_zeros(z)
    C = R.w4a16(t(A, dev), i32(packed, dev), t(s, dev), i32(qz, dev), group, config=cfg).cpu().numpy()
    ref = orc.w4a16_rows(A, orc.w4_dequant(packed, s, qz, group))
    w4_close(C, ref)


def test_w4a16_exact_small_integers(dev, orc):
    # scale 1, A small integers: every product and sum is exact in FP32
    M, N, K = 20, 64, 256
    rng = np.random.default_rng(5)
    q = rng.integers(0, 16, (K, N)).astype(np.uint8)
    z = rng.integers(0, 16, (K // 128, N)).astype(np.uint8)
    s = np.ones((K // 128, N), np.float16)
    A = rng.integers(-2, 3, (M, K)).astype(np.float16)
    packed, qz = orc.w4_pack(q), orc.w4_pack_zeros(z)
    ref = orc.w4a16_rows(A, orc.w4_dequant(packed, s, qz, 128))
    for cfg in (W4A16Config(1, 64, 16, 1), W4A16Config(1, 64, 16, 4), W4A16Config(2, 64, 64, 1)):
        C = R.w4a16(t(A, dev), i32(packed, dev), t(s, dev), i32(qz, dev), 128, config=cfg).cpu().numpy()
        assert np.array_equal(C, ref), cfg


def test_w4a16_symmetric_zero_point(dev, orc):
    M, N, K = 16, 256, 512
    A, q, s, _ = inputs.w4a16_problem(M, N, K, 128, seed=3)
    packed = inputs.pack_w4(q)
    ref = orc.w4a16_rows(A, orc.w4_dequant(packed, s, None, 128))
    for cfg in (None, W4A16Config(2, 64, 64, 1)):
        C = R.w4a16(t(A, dev), i32(packed, dev), t(s, dev), None, 128, config=cfg).cpu().numpy()
        w4_close(C, ref)


KIND3 = [W4A16Config(3, 256, 256, 1), W4A16Config(3, 128, 256, 1), W4A16Config(3, 256, 128, 1),
         W4A16Config(3, 128, 128, 1), W4A16Config(3, 128, 128, 3), W4A16Config(3, 256, 256, 2),
         W4A16Config(3, 256, 64, 1), W4A16Config(3, 256, 64, 2), W4A16Config(3, 128, 64, 1),
         W4A16Config(3, 64, 64, 1), W4A16Config(3, 128, 64, 3)]


@pytest.mark.parametrize("cfg", KIND3)
@pytest.mark.parametrize("M", [1, 65, 130, 300])
@pytest.mark.parametrize("group", [64, 128, 256])
def test_w4a16_pipelined(dev, orc, cfg, M, group):
    N, K = 544, 1024
    A, q, s, z = inputs.w4a16_problem(M, N, K, group, seed=91 + M)
    packed, qz = inputs.pack_w4(q), inputs.pack_zeros(z)
    C = R.w4a16(t(A, dev), i32(packed, dev), t(s, dev), i32(qz, dev), group, config=cfg).cpu().numpy()
    ref = orc.w4a16_rows(A, orc.w4_dequant(packed, s, qz, group))
    w4_close(C, ref)


def test_w4a16_pipelined_exact_and_symmetric(dev, orc):
    M, N, K = 200, 288, 512
    rng = np.random.default_rng(15)
    q = rng.integers(0, 16, (K, N)).astype(np.uint8)
    z = rng.integers(0, 16, (K // 128, N)).astype(np.uint8)
    s = np.ones((K // 128, N), np.float16)
    A = rng.integers(-2, 3, (M, K)).astype(np.float16)
    packed, qz = orc.w4_pack(q), orc.w4_pack_zeros(z)
    for cfg in KIND3:
        C = R.w4a16(t(A, dev), i32(packed, dev), t(s, dev), i32(qz, dev), 128, config=cfg).cpu().numpy()
        assert np.array_equal(C, orc.w4a16_rows(A, orc.w4_dequant(packed, s, qz, 128))), cfg
        C = R.w4a16(t(A, dev), i32(packed, dev), t(s, dev), None, 128, config=cfg).cpu().numpy()
        assert np.array_equal(C, orc.w4a16_rows(A, orc.w4_dequant(packed, s, None, 128))), cfg


def test_w4a16_pipelined_rejects_unsupported(dev):
    from libthom_amd._lib import LthipError
    M, N, K = 64, 520, 256
    A, q, s, z = inputs.w4a16_problem(M, N, K, 128, seed=5)
    with pytest.raises(LthipError):
        R.w4a16(t(A, dev), i32(inputs.pack_w4(q), dev), t(s, dev), i32(inputs.pack_zeros(z), dev), 128,
                config=W4A16Config(3, 128, 128, 1))


KIND4 = [W4A16Config(4, 64, 0, 1), W4A16Config(4, 32, 0, 1)]


@pytest.mark.parametrize("cfg", KIND4)
@pytest.mark.parametrize("M", [1, 2, 3, 4])
@pytest.mark.parametrize("group", [32, 64, 128, 256])
def test_w4a16_decode(dev, orc, cfg, M, group):
    from libthom_amd._lib import LthipError
    N, K = 520, 1024
    A, q, s, z = inputs.w4a16_problem(M, N, K, group, seed=131 + M)
    packed, qz = inputs.pack_w4(q), inputs.pack_zeros(z)
    args = (t(A, dev), i32(packed, dev), t(s, dev), i32(qz, dev), group)
    if group % (2048 // cfg.tile_n) != 0:
        with pytest.raises(LthipError):
            R.w4a16(*args, config=cfg)
        return
    C = R.w4a16(*args, config=cfg).cpu().numpy()
    w4_close(C, orc.w4a16_rows(A, orc.w4_dequant(packed, s, qz, group)))


@pytest.mark.parametrize("cfg", KIND4)
@pytest.mark.parametrize("N,K", [(8, 128), (24, 256), (4104, 384), (264, 32768)])
def test_w4a16_decode_edges(dev, orc, cfg, N, K):
    # narrower than one block, ragged last block, K at the A-staging bound
    # (dynamic LDS above 64 KiB)
    M = 1
    A, q, s, z = inputs.w4a16_problem(M, N, K, 128, seed=N + K)
    packed, qz = inputs.pack_w4(q), inputs.pack_zeros(z)
    C = R.w4a16(t(A, dev), i32(packed, dev), t(s, dev), i32(qz, dev), 128, config=cfg).cpu().numpy()
    w4_close(C, orc.w4a16_rows(A, orc.w4_dequant(packed, s, qz, 128)))


def test_w4a16_decode_exact_and_symmetric(dev, orc):
    # scale a power of two, A small integers: every product and FP32 sum is
    # exact, so the decode kernel matches the oracle bit for bit
    N, K = 200, 1024
    rng = np.random.default_rng(21)
    q = rng.integers(0, 16, (K, N)).astype(np.uint8)
    z = rng.integers(0, 16, (K // 128, N)).astype(np.uint8)
    packed, qz = orc.w4_pack(q), orc.w4_pack_zeros(z)
    for M in (1, 2, 3, 4):
        A = rng.integers(-3, 4, (M, K)).astype(np.float16)
        for scale in (1.0, 0.5):
            s = np.full((K // 128, N), scale, np.float16)
            for cfg in KIND4:
                C = R.w4a16(t(A, dev), i32(packed, dev), t(s, dev), i32(qz, dev), 128, config=cfg).cpu().numpy()
                assert np.array_equal(C, orc.w4a16_rows(A, orc.w4_dequant(packed, s, qz, 128))), (M, scale, cfg)
                C = R.w4a16(t(A, dev), i32(packed, dev), t(s, dev), None, 128, config=cfg).cpu().numpy()
                assert np.array_equal(C, orc.w4a16_rows(A, orc.w4_dequant(packed, s, None, 128))), (M, scale, cfg)


def test_w4a16_decode_auto_matches_stream(dev, orc):
    # the automatic choice at M <= 2 is the decode kernel, at M = 3, 4 the
    # skinny MFMA kernel; both within the product tolerance, as is the MFMA
    # stream kernel
    for M, N in ((1, 11008), (2, 4096), (3, 4096), (4, 10240)):
        K = 1024
        A, q, s, z = inputs.w4a16_problem(M, N, K, 128, seed=7 * M)
        packed, qz = inputs.pack_w4(q), inputs.pack_zeros(z)
        args = (t(A, dev), i32(packed, dev), t(s, dev), i32(qz, dev), 128)
        ref = orc.w4a16_rows(A, orc.w4_dequant(packed, s, qz, 128))
        w4_close(R.w4a16(*args).cpu().numpy(), ref)
        w4_close(R.w4a16(*args, config=W4A16Config(1, 64, 16, 1)).cpu().numpy(), ref)


def test_w4a16_decode_rejects_unsupported(dev):
    from libthom_amd._lib import LthipError
    A, q, s, z = inputs.w4a16_problem(5, 256, 512, 128, seed=9)
    packed, qz = i32(inputs.pack_w4(q), dev), i32(inputs.pack_zeros(z), dev)
    At, st = t(A, dev), t(s, dev)
    for M, cfg in ((5, W4A16Config(4, 64, 0, 1)),      # M > 4
                   (1, W4A16Config(4, 48, 0, 1)),      # tile_n not built
                   (1, W4A16Config(4, 64, 0, 2))):     # split-K
        with pytest.raises(LthipError):
            R.w4a16(At[:M], packed, st, qz, 128, config=cfg)
    A4, q4, s4, z4 = inputs.w4a16_problem(4, 64, 16384, 128, seed=10)   # M * K above the staging bound
    with pytest.raises(LthipError):
        R.w4a16(t(A4, dev), i32(inputs.pack_w4(q4), dev), t(s4, dev), i32(inputs.pack_zeros(z4), dev), 128,
                config=W4A16Config(4, 64, 0, 1))


KIND5 = W4A16Config(5, 64, 0, 1)
KIND5_32 = W4A16Config(5, 32, 0, 1)


@pytest.mark.parametrize("cfg", [KIND5, KIND5_32], ids=["cols64", "cols32"])
@pytest.mark.parametrize("M", [1, 5, 16, 17, 33, 64])
@pytest.mark.parametrize("group", [32, 64, 128, 256])
def test_w4a16_skinny(dev, orc, cfg, M, group):
    N, K = 520, 1024
    A, q, s, z = inputs.w4a16_problem(M, N, K, group, seed=171 + M)
    packed, qz = inputs.pack_w4(q), inputs.pack_zeros(z)
    C = R.w4a16(t(A, dev), i32(packed, dev), t(s, dev), i32(qz, dev), group, config=cfg).cpu().numpy()
    w4_close(C, orc.w4a16_rows(A, orc.w4_dequant(packed, s, qz, group)))


@pytest.mark.parametrize("cfg", [KIND5, KIND5_32], ids=["cols64", "cols32"])
@pytest.mark.parametrize("N,K", [(8, 128), (24, 256), (4104, 384)])
def test_w4a16_skinny_edges(dev, orc, cfg, N, K):
    M = 20
    A, q, s, z = inputs.w4a16_problem(M, N, K, 128, seed=N + 3 * K)
    packed, qz = inputs.pack_w4(q), inputs.pack_zeros(z)
    C = R.w4a16(t(A, dev), i32(packed, dev), t(s, dev), i32(qz, dev), 128, config=cfg).cpu().numpy()
    w4_close(C, orc.w4a16_rows(A, orc.w4_dequant(packed, s, qz, 128)))


def test_w4a16_skinny_exact_and_symmetric(dev, orc):
    N, K = 200, 1024
    rng = np.random.default_rng(23)
    q = rng.integers(0, 16, (K, N)).astype(np.uint8)
    z = rng.integers(0, 16, (K // 64, N)).astype(np.uint8)
    packed, qz = orc.w4_pack(q), orc.w4_pack_zeros(z)
    for M in (3, 16, 40):
        A = rng.integers(-3, 4, (M, K)).astype(np.float16)
        for scale in (1.0, 0.25):
            s = np.full((K // 64, N), scale, np.float16)
            C = R.w4a16(t(A, dev), i32(packed, dev), t(s, dev), i32(qz, dev), 64, config=KIND5).cpu().numpy()
            assert np.array_equal(C, orc.w4a16_rows(A, orc.w4_dequant(packed, s, qz, 64))), (M, scale)
            C = R.w4a16(t(A, dev), i32(packed, dev), t(s, dev), None, 64, config=KIND5).cpu().numpy()
            assert np.array_equal(C, orc.w4a16_rows(A, orc.w4_dequant(packed, s, None, 64))), (M, scale)


def test_w4a16_skinny_rejects_unsupported(dev):
    from libthom_amd._lib import LthipError
    A, q, s, z = inputs.w4a16_problem(65, 256, 512, 128, seed=19)
    packed, qz = i32(inputs.pack_w4(q), dev), i32(inputs.pack_zeros(z), dev)
    At, st = t(A, dev), t(s, dev)
    for M, cfg in ((65, KIND5), (16, W4A16Config(5, 48, 0, 1)), (16, W4A16Config(5, 64, 0, 2))):
        with pytest.raises(LthipError):
            R.w4a16(At[:M], packed, st, qz, 128, config=cfg)
    A2, q2, s2, z2 = inputs.w4a16_problem(8, 64, 512, 16, seed=20)   # group 16
    with pytest.raises(LthipError):
        R.w4a16(t(A2, dev), i32(inputs.pack_w4(q2), dev), t(s2, dev), i32(inputs.pack_zeros(z2), dev), 16,
                config=KIND5)


KIND6 = W4A16Config(6, 128, 128, 1)


@pytest.mark.parametrize("M", [1, 65, 128, 130, 256, 300])
@pytest.mark.parametrize("group", [64, 128, 256])
@pytest.mark.parametrize("N,K", [(544, 1024), (2080, 4096)])
def test_w4a16_streamk(dev, orc, M, group, N, K):
    # N % 128 == 32: a ragged last column tile; K = 4096 at N = 2080 gives
    # ranges that start and end inside tiles and tiles shared by 3-4 blocks
    A, q, s, z = inputs.w4a16_problem(M, N, K, group, seed=211 + M + N)
    packed, qz = inputs.pack_w4(q), inputs.pack_zeros(z)
    C = R.w4a16(t(A, dev), i32(packed, dev), t(s, dev), i32(qz, dev), group, config=KIND6).cpu().numpy()
    w4_close(C, orc.w4a16_rows(A, orc.w4_dequant(packed, s, qz, group)))


@pytest.mark.parametrize("K", [512, 1408, 2048])
def test_w4a16_streamk_exact_and_symmetric(dev, orc, K):
    # scale 1, A small integers: every FP32 sum is an integer below 2^16, so the
    # split tiles' partials add up exactly and the result equals the oracle bit
    # for bit.  K = 512: tiles shared by two blocks; 1408: block ranges that
    # cross tile edges; 2048: tiles shared by eight blocks
    M, N = 200, 288
    rng = np.random.default_rng(K)
    q = rng.integers(0, 16, (K, N)).astype(np.uint8)
    z = rng.integers(0, 16, (K // 128, N)).astype(np.uint8)
    s = np.ones((K // 128, N), np.float16)
    A = rng.integers(-2, 3, (M, K)).astype(np.float16)
    packed, qz = orc.w4_pack(q), orc.w4_pack_zeros(z)
    for zeros in (qz, None):
        want = orc.w4a16_rows(A, orc.w4_dequant(packed, s, zeros, 128))
        C = R.w4a16(t(A, dev), i32(packed, dev), t(s, dev), None if zeros is None else i32(zeros, dev), 128,
                    config=KIND6).cpu().numpy()
        assert np.array_equal(C.view(np.uint16), want.view(np.uint16)), (K, zeros is None)


def test_w4a16_streamk_repeatable(dev, orc):
    # the arrival counters return to zero after each call: repeated calls,
    # and a call of another shape in between, give identical bits
    shapes = [(128, 2080, 4096), (200, 544, 1408)]
    data = []
    for M, N, K in shapes:
        A, q, s, z = inputs.w4a16_problem(M, N, K, 128, seed=M + K)
        data.append((t(A, dev), i32(inputs.pack_w4(q), dev), t(s, dev), i32(inputs.pack_zeros(z), dev)))
    first = [R.w4a16(*d, 128, config=KIND6).clone() for d in data]
    for _ in range(3):
        for d, f in zip(data, first):
            assert torch.equal(R.w4a16(*d, 128, config=KIND6), f)


@pytest.mark.parametrize("M,N,K,group", [(65, 512, 2048, 32), (261, 512, 3072, 32), (130, 1856, 2048, 16),
                                          (300, 704, 1024, 32), (1100, 640, 1024, 32), (96, 3520, 1536, 8)])
def test_w4a16_auto_kind2_fallback(dev, orc, M, N, K, group):
    # group not a multiple of 64, M > 64: the automatic choice runs the
    # LDS-dequantising tile with its own tile and K split (split 8 at
    # 261 x 512 x 3072); with the queried workspace and with none (the split
    # shrinks to 1) both meet the product tolerance
    from libthom_amd._lib import call, lib
    A, q, s, z = inputs.w4a16_problem(M, N, K, group, seed=M + N)
    packed, qz = inputs.pack_w4(q), inputs.pack_zeros(z)
    ref = orc.w4a16_rows(A, orc.w4_dequant(packed, s, qz, group))
    args = (t(A, dev), i32(packed, dev), t(s, dev), i32(qz, dev), group)
    if (M, N, K) == (261, 512, 3072):
        assert lib().lthip_w4a16_workspace_size(M, N, K, None) >= 8 * M * N * 4
    w4_close(R.w4a16(*args).cpu().numpy(), ref)
    C = torch.empty(M, N, dtype=torch.float16, device=dev)
    call("lthip_w4a16_gemm", args[0].data_ptr(), args[1].data_ptr(), args[2].data_ptr(), args[3].data_ptr(),
         C.data_ptr(), M, N, K, group, None, None, 0, torch.cuda.current_stream(dev).cuda_stream)
    w4_close(C.cpu().numpy(), ref)


def test_w4a16_streamk_rejects_and_auto_falls_back(dev, orc):
    from libthom_amd._lib import LthipError
    A, q, s, z = inputs.w4a16_problem(128, 520, 512, 128, seed=31)   # N % 32 != 0
    args = (t(A, dev), i32(inputs.pack_w4(q), dev), t(s, dev), i32(inputs.pack_zeros(z), dev), 128)
    with pytest.raises(LthipError):
        R.w4a16(*args, config=KIND6)
    A2, q2, s2, z2 = inputs.w4a16_problem(128, 544, 512, 32, seed=32)   # group 32
    packed2, qz2 = inputs.pack_w4(q2), inputs.pack_zeros(z2)
    args2 = (t(A2, dev), i32(packed2, dev), t(s2, dev), i32(qz2, dev), 32)
    for cfg in (KIND6, W4A16Config(6, 128, 128, 2), W4A16Config(6, 256, 128, 1)):
        with pytest.raises(LthipError):
            R.w4a16(*args2, config=cfg)
    # the automatic choice at M = 128 is kind 6; where kind 6 does not take
    # the shape it falls back (group 32: kind 2), within the product tolerance
    w4_close(R.w4a16(*args2).cpu().numpy(), orc.w4a16_rows(A2, orc.w4_dequant(packed2, s2, qz2, 32)))
