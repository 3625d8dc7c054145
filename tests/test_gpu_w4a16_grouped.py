"""GPU parity: grouped w4a16 (many problems in one persistent launch).

Each problem's output is checked against the oracle with the reference's
criterion (w4a16.cpp:194-197): |r - ref| <= 1e-2 |ref| + 1e-3 RMS(C_ref),
against the oracle's FP32 k-ascending sum over the oracle-dequantised
weights; a grouped tile is also bit-identical to the same tile run alone.
"""
import ctypes

import numpy as np
import pytest
import torch

from libthom_amd import W4A16Config, inputs, routines as R
from libthom_amd._lib import LthipError, call

pytestmark = pytest.mark.gpu


def t(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def i32(a, dev):
    return t(np.ascontiguousarray(a).view(np.int32), dev)


def w4_close(got, ref):
    got = got.astype(np.float32)
    ref = ref.astype(np.float32)
    rms = float(np.sqrt(np.mean(ref.astype(np.float64) ** 2)))
    bad = np.abs(got - ref) > 1e-2 * np.abs(ref) + 1e-3 * rms
    assert not bad.any(), f"{bad.sum()} of {bad.size} outside 1% + 1e-3 RMS; max abs err {np.abs(got - ref).max()}"


def make(dims, dev, seed=0):
    """dims: (M, N, K, group, asymmetric) -> host arrays and device problems"""
    host, probs = [], []
    for i, (M, N, K, group, asym) in enumerate(dims):
        A, q, s, z = inputs.w4a16_problem(M, N, K, group, seed=1000 + 31 * seed + i)
        packed = inputs.pack_w4(q)
        qz = inputs.pack_zeros(z) if asym else None
        host.append((A, packed, s, qz, group))
        C = torch.full((M, N), float("nan"), dtype=torch.float16, device=dev)
        probs.append((t(A, dev), i32(packed, dev), t(s, dev), None if qz is None else i32(qz, dev), C, group))
    return host, probs


def check_all(orc, host, probs):
    for (A, packed, s, qz, group), prob in zip(host, probs):
        if A.shape[0] == 0 or packed.shape[1] == 0:
            continue
        ref = orc.w4a16_rows(A, orc.w4_dequant(packed, s, qz, group))
        w4_close(prob[4].cpu().numpy(), ref)


MIXED = [(1, 544, 1024, 128, True), (5, 256, 512, 64, False), (16, 1056, 256, 128, True),
         (64, 96, 2048, 256, True), (65, 544, 512, 128, False), (100, 288, 1024, 64, True),
         (128, 256, 128, 128, True), (129, 1024, 512, 128, True), (300, 544, 256, 64, False),
         (513, 320, 1024, 128, True)]


def test_grouped_mixed_vs_oracle(dev, orc):
    host, probs = make(MIXED, dev)
    plan = R.GroupedW4A16(probs)
    info = plan.info()
    # tiles of 256 columns; 256, 128 or 64 rows by M
    rows = [256 if M > 128 else (128 if M > 64 else 64) for M, *_ in MIXED]
    assert info["work_items"] == sum(-(-M // r) * -(-N // 256) for (M, N, *_), r in zip(MIXED, rows))
    plan.run()
    torch.cuda.synchronize()
    check_all(orc, host, probs)
    plan.close()


@pytest.mark.parametrize("tile_m", [256, 128, 64])
def test_grouped_forced_tile(dev, orc, tile_m):
    dims = [(1, 544, 1024, 128, True), (70, 256, 512, 64, False), (300, 544, 256, 128, True)]
    host, probs = make(dims, dev, seed=tile_m)
    plan = R.GroupedW4A16(probs, W4A16Config(3, 256, tile_m, 1))
    assert plan.info()["work_items"] == sum(-(-M // tile_m) * -(-N // 256) for M, N, *_ in dims)
    plan.run()
    torch.cuda.synchronize()
    check_all(orc, host, probs)
    plan.close()


def single_kind3(A, W, S, Z, C, group, tile_m, dev):
    # one problem through the single entry point, whole tiles (no workspace,
    # so no tail split): the grouped launch runs the same tile code
    M, K = A.shape
    N = W.shape[1]
    cfg = W4A16Config(3, 256, tile_m, 1)
    call("lthip_w4a16_gemm", A.data_ptr(), W.data_ptr(), S.data_ptr(), None if Z is None else Z.data_ptr(),
         C.data_ptr(), M, N, K, group, ctypes.byref(cfg), None, 0, torch.cuda.current_stream(dev).cuda_stream)


def test_grouped_bit_identical_to_single(dev):
    dims = [(33, 544, 1024, 128, True), (90, 256, 512, 64, False), (400, 288, 512, 128, True)]
    _, probs = make(dims, dev, seed=5)
    plan = R.GroupedW4A16(probs)
    plan.run()
    torch.cuda.synchronize()
    for (A, W, S, Z, C, group), (M, *_) in zip(probs, dims):
        C1 = torch.empty_like(C)
        single_kind3(A, W, S, Z, C1, group, 256 if M > 128 else (128 if M > 64 else 64), dev)
        torch.cuda.synchronize()
        assert torch.equal(C, C1), M
    plan.close()


def test_grouped_configured_shapes(dev, orc):
    # the four configured activations (M = 1, 16, 128, 2048), each with its
    # own 11008 x 4096 weights, in one launch
    dims = [(M, 11008, 4096, 128, True) for M in (1, 16, 128, 2048)]
    host, probs = make(dims, dev, seed=9)
    plan = R.GroupedW4A16(probs)
    plan.run()
    torch.cuda.synchronize()
    for (A, packed, s, qz, group), prob in zip(host, probs):
        M = A.shape[0]
        rows = M if M <= 128 else 64
        row0 = 0 if M <= 128 else M - 64
        ref = orc.w4a16_rows(A, orc.w4_dequant(packed, s, qz, group), row0, rows)
        w4_close(prob[4].cpu().numpy()[row0:row0 + rows], ref)
    plan.close()


def test_grouped_graph_replays(dev):
    _, probs = make(MIXED[:6], dev, seed=3)
    plan = R.GroupedW4A16(probs)
    side = torch.cuda.Stream(device=dev)
    plan.run(side)
    side.synchronize()
    eager = [p[4].clone() for p in probs]
    plan.capture(side)
    for rep in range(4):
        with torch.cuda.stream(side):
            for p in probs:
                p[4].zero_()                     # fill kernel, not a memset node
            x = torch.randn(1000, 1000, device=dev) @ torch.randn(1000, 1000, device=dev)   # unrelated eager work
        plan.replay(side)
        side.synchronize()
        for p, e in zip(probs, eager):
            assert torch.equal(p[4], e), rep
        del x
    # and an eager relaunch after the replays (the queue was left at zero)
    plan.run(side)
    side.synchronize()
    for p, e in zip(probs, eager):
        assert torch.equal(p[4], e)
    plan.close()


def test_grouped_empty_problems(dev, orc):
    host, probs = make([(0, 256, 512, 128, True), (3, 256, 512, 128, True)], dev, seed=4)
    plan = R.GroupedW4A16(probs)
    assert plan.info()["work_items"] == 1
    plan.run()
    torch.cuda.synchronize()
    check_all(orc, host[1:], probs[1:])
    plan.close()


def test_grouped_rejects(dev):
    _, probs = make([(8, 544, 512, 128, True)], dev, seed=6)
    A, W, S, Z, C, group = probs[0]
    # group not a multiple of 64; K not a multiple of group (512 % 192)
    for g in (32, 192):
        with pytest.raises(LthipError):
            R.GroupedW4A16([(A, W, S, Z, C, g)])
    # N % 32 != 0
    _, p2 = make([(8, 520, 512, 128, True)], dev, seed=7)
    with pytest.raises(LthipError):
        R.GroupedW4A16(p2)
    # misaligned activations (2-byte offset)
    Abig = torch.empty(A.numel() + 8, dtype=torch.float16, device=dev)
    Amis = Abig[1:1 + A.numel()].view(A.shape)
    with pytest.raises(LthipError):
        R.GroupedW4A16([(Amis, W, S, Z, C, group)])
    # configs the grouped launch does not take
    for cfg in (W4A16Config(2, 128, 128, 1), W4A16Config(3, 128, 128, 1), W4A16Config(3, 256, 32, 1),
                W4A16Config(3, 256, 128, 2)):
        with pytest.raises(LthipError):
            R.GroupedW4A16(probs, cfg)


def test_grouped_many_small_problems(dev, orc):
    # 300 problems (the table's binary search over many entries, items of all
    # three tile classes interleaved), each checked against the oracle
    rng = np.random.default_rng(12)
    dims = [(int(rng.integers(1, 300)), int(rng.choice([32, 96, 256, 288])), int(rng.choice([64, 128, 256])), 64,
             bool(rng.integers(0, 2))) for _ in range(300)]
    host, probs = make(dims, dev, seed=13)
    plan = R.GroupedW4A16(probs)
    rows = [256 if M > 128 else (128 if M > 64 else 64) for M, *_ in dims]
    assert plan.info()["work_items"] == sum(-(-M // r) * -(-N // 256) for (M, N, *_), r in zip(dims, rows))
    plan.run()
    torch.cuda.synchronize()
    check_all(orc, host, probs)
    plan.close()
