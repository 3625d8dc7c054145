"""Bit-exact w4a16 on inputs whose FP32 sums are exact (tests/_w4_exact.py).

The reference's weight is fp16((q - z) s), rounded once (quantize.cpp:58-65),
and its product the FP32 sum over those weights (w4a16.cpp:178-184).  With
A in {-2..2}, K <= 64 and non-power-of-two scales every FP32 sum is exact, so
every kind, tile and the grouped launch must equal the oracle bit for bit; the
scale-after-the-sum shortcut differs from it on these inputs
(tests/test_oracle_w4.py::test_exact_inputs_expose_shortcut).
"""
import numpy as np
import pytest
import torch

from libthom_amd import W4A16Config, routines as R
from _w4_exact import exact_problem

pytestmark = pytest.mark.gpu

N = 288   # ragged for 64-, 32- and 128-column blocks; N % 32 == 0 for kind 3

CASES = (
    [(W4A16Config(4, tn, 0, 1), M, 2048 // tn) for tn in (64, 32) for M in (1, 2, 3, 4)] +
    [(W4A16Config(5, tn, 0, 1), M, g) for tn in (64, 32) for M in (1, 2, 16, 17, 40, 64) for g in (32, 64)] +
    [(W4A16Config(1, 64, 16, sp), M, 32) for sp in (1, 2) for M in (1, 16, 100)] +
    [(W4A16Config(2, bn, bm, 1), M, 32) for bm, bn in ((64, 64), (128, 128), (256, 256)) for M in (1, 70, 300)] +
    [(W4A16Config(3, bn, bm, 1), M, 64) for bm, bn in ((128, 128), (256, 256), (64, 256), (64, 128), (64, 64)) for M in (1, 70, 300)] +
    [(W4A16Config(6, 128, 128, 1), M, 64) for M in (1, 70, 128, 300)] +
    [(None, M, 32) for M in (1, 2, 3, 16, 64, 65, 200)] + [(None, M, 64) for M in (1, 16, 128, 300)]
)


def _dev(a, dev, dt=None):
    a = np.ascontiguousarray(a)
    return torch.from_numpy(a.view(dt) if dt is not None else a).to(dev)


@pytest.mark.parametrize("cfg,M,group", CASES, ids=lambda v: str(v) if not hasattr(v, "kind") else
                         f"k{v.kind}_{v.tile_m}x{v.tile_n}_s{v.split_k}")
@pytest.mark.parametrize("zeros", [True, False])
def test_w4a16_bit_exact(dev, orc, cfg, M, group, zeros):
    K = 64
    A, q, z, s = exact_problem(M, N, K, group, seed=1000 + 7 * M + group)
    packed = orc.w4_pack(q)
    qz = orc.w4_pack_zeros(z) if zeros else None
    want = orc.w4a16_rows(A, orc.w4_dequant(packed, s, qz, group))
    C = R.w4a16(_dev(A, dev), _dev(packed, dev, np.int32), _dev(s, dev),
                None if qz is None else _dev(qz, dev, np.int32), group, config=cfg).cpu().numpy()
    diff = int((C.view(np.uint16) != want.view(np.uint16)).sum())
    assert diff == 0, f"{diff} of {C.size} outputs differ from the oracle"


def test_w4a16_grouped_bit_exact(dev, orc):
    # one launch: 64-, 128- and 256-row tiles, each problem against the oracle
    host, probs = [], []
    for i, M in enumerate((1, 16, 100, 300)):
        A, q, z, s = exact_problem(M, N, 64, 64, seed=2000 + i)
        packed, qz = orc.w4_pack(q), orc.w4_pack_zeros(z)
        host.append(orc.w4a16_rows(A, orc.w4_dequant(packed, s, qz, 64)))
        probs.append((_dev(A, dev), _dev(packed, dev, np.int32), _dev(s, dev), _dev(qz, dev, np.int32),
                      torch.empty(M, N, dtype=torch.float16, device=dev), 64))
    plan = R.GroupedW4A16(probs)
    plan.run()
    torch.cuda.synchronize()
    for want, prob in zip(host, probs):
        got = prob[4].cpu().numpy()
        assert np.array_equal(got.view(np.uint16), want.view(np.uint16)), prob[0].shape
    plan.close()
