"""GPU parity: grouped/batched hgemm in one persistent launch and its hipGraph.

Each problem's result is bit-identical to the single-problem hgemm with the
same tile and split-K (same kernel body, same summation order), and within the
reference's hgemm criterion of the oracle.  Graph replays, interleaved with
eager work, reproduce the eager result every time.
"""
import numpy as np
import pytest
import torch

from libthom_amd import HgemmConfig, routines as R

pytestmark = pytest.mark.gpu


def make_problems(dims, dev, seed, splits=None):
    g = torch.Generator(device="cpu").manual_seed(seed)
    probs = []
    for i, (M, N, K) in enumerate(dims):
        A = (torch.rand(M, K, generator=g) * 2 - 1).half().to(dev)
        B = (torch.rand(K, N, generator=g) * 2 - 1).half().to(dev)
        C = torch.zeros(M, N, dtype=torch.float16, device=dev)
        probs.append((A, B, C, splits[i] if splits else 1))
    return probs


def single(A, B, cfg, split):
    c = HgemmConfig(cfg.bm, cfg.bn, cfg.bk, split, cfg.stages)
    return R.hgemm(A, B, config=c)


@pytest.mark.parametrize("tile", [(128, 128, 64, 2), (64, 64, 64, 2), (256, 256, 64, 2), (128, 128, 32, 4),
                                  (256, 256, 32, 4), (256, 128, 64, 3)])
def test_grouped_matches_single(dev, orc, tile):
    dims = [(64, 64, 64), (100, 200, 72), (1000, 48, 512), (48, 1000, 4096), (512, 512, 512), (16, 16, 16),
            (257, 136, 1000), (1024, 1024, 256), (80, 96, 2048), (0, 64, 64)]
    splits = [1, 1, 2, 8, 1, 1, 3, 1, 4, 1]
    probs = make_problems(dims, dev, 1, splits)
    bm, bn, bk, st = tile
    cfg = HgemmConfig(bm, bn, bk, 1, st)
    plan = R.GroupedHgemm(probs, cfg)
    plan.run()
    torch.cuda.synchronize()
    for (A, B, C, sp), (M, N, K) in zip(probs, dims):
        if M == 0:
            continue
        assert torch.equal(C, single(A, B, cfg, sp)), (M, N, K, sp)
    A, B, C, _ = probs[1]
    ref = orc.hgemm(A.cpu().numpy(), B.cpu().numpy())
    got = C.cpu().numpy().astype(np.float32)
    diff = np.abs(got - ref.astype(np.float32))
    assert np.all((diff <= 1e-3) | (diff <= 1e-2 * np.maximum(np.abs(ref.astype(np.float32)), 1e-3) +
                                     1e-2 * (np.abs(A.cpu().numpy().astype(np.float32)) @
                                             np.abs(B.cpu().numpy().astype(np.float32)))))
    plan.close()


@pytest.mark.parametrize("stages", [2, 4])
def test_grouped_graph_replays(dev, stages):
    dims = [(128, 256, 512), (64, 64, 4096), (300, 104, 200), (512, 64, 1024)]
    probs = make_problems(dims, dev, 2, [1, 4, 1, 2])
    plan = R.GroupedHgemm(probs, HgemmConfig(128, 128, 64 if stages == 2 else 32, 1, stages))
    side = torch.cuda.Stream(device=dev)
    plan.run(side)
    side.synchronize()
    eager = [C.clone() for _, _, C, _ in probs]
    plan.capture(side)
    for rep in range(4):
        with torch.cuda.stream(side):
            for _, _, C, _ in probs:
                C.zero_()                        # fill kernel, not a memset node
            x = torch.randn(1000, 1000, device=dev) @ torch.randn(1000, 1000, device=dev)   # unrelated eager work
        plan.replay(side)
        side.synchronize()
        for (_, _, C, _), e in zip(probs, eager):
            assert torch.equal(C, e), rep
        del x
    plan.close()


def test_batched_configuration(dev):
    dims = R.batched_problem_dims()
    assert len(dims) == 256 and all(64 <= d <= 4096 and d % 16 == 0 for p in dims for d in p)
    from libthom_amd import tuner
    cfg, policy = tuner.best_batched(len(dims))
    probs, _ = tuner.batched_problems(dims, dev, policy, cfg)
    plan = R.GroupedHgemm(probs, cfg)
    plan.run()
    torch.cuda.synchronize()
    rng = np.random.default_rng(0)
    for i in rng.choice(256, 12, replace=False):
        A, B, C, split = probs[i]
        assert torch.equal(C, single(A, B, cfg, split)), dims[i]
    plan.close()


def _padding(M, N, bm, bn):
    return 1.0 - (M * N) / (((M + bm - 1) // bm) * bm * ((N + bn - 1) // bn) * bn)


def test_batched_configuration_vs_oracle(dev, orc):
    """The configured batched launch (256 log-uniform problems, the tuned tile
    and split policy) against the oracle (hgemm.cu:137-152) with the reference's
    criterion (hgemm.cu:165-168), on sampled rows of at least 16 problems: the
    largest K, every split-K problem, the six most edge-padded, the largest and
    smallest by flops, and random ones."""
    from libthom_amd import tuner
    from test_gpu_gemm import hgemm_check
    dims = R.batched_problem_dims()
    cfg, policy = tuner.best_batched(len(dims))
    probs, _ = tuner.batched_problems(dims, dev, policy, cfg)
    plan = R.GroupedHgemm(probs, cfg)
    plan.run()
    torch.cuda.synchronize()
    flops = [M * N * K for M, N, K in dims]
    pick = {int(np.argmax([K for _, _, K in dims])), int(np.argmax(flops)), int(np.argmin(flops))}
    pick |= {i for i, p in enumerate(probs) if p[3] > 1}
    pick |= set(int(i) for i in np.argsort([-_padding(M, N, cfg.bm, cfg.bn) for M, N, _ in dims])[:6])
    rng = np.random.default_rng(3)
    while len(pick) < 20:
        pick.add(int(rng.integers(0, len(dims))))
    for i in sorted(pick):
        A, B, C, split = probs[i]
        M, N, K = dims[i]
        rows = np.unique(np.concatenate([[0, M - 1], rng.integers(0, M, 14)]))
        An = A.cpu().numpy()[rows]
        Bn = B.cpu().numpy()
        ref = orc.hgemm(An, Bn)
        print(f"batched problem {i}: {M}x{N}x{K} split {split}", flush=True)
        hgemm_check(C.cpu().numpy()[rows], ref, An, Bn)
    plan.close()
