"""GPU parity of the cost-model kernels against the oracle, and training.

Forward and gradients (FP32 on the GPU) match the oracle's double-precision
restatement, whose gradients are themselves pinned by finite differences
(test_oracle_routines.py); full-batch Adam on the GPU fits a known cost surface.
"""
import numpy as np
import pytest
import torch

from libthom_amd import tuner
from libthom_amd._lib import call, lib

pytestmark = pytest.mark.gpu


def f32(a, dev):
    return torch.tensor(np.ascontiguousarray(a), dtype=torch.float32, device=dev)


def test_forward_and_gradients_vs_oracle(dev, orc):
    rng = np.random.default_rng(11)
    B, I, H, O = 301, 4, 64, 14
    x = rng.standard_normal((B, I))
    W1 = rng.standard_normal((H, I)) * 0.7
    b1 = rng.standard_normal(H) * 0.1
    W2 = rng.standard_normal((O, H)) * 0.2
    b2 = rng.standard_normal(O) * 0.1
    t = rng.standard_normal((B, O))
    mask = (rng.uniform(size=(B, O)) > 0.25).astype(np.uint8)
    hid_ref, y_ref = orc.mlp_forward(x, W1, b1, W2, b2)
    loss_ref, g_ref = orc.mlp_loss_grad(x, W1, b1, W2, b2, t, mask)

    X, P1, Q1, P2, Q2, T = (f32(a, dev) for a in (x, W1, b1, W2, b2, t))
    M = torch.tensor(mask, device=dev)
    hid = torch.empty(B, H, device=dev)
    y = torch.empty(B, O, device=dev)
    s = torch.cuda.current_stream().cuda_stream
    call("lthip_mlp_forward", X.data_ptr(), P1.data_ptr(), Q1.data_ptr(), P2.data_ptr(), Q2.data_ptr(),
         hid.data_ptr(), y.data_ptr(), B, I, H, O, s)
    assert np.allclose(y.cpu().numpy(), y_ref, rtol=1e-4, atol=1e-5)
    g = [torch.empty_like(p) for p in (P1, Q1, P2, Q2)]
    loss = torch.zeros(1, device=dev)
    need = lib().lthip_mlp_workspace_size(B, H, O)
    ws = torch.empty(need, dtype=torch.uint8, device=dev)
    call("lthip_mlp_loss_grad", X.data_ptr(), P1.data_ptr(), Q1.data_ptr(), P2.data_ptr(), Q2.data_ptr(),
         T.data_ptr(), M.data_ptr(), B, I, H, O, *[a.data_ptr() for a in g], loss.data_ptr(), ws.data_ptr(),
         ws.numel(), s)
    assert loss.item() == pytest.approx(loss_ref, rel=1e-4)
    for got, want in zip(g, g_ref):
        scale = np.abs(want).max()
        assert np.max(np.abs(got.cpu().numpy() - want)) <= 1e-4 * scale + 1e-6


def test_training_fits_a_known_cost_surface(dev):
    # synthetic "timings": log t_c = a_c . f(shape) + b_c, a smooth surface the
    # 2-layer network must fit; the argmin of the prediction must match the truth
    rng = np.random.default_rng(3)
    shapes = tuner.random_shapes(600, seed=9)
    A = rng.standard_normal((len(tuner.W4A16_SPACE), 4)) * 0.5
    bias = rng.standard_normal(len(tuner.W4A16_SPACE))
    db = {}
    for s in shapes:
        f = tuner.features(*s) / 10.0
        db[s] = {cfg: float(np.exp(A[i] @ f + bias[i])) for i, cfg in enumerate(tuner.W4A16_SPACE)
                 if tuner.valid(cfg, s[0], s[2], s[1], s[3])}
    model = tuner.CostModel(device="cuda:0")
    hist = model.fit(db, shapes[:500], steps=3000, log_every=500)
    assert hist[-1] < 0.02 * hist[0]
    res = tuner.evaluate(db, shapes[500:], model.choose(shapes[500:]))
    assert res["mean_regret"] < 0.05, res
    # the GPU forward and the numpy forward agree
    assert np.allclose(model.predict_logtime(shapes[:20]), model.predict_logtime_host(shapes[:20]), atol=1e-4)
