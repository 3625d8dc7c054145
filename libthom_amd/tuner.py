"""Autotuner and learned cost model for w4a16 (and the batched hgemm tile).

Exhaustive tuner: times every configuration of the search space on a shape
(median of launches, caches flushed between launches) and appends the timings
to the tuning database (CSV, the columns the CLI's `libthom tune w4a16`
writes).  Learned tuner: a 2-layer MLP (4 -> 64 -> one output per
configuration) maps (log2 M, log2 N, log2 K, log2 group), standardised, to the
standardised log-time of every configuration; the argmin over the valid
configurations is the choice.  It is trained on the GPU with the library's own
forward/backward/Adam kernels (lthip_mlp_*), from the database of random-shape
sweeps.  Shapes outside the training range fall back to the exhaustive tuner.

The reference's tuner is the exhaustive search of hgemm.cu:258-323 whose
winners were copied by hand into a size-keyed table (hgemm_better_reuse.cu:
357-376); "the method without the network" here is the library's built-in
size-based choice (lthip_w4a16_gemm with a null config).

Batched hgemm: one tile shape (tile, LDS stages) for the whole batch plus a
split-K policy per problem, chosen by timing every candidate on the batch
(graph replays) -> tuning/batched.csv, read back by best_batched() and by the
CLI's hgemm_batched routine.

Usage (on a GPU box):
  python -m libthom_amd.tuner batched --count 256
  python -m libthom_amd.tuner sweep --shapes 2000 --out libthom_amd/tuning/w4a16_sweep.csv
  python -m libthom_amd.tuner train --db libthom_amd/tuning/w4a16_sweep.csv --out libthom_amd/tuning/w4a16_mlp.json
  python -m libthom_amd.tuner eval --db ... --model ...
"""
from __future__ import annotations

import argparse
import csv
import ctypes
import json
import math
import os
import sys
import time
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from ._lib import W4A16Config

HERE = os.path.dirname(os.path.abspath(__file__))
TUNING_DIR = os.path.join(HERE, "tuning")
DB_COLUMNS = ["M", "N", "K", "group", "kind", "tile_n", "tile_m", "split_k", "ms"]

# the search space: index i is output i of the cost model
# (kind, tile_n, tile_m, split_k); kind 3 with split_k 1 lets the library
# split K for the last partial round of tiles
W4A16_SPACE: List[Tuple[int, int, int, int]] = (
    [(1, 64, 16, s) for s in (1, 2, 3, 4, 6, 8, 12, 16)] +
    [(2, 64, 64, s) for s in (1, 2, 4, 8)] + [(2, 128, 64, s) for s in (1, 2, 4, 8)] +
    [(2, 128, 128, s) for s in (1, 2, 4)] +
    [(2, bn, bm, 1) for bm, bn in ((128, 256), (256, 128), (256, 256))] +
    [(3, 128, 128, 1), (3, 128, 128, 2), (3, 256, 128, 1), (3, 128, 256, 1), (3, 256, 256, 1)] +
    [(4, 64, 0, 1), (4, 32, 0, 1), (5, 64, 0, 1)] +
    [(3, 128, 64, 1), (3, 128, 64, 2), (3, 64, 64, 1), (5, 32, 0, 1)] +
    [(6, 128, 128, 1)])


def valid(cfg: Tuple[int, int, int, int], M: int, K: int, N: Optional[int] = None,
          group: Optional[int] = None) -> bool:
    """Whether the library runs cfg on the shape; kind 3 needs N % 32 == 0 and
    group % 64 == 0, kind 4 M <= 4, group % (2048 / tile_n) == 0 and M K <=
    32768, kind 5 M <= 64 and group % 32 == 0, kind 6 kind 3's conditions
    and at most 16384 tiles (kinds 3-6: false when N or group is not given)."""
    kind, tile_n, _, split = cfg
    if kind == 1:
        return M <= 128 and split <= max(1, K // 32)
    if kind == 4:
        return N is not None and group is not None and M <= 4 and group % (2048 // tile_n) == 0 and \
            M * K <= 32768 and (K // 8) * N * 4 < 2 ** 31 and split == 1
    if kind == 5:
        return N is not None and group is not None and M <= 64 and N >= 8 and group % 32 == 0 and \
            group <= 256 and (K // 8) * N * 4 < 2 ** 31 and split == 1
    if kind == 3:
        return N is not None and group is not None and N % 32 == 0 and group % 64 == 0 and \
            split <= max(1, K // 256)
    if kind == 6:
        return N is not None and group is not None and N % 32 == 0 and group % 64 == 0 and split == 1 and \
            ((M + 127) // 128) * ((N + 127) // 128) <= 16384
    return split <= max(1, K // 256)


def features(M, N, K, group) -> np.ndarray:
    return np.log2(np.array([M, N, K, group], dtype=np.float64))


# ---------------------------------------------------------------------------
# database
def read_db(path: str) -> Dict[Tuple[int, int, int, int], Dict[Tuple[int, int, int, int], float]]:
    """{(M, N, K, group): {config: best ms}}"""
    out: Dict = {}
    with open(path) as f:
        for row in csv.DictReader(f):
            shape = (int(row["M"]), int(row["N"]), int(row["K"]), int(row["group"]))
            cfg = (int(row["kind"]), int(row["tile_n"]), int(row["tile_m"]), int(row["split_k"]))
            ms = float(row["ms"])
            d = out.setdefault(shape, {})
            d[cfg] = min(ms, d.get(cfg, math.inf))
    return out


def append_db(path: str, shape, cfg, ms: float) -> None:
    new = not os.path.exists(path)
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "a") as f:
        if new:
            f.write(",".join(DB_COLUMNS) + "\n")
        f.write(",".join(str(v) for v in (*shape, *cfg)) + f",{ms:.6f}\n")


def random_shapes(n: int, seed: int = 4321):
    """M log-uniform in [1, 4096]; N a multiple of 64 log-uniform in [512, 16384];
    K a multiple of 128 log-uniform in [512, 8192]; group in {32, 64, 128}."""
    rng = np.random.default_rng(seed)
    shapes = []
    for _ in range(n):
        M = int(round(2 ** rng.uniform(0, 12)))
        N = int(round(2 ** rng.uniform(9, 14) / 64)) * 64
        K = int(round(2 ** rng.uniform(9, 13) / 128)) * 128
        G = int(rng.choice([32, 64, 128]))
        shapes.append((max(M, 1), N, K, G))
    return shapes


# ---------------------------------------------------------------------------
# exhaustive tuner (GPU)
class Sweeper:
    """Times w4a16 configurations on random data of any shape up to the maxima."""

    def __init__(self, max_M=4096, max_N=16384, max_K=8192, flush=True):
        import torch
        from . import routines as R
        self.torch, self.R = torch, R
        dev = torch.device("cuda:0")
        self.dev = dev
        g = torch.Generator(device="cpu").manual_seed(7)
        self.A = (torch.randn(max_M * max_K, generator=g) * 0.5).half().to(dev)
        self.W = torch.randint(-2 ** 31, 2 ** 31 - 1, ((max_K // 8) * max_N,), dtype=torch.int32, generator=g).to(dev)
        self.S = (torch.rand((max_K // 32) * max_N, generator=g) * 0.015 + 0.005).half().to(dev)
        self.Z = torch.randint(-2 ** 31, 2 ** 31 - 1, ((max_K // 32) * (max_N // 8),), dtype=torch.int32,
                               generator=g).to(dev)
        self.C = torch.empty(max_M * max_N, dtype=torch.float16, device=dev)
        self.flush_buf = torch.empty(128 * 1024 * 1024, dtype=torch.float32, device=dev) if flush else None
        self.flush_src = torch.empty_like(self.flush_buf) if flush else None

    def views(self, M, N, K, group):
        return (self.A[:M * K].view(M, K), self.W[:(K // 8) * N].view(K // 8, N),
                self.S[:(K // group) * N].view(K // group, N), self.Z[:(K // group) * (N // 8)].view(K // group, N // 8),
                self.C[:M * N].view(M, N))

    def time(self, shape, cfg: Optional[Tuple[int, int, int, int]], runs=3):
        torch, R = self.torch, self.R
        M, N, K, group = shape
        A, W, S, Z, C = self.views(M, N, K, group)
        config = W4A16Config(*cfg) if cfg is not None else None
        R.w4a16(A, W, S, Z, group, C, config)   # warm-up (and first-use workspace)
        ts = []
        for _ in range(runs):
            if self.flush_buf is not None:
                R.copy_f32(self.flush_buf, self.flush_src)
            s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            s.record()
            R.w4a16(A, W, S, Z, group, C, config)
            e.record()
            e.synchronize()
            ts.append(s.elapsed_time(e))
        return float(np.median(ts))

    def exhaustive(self, shape, runs=3, db: Optional[str] = None):
        M, N, K, group = shape
        best, best_ms = None, math.inf
        for cfg in W4A16_SPACE:
            if not valid(cfg, M, K, N, group):
                continue
            ms = self.time(shape, cfg, runs)
            if db:
                append_db(db, shape, cfg, ms)
            if ms < best_ms:
                best, best_ms = cfg, ms
        return best, best_ms


# ---------------------------------------------------------------------------
# learned cost model
class CostModel:
    """2-layer MLP, parameters on the GPU, trained with the lthip_mlp_* kernels."""

    def __init__(self, n_in=4, hidden=64, n_out=len(W4A16_SPACE), seed=0, device="cuda:0"):
        import torch
        self.torch = torch
        self.dev = torch.device(device)
        rng = np.random.default_rng(seed)
        self.I, self.H, self.O = n_in, hidden, n_out
        self.params = {
            "W1": rng.standard_normal((hidden, n_in)) * math.sqrt(2.0 / n_in),
            "b1": np.zeros(hidden),
            "W2": rng.standard_normal((n_out, hidden)) * math.sqrt(1.0 / hidden),
            "b2": np.zeros(n_out),
        }
        self.x_mean = np.zeros(n_in)
        self.x_std = np.ones(n_in)
        self.y_mean = 0.0
        self.y_std = 1.0
        self.lo = np.full(n_in, -np.inf)
        self.hi = np.full(n_in, np.inf)
        self.history: List[float] = []

    # -- data --
    def _xy(self, db, shapes):
        X = np.stack([features(*s) for s in shapes])
        Y = np.zeros((len(shapes), self.O))
        mask = np.zeros((len(shapes), self.O), dtype=np.uint8)
        for r, s in enumerate(shapes):
            for c, cfg in enumerate(W4A16_SPACE):
                ms = db[s].get(cfg)
                if ms is not None and valid(cfg, s[0], s[2], s[1], s[3]):
                    Y[r, c] = math.log(ms)
                    mask[r, c] = 1
        return X, Y, mask

    def fit(self, db, shapes, steps=4000, lr=3e-3, log_every=500):
        """Full-batch Adam on the GPU (lthip_mlp_loss_grad + lthip_adam_step)."""
        torch = self.torch
        from ._lib import call, lib
        X, Y, mask = self._xy(db, shapes)
        self.x_mean, self.x_std = X.mean(0), X.std(0) + 1e-9
        self.lo, self.hi = X.min(0), X.max(0)
        vals = Y[mask.astype(bool)]
        self.y_mean, self.y_std = float(vals.mean()), float(vals.std() + 1e-9)
        Xn = (X - self.x_mean) / self.x_std
        Yn = (Y - self.y_mean) / self.y_std
        dev = self.dev
        f32 = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float32, device=dev)
        xs, ts = f32(Xn), f32(Yn)
        ms = torch.tensor(mask, dtype=torch.uint8, device=dev)
        P = {k: f32(v) for k, v in self.params.items()}
        G = {k: torch.zeros_like(v) for k, v in P.items()}
        Mo = {k: torch.zeros_like(v) for k, v in P.items()}
        Ve = {k: torch.zeros_like(v) for k, v in P.items()}
        loss = torch.zeros(1, dtype=torch.float32, device=dev)
        B = xs.shape[0]
        need = lib().lthip_mlp_workspace_size(B, self.H, self.O)
        ws = torch.empty(need, dtype=torch.uint8, device=dev)
        stream = torch.cuda.current_stream(dev).cuda_stream
        self.history = []
        for step in range(1, steps + 1):
            call("lthip_mlp_loss_grad", xs.data_ptr(), P["W1"].data_ptr(), P["b1"].data_ptr(), P["W2"].data_ptr(),
                 P["b2"].data_ptr(), ts.data_ptr(), ms.data_ptr(), B, self.I, self.H, self.O, G["W1"].data_ptr(),
                 G["b1"].data_ptr(), G["W2"].data_ptr(), G["b2"].data_ptr(), loss.data_ptr(), ws.data_ptr(),
                 ws.numel(), stream)
            rate = lr if step < steps * 0.7 else lr * 0.1
            for k in P:
                call("lthip_adam_step", P[k].data_ptr(), G[k].data_ptr(), Mo[k].data_ptr(), Ve[k].data_ptr(),
                     P[k].numel(), rate, 0.9, 0.999, 1e-8, step, stream)
            if step % log_every == 0 or step == 1:
                self.history.append(float(loss.item()))
        self.params = {k: v.double().cpu().numpy() for k, v in P.items()}
        return self.history

    # -- inference --
    def predict_logtime(self, shapes) -> np.ndarray:
        """Forward pass through lthip_mlp_forward on the GPU; log-ms per config."""
        torch = self.torch
        from ._lib import call
        X = (np.stack([features(*s) for s in shapes]) - self.x_mean) / self.x_std
        dev = self.dev
        f32 = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float32, device=dev)
        x = f32(X)
        P = {k: f32(v) for k, v in self.params.items()}
        hid = torch.empty(len(shapes), self.H, dtype=torch.float32, device=dev)
        y = torch.empty(len(shapes), self.O, dtype=torch.float32, device=dev)
        call("lthip_mlp_forward", x.data_ptr(), P["W1"].data_ptr(), P["b1"].data_ptr(), P["W2"].data_ptr(),
             P["b2"].data_ptr(), hid.data_ptr(), y.data_ptr(), len(shapes), self.I, self.H, self.O,
             torch.cuda.current_stream(dev).cuda_stream)
        return y.double().cpu().numpy() * self.y_std + self.y_mean

    def predict_logtime_host(self, shapes) -> np.ndarray:
        """The same forward pass in numpy (for machines without a GPU)."""
        X = (np.stack([features(*s) for s in shapes]) - self.x_mean) / self.x_std
        h = np.maximum(X @ self.params["W1"].T + self.params["b1"], 0.0)
        return (h @ self.params["W2"].T + self.params["b2"]) * self.y_std + self.y_mean

    def in_range(self, shape) -> bool:
        f = features(*shape)
        return bool(np.all(f >= self.lo - 1e-9) and np.all(f <= self.hi + 1e-9))

    def choose(self, shapes, logt=None):
        logt = self.predict_logtime(shapes) if logt is None else logt
        out = []
        for s, row in zip(shapes, logt):
            cands = [(row[i], cfg) for i, cfg in enumerate(W4A16_SPACE) if valid(cfg, s[0], s[2], s[1], s[3])]
            out.append(min(cands)[1])
        return out

    # -- persistence --
    def to_json(self) -> dict:
        return {"space": W4A16_SPACE, "I": self.I, "H": self.H, "O": self.O,
                "params": {k: v.tolist() for k, v in self.params.items()},
                "x_mean": self.x_mean.tolist(), "x_std": self.x_std.tolist(), "y_mean": self.y_mean,
                "y_std": self.y_std, "lo": self.lo.tolist(), "hi": self.hi.tolist(), "history": self.history}

    def save(self, path):
        with open(path, "w") as f:
            json.dump(self.to_json(), f)

    @classmethod
    def load(cls, path, device="cuda:0"):
        d = json.load(open(path))
        m = cls.__new__(cls)
        try:
            import torch
            m.torch = torch
            m.dev = torch.device(device)
        except Exception:
            m.torch, m.dev = None, None
        m.I, m.H, m.O = d["I"], d["H"], d["O"]
        m.params = {k: np.array(v) for k, v in d["params"].items()}
        m.x_mean, m.x_std = np.array(d["x_mean"]), np.array(d["x_std"])
        m.y_mean, m.y_std = d["y_mean"], d["y_std"]
        m.lo, m.hi = np.array(d["lo"]), np.array(d["hi"])
        m.history = d.get("history", [])
        assert [tuple(c) for c in d["space"]] == W4A16_SPACE
        return m


class LearnedTuner:
    """The model's choice inside its training range, the exhaustive tuner outside."""

    def __init__(self, model: CostModel, sweeper: Optional[Sweeper] = None, db: Optional[str] = None):
        self.model, self.sweeper, self.db = model, sweeper, db

    def config(self, shape):
        if self.model.in_range(shape):
            return self.model.choose([shape])[0], "model"
        if self.sweeper is None:
            raise RuntimeError(f"{shape} is outside the model's training range and no exhaustive tuner is attached")
        return self.sweeper.exhaustive(shape, db=self.db)[0], "exhaustive"


# ---------------------------------------------------------------------------
# evaluation: regret against the exhaustive optimum, gain over the default
def split_shapes(shapes, test_frac=0.2, seed=0):
    rng = np.random.default_rng(seed)
    idx = rng.permutation(len(shapes))
    n_test = int(len(shapes) * test_frac)
    test = [shapes[i] for i in idx[:n_test]]
    train = [shapes[i] for i in idx[n_test:]]
    return train, test


def evaluate(db, shapes, chosen, default_ms: Optional[Dict] = None):
    """Mean/median/p90 regret = t(chosen)/t(best) - 1 over shapes; and, when the
    default's timings are given, mean speed-up t(default)/t(chosen)."""
    regrets, gains = [], []
    for s, cfg in zip(shapes, chosen):
        times = db[s]
        best = min(times.values())
        t = times.get(tuple(cfg))
        if t is None:
            continue
        regrets.append(t / best - 1.0)
        if default_ms is not None and s in default_ms:
            gains.append(default_ms[s] / t)
    r = np.array(regrets)
    out = {"shapes": int(r.size), "mean_regret": float(r.mean()), "median_regret": float(np.median(r)),
           "p90_regret": float(np.percentile(r, 90)), "within_5pct": float((r <= 0.05).mean())}
    if gains:
        g = np.array(gains)
        out.update({"mean_speedup_vs_default": float(g.mean()), "geomean_speedup_vs_default":
                    float(np.exp(np.log(g).mean()))})
    return out


def read_default(path):
    out = {}
    if os.path.exists(path):
        with open(path) as f:
            for row in csv.DictReader(f):
                out[(int(row["M"]), int(row["N"]), int(row["K"]), int(row["group"]))] = float(row["ms"])
    return out


BATCHED_COLUMNS = ["count", "bm", "bn", "bk", "stages", "policy", "ms"]
# split-K policies for the batched hgemm: 0 none; 1 split problems whose tile
# grid is small next to their K (2 splits from K >= 1024, 4 from K >= 2048)
BATCHED_POLICIES = (0, 1)


def batched_split(policy: int, M: int, N: int, K: int, bm: int, bn: int) -> int:
    if policy == 0:
        return 1
    tiles = ((M + bm - 1) // bm) * ((N + bn - 1) // bn)
    if tiles < 16 and K >= 1024:
        return 4 if K >= 2048 else 2
    return 1


def best_batched(count: int = 256, path: Optional[str] = None):
    """(HgemmConfig, policy) of the fastest row for `count` problems in the
    batched database, or the built-in choice when there is none."""
    from ._lib import HgemmConfig
    path = path or os.path.join(TUNING_DIR, "batched.csv")
    best = None
    if os.path.exists(path):
        with open(path) as f:
            for r in csv.DictReader(f):
                if int(r["count"]) != count:
                    continue
                ms = float(r["ms"])
                if best is None or ms < best[0]:
                    best = (ms, int(r["bm"]), int(r["bn"]), int(r["bk"]), int(r["stages"]), int(r["policy"]))
    if best is None:
        return HgemmConfig(256, 256, 32, 1, 4), 0
    return HgemmConfig(best[1], best[2], best[3], 1, best[4]), best[5]


def batched_problems(dims, device, split_policy: int, cfg, tensors=None):
    """Problem list for GroupedHgemm on the seeded batched inputs (U[-1,1),
    stream ids S_BATCH_BASE + 2i and + 2i + 1); `tensors` reuses buffers."""
    import torch
    from . import inputs, routines as R
    if tensors is None:
        tensors = []
        for i, (M, N, K) in enumerate(dims):
            A = torch.empty(M, K, dtype=torch.float16, device=device)
            B = torch.empty(K, N, dtype=torch.float16, device=device)
            R.fill_uniform(A, -1, 1, 1234, inputs.S_BATCH_BASE + 2 * i)
            R.fill_uniform(B, -1, 1, 1234, inputs.S_BATCH_BASE + 2 * i + 1)
            tensors.append((A, B, torch.empty(M, N, dtype=torch.float16, device=device)))
    probs = [(A, B, C, batched_split(split_policy, M, N, K, cfg.bm, cfg.bn))
             for (A, B, C), (M, N, K) in zip(tensors, dims)]
    return probs, tensors


def time_batched(probs, cfg, device, replays: int = 10, warm_ms: float = 100.0) -> float:
    """Median ms per graph replay of the grouped launch, after at least
    warm_ms of untimed replays (under a new MFMA load the clocks dip for tens
    of milliseconds and recover, which would favour the later candidates)."""
    import torch
    from . import routines as R
    plan = R.GroupedHgemm(probs, cfg)
    side = torch.cuda.Stream(device=device)
    try:
        plan.run(side)
        plan.capture(side)
        times = []
        with torch.cuda.stream(side):
            spent = 0.0
            while spent < warm_ms:
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record(side)
                for _ in range(8):
                    plan.replay(side)
                e1.record(side)
                e1.synchronize()
                spent += max(e0.elapsed_time(e1), 0.01)
            for _ in range(replays):
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record(side)
                plan.replay(side)
                e1.record(side)
                e1.synchronize()
                times.append(e0.elapsed_time(e1))
        return float(np.median(times))
    finally:
        side.synchronize()
        plan.close()


def tune_batched(count: int = 256, out: Optional[str] = None, replays: int = 10, device: str = "cuda:0"):
    """Exhaustive search over every compiled tile x split policy; appends one
    row per candidate to the batched database and returns the best."""
    from . import routines as R
    out = out or os.path.join(TUNING_DIR, "batched.csv")
    dims = R.batched_problem_dims(count)
    tensors = None
    best = None
    new = not os.path.exists(out)
    with open(out, "a", newline="") as f:
        w = csv.writer(f)
        if new:
            w.writerow(BATCHED_COLUMNS)
        for cfg in R.hgemm_configs():
            for policy in BATCHED_POLICIES:
                probs, tensors = batched_problems(dims, device, policy, cfg, tensors)
                ms = time_batched(probs, cfg, device, replays)
                w.writerow([count, cfg.bm, cfg.bn, cfg.bk, cfg.stages, policy, f"{ms:.6f}"])
                f.flush()
                print(f"[batched] {cfg!r} policy {policy}: {ms:.3f} ms", flush=True)
                if best is None or ms < best[0]:
                    best = (ms, cfg, policy)
    return best


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m libthom_amd.tuner")
    sub = ap.add_subparsers(dest="cmd", required=True)
    sw = sub.add_parser("sweep")
    sw.add_argument("--shapes", type=int, default=2000)
    sw.add_argument("--seed", type=int, default=4321)
    sw.add_argument("--runs", type=int, default=3)
    sw.add_argument("--out", default=os.path.join(TUNING_DIR, "w4a16_sweep.csv"))
    sw.add_argument("--default-out", default=os.path.join(TUNING_DIR, "w4a16_default.csv"))
    sw.add_argument("--budget-s", type=float, default=480.0)
    tr = sub.add_parser("train")
    tr.add_argument("--db", default=os.path.join(TUNING_DIR, "w4a16_sweep.csv"))
    tr.add_argument("--out", default=os.path.join(TUNING_DIR, "w4a16_mlp.json"))
    tr.add_argument("--steps", type=int, default=4000)
    bt = sub.add_parser("batched")
    bt.add_argument("--count", type=int, default=256)
    bt.add_argument("--out", default=os.path.join(TUNING_DIR, "batched.csv"))
    bt.add_argument("--replays", type=int, default=10)
    ev = sub.add_parser("eval")
    ev.add_argument("--db", default=os.path.join(TUNING_DIR, "w4a16_sweep.csv"))
    ev.add_argument("--model", default=os.path.join(TUNING_DIR, "w4a16_mlp.json"))
    ev.add_argument("--default", default=os.path.join(TUNING_DIR, "w4a16_default.csv"))
    a = ap.parse_args(argv)

    if a.cmd == "batched":
        ms, cfg, policy = tune_batched(a.count, a.out, a.replays)
        print(f"[batched] best {cfg!r} policy {policy}: {ms:.3f} ms", flush=True)
        return 0
    if a.cmd == "sweep":
        sweeper = Sweeper()
        done = set(read_db(a.out).keys()) if os.path.exists(a.out) else set()
        t0 = time.time()
        shapes = random_shapes(a.shapes, a.seed)
        for i, s in enumerate(shapes):
            if s in done:
                continue
            if time.time() - t0 > a.budget_s:
                print(f"[sweep] budget reached after {i} shapes", flush=True)
                break
            sweeper.exhaustive(s, a.runs, a.out)
            append_db(a.default_out, s, (0, 0, 0, 0), sweeper.time(s, None, a.runs))
            if i % 50 == 0:
                print(f"[sweep] {i}/{len(shapes)} shapes, {time.time() - t0:.0f} s", flush=True)
        return 0
    if a.cmd == "train":
        db = read_db(a.db)
        shapes = sorted(db)
        train, test = split_shapes(shapes)
        model = CostModel()
        hist = model.fit(db, train, steps=a.steps)
        model.save(a.out)
        print(json.dumps({"train_shapes": len(train), "loss_history": hist}), flush=True)
        return 0
    if a.cmd == "eval":
        db = read_db(a.db)
        shapes = sorted(db)
        _, test = split_shapes(shapes)
        model = CostModel.load(a.model)
        chosen = model.choose(test)
        print(json.dumps(evaluate(db, test, chosen, read_default(a.default))), flush=True)
        return 0
    return 1


if __name__ == "__main__":
    sys.exit(main())
