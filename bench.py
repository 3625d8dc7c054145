"""Benchmark: % of HBM/MFMA roofline per routine on one MI355X.

The headline line measures the sgemv configuration (y = A x, A FP32 row-major
16384 x 8192, U[-1,1) from seed 1234): one step = one sgemv over the resident
matrix.  `value` is the achieved fraction of the 8 TB/s HBM peak, in percent;
`roofline` carries the absolute GB/s.  The `routines` object adds the other
routines at their configured shapes (reduction n=1e6 and 5e8, sdot, native
dgemm 4096^3, hgemm 4096^3, w4a16 M in {1,16,128,2048}).  `cpu_baseline`
times the oracle's sgemv on the host (all usable cores, up to 16, and one
thread), with the core count and CPU model.  `measured_peaks` holds
what this device delivers (MFMA probe loops, a read-only HBM stream and a
copy); every routine carries its fraction of the datasheet peak and
`frac_measured`, its fraction of the measured one, and `max_rel_err_vs_fp64`,
its largest error against an FP64 evaluation of the same inputs (divided by
the sum of absolute terms, |A||B| or |A||x|; reduction: by the sum).
`vendor_reference` times
torch.matmul (hipBLASLt / rocBLAS) on the GEMM shapes, for context only.
`sizes_sweeps` holds the reference's own benchmark sweeps (each routine's
SIZES through `libthom bench <routine> hip`, read back from its CSVs).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--quick]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK = 78.6      # TFLOP/s dense MFMA
FP16_PEAK = 2500.0    # TFLOP/s dense MFMA
FP8_PEAK = 5000.0
M_SG, N_SG = 16384, 8192
OZ_SLICES = (4, 6, 8, 10, 11, 12, 13)   # Ozaki slice counts in the emulated-vs-native study
SG_BYTES = 4 * (M_SG * N_SG + N_SG + M_SG)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _orc():
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle
    return _oracle


def dot2_err(got, A, B, rows, cols):
    """Largest |got_i - (A B)[rows_i, cols_i]| / (|A||B|)[rows_i, cols_i] over the
    sampled entries, the exact product taken from the oracle's compensated dot
    product (Dot2, twice the FP64 working precision) on the host.  A, B are
    host arrays (converted to float64), got the device results at the entries."""
    import numpy as np
    orc = _orc()
    r = np.unique(np.asarray(rows, np.int64))
    c = np.unique(np.asarray(cols, np.int64))
    Ar = np.asarray(A[r], dtype=np.float64)
    Bc = np.asarray(B[:, c], dtype=np.float64)
    ri = np.searchsorted(r, rows).astype(np.int32)
    ci = np.searchsorted(c, cols).astype(np.int32)
    hi, lo = orc.dgemm_dot2(Ar, Bc, ri, ci)
    scale = np.einsum("ik,ki->i", np.abs(Ar[ri]), np.abs(Bc[:, ci]))
    return float(np.max(np.abs(np.asarray(got, np.float64) - (hi + lo)) / scale))


def sample_entries(m, n, count, seed=0):
    import numpy as np
    rng = np.random.default_rng(seed)
    return rng.integers(0, m, count), rng.integers(0, n, count)


class Timer:
    """HIP events on torch's current stream (the stream the kernels run on)."""

    def __init__(self, torch):
        self.torch = torch

    def per_launch_ms(self, fn, iters, flush=None):
        t = self.torch
        if flush is None:
            s, e = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
            s.record()
            for _ in range(iters):
                fn()
            e.record()
            e.synchronize()
            return s.elapsed_time(e) / iters
        total = 0.0
        for _ in range(iters):
            flush()
            s, e = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
            s.record()
            fn()
            e.record()
            e.synchronize()
            total += s.elapsed_time(e)
        return total / iters

    def steady_ms(self, fn, iters, warm_ms=100.0):
        """Per-launch ms of `iters` back-to-back calls after at least warm_ms of
        untimed calls of the same work: under a new MFMA load the clocks dip for
        the first tens of milliseconds and recover (a batched-hgemm replay runs
        605 us, then up to 780 us, then 610 us again within ~30 replays)."""
        t = self.torch
        spent, n = 0.0, 1
        while spent < warm_ms:   # batches of back-to-back calls, doubled up to ~5 ms
            s, e = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
            s.record()
            for _ in range(n):
                fn()
            e.record()
            e.synchronize()
            el = max(s.elapsed_time(e), 0.01)
            spent += el
            if el < 5.0:
                n *= 2
        return self.per_launch_ms(fn, iters)


def traffic_from_profiles(kernel_key):
    """HBM bytes per launch measured with rocprofv3 --pmc (profiles/traffic.json)."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    try:
        return json.load(open(path)).get(kernel_key, {}).get("bytes_per_launch")
    except Exception:
        return None


# bytes per element/problem of each routine, for GB/s next to the CLI's GFLOP/s
def _sweep_rate(routine, sizes, gflops):
    if routine in ("reduction", "sdot"):
        n = sizes[0]
        flop = (n - 1) if routine == "reduction" else 2 * n
        byts = (4 if routine == "reduction" else 8) * n
        return {"GB/s": round(gflops * byts / flop, 1)}
    if routine == "sgemv":
        m, n = sizes
        return {"GB/s": round(gflops * 4 * (m * n + m + n) / (2 * m * n), 1)}
    return {"TFLOP/s": round(gflops / 1e3, 2)}


def cli_sizes_sweeps(timeout_s=240):
    """The reference's own benchmark sweeps: `libthom bench <routine> hip` runs
    each routine's SIZES (reduction.cpp:9, sdot.cpp:19, sgemv.cpp:18, dgemm.cpp:17,
    hgemm.cu:17-27, w4a16.cpp:17-27) in the reference's bench mode (BENCH_RUNS
    timed runs after WARMUP_RUNS, caches flushed per run) and writes
    bench_<routine>_hip.csv (routines.h:157-158); the rows are read back here."""
    import csv
    import subprocess
    import tempfile
    exe = os.path.join(ROOT, "libthom_amd", "bin", "libthom")
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        for routine in ("reduction", "sdot", "sgemv", "dgemm", "hgemm", "w4a16"):
            log(f"libthom bench {routine} hip")
            t0 = time.perf_counter()
            r = subprocess.run([exe, "bench", routine, "hip"], cwd=tmp, capture_output=True, text=True,
                               timeout=timeout_s)
            if r.returncode != 0:
                raise RuntimeError(f"libthom bench {routine} hip failed: {r.stdout[-800:]} {r.stderr[-800:]}")
            rows = []
            with open(os.path.join(tmp, f"bench_{routine}_hip.csv")) as f:
                for row in csv.DictReader(f):
                    dims = [int(row[k]) for k in ("size",) if k in row] or \
                           [int(row[k]) for k in ("M", "N", "K") if k in row]
                    g = float(row["gflops"])
                    e = {"size": dims, "GFLOP/s": round(g, 2), **_sweep_rate(routine, dims, g)}
                    if routine == "w4a16":
                        e.update({"us": float(row["time_us"]), "GB/s": float(row["gbps"]),
                                  "dense_fp16_us": float(row["cublas_time_us"]),
                                  "speedup_vs_dense_fp16": float(row["speedup_vs_cublas"])})
                    rows.append(e)
            out[routine] = {"rows": rows, "seconds": round(time.perf_counter() - t0, 1)}
    return out


def host_cpu():
    """(CPU model string, CPUs this process may run on)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return model, n


def cpu_baseline(seconds=10.0):
    """The oracle's sgemv (sgemv.cpp:132-147 restated: FP64 dot per row) on the
    host, on all the cores this process may use (row blocks, one per thread;
    the oracle call releases the GIL) and on one thread."""
    import numpy as np
    from concurrent.futures import ThreadPoolExecutor
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as orc
    from libthom_amd import inputs
    A = inputs.uniform_f32(M_SG * N_SG, -1.0, 1.0, 1234, inputs.S_SGEMV_A)
    x = inputs.uniform_f32(N_SG, -1.0, 1.0, 1234, inputs.S_SGEMV_X)
    y = np.zeros(M_SG, np.float32)
    model, ncpu = host_cpu()
    threads = max(1, min(16, ncpu))   # a GPU box's share is 16 CPUs per GPU

    def timed(fn, budget):
        reps, t0 = 0, time.perf_counter()
        while True:
            fn()
            reps += 1
            el = time.perf_counter() - t0
            if el >= budget:
                return reps, el

    bounds = [M_SG * t // threads for t in range(threads + 1)]

    def block(t):
        r0, r1 = bounds[t], bounds[t + 1]
        return orc.sgemv(True, r1 - r0, N_SG, 1.0, A[r0 * N_SG:r1 * N_SG], N_SG, x, 1, 0.0, y[r0:r1], 1)

    with ThreadPoolExecutor(threads) as pool:
        full = np.concatenate(list(pool.map(block, range(threads))))
        assert np.array_equal(full, orc.sgemv(True, M_SG, N_SG, 1.0, A, N_SG, x, 1, 0.0, y, 1))
        reps_mt, el_mt = timed(lambda: list(pool.map(block, range(threads))), seconds / 2)
    reps_1, el_1 = timed(lambda: orc.sgemv(True, M_SG, N_SG, 1.0, A, N_SG, x, 1, 0.0, y, 1), seconds / 2)
    gbs_mt = SG_BYTES * reps_mt / el_mt / 1e9
    gbs_1 = SG_BYTES * reps_1 / el_1 / 1e9
    return {"value": round(gbs_mt, 3), "unit": "GB/s", "cores": threads, "kind": "port",
            "sample": f"full 16384x8192 sgemv, {reps_mt} runs in {el_mt:.1f} s on {threads} threads "
                      f"(oracle/oracle.c orc_sgemv on row blocks, FP64 dot per row)",
            "single_thread": {"value": round(gbs_1, 3), "unit": "GB/s", "cores": 1,
                              "sample": f"full sgemv, {reps_1} runs in {el_1:.1f} s"},
            "host_cpus": ncpu, "cpu_model": model,
            "per_routine": cpu_oracle_routines()}


def cpu_oracle_routines(budget_s=2.0):
    """The oracle's other checks timed on the host on bounded samples of each
    configuration (threads as stated), so each routine's CPU path has a rate."""
    import numpy as np
    import _oracle as orc
    from libthom_amd import inputs
    threads = orc.THREADS

    def rate(fn):
        reps, t0 = 0, time.perf_counter()
        while True:
            fn()
            reps += 1
            el = time.perf_counter() - t0
            if el >= budget_s:
                return reps / el

    out = {}
    x = inputs.uniform_f32(1_000_000, 0.0, 1.0, 1234, inputs.S_REDUCE)
    r = rate(lambda: orc.sum_f64(x))
    out["reduction_n1000000"] = {"GB/s": round(4e6 * r / 1e9, 3), "cores": 1, "sample": "full n = 1e6, FP64 sum"}
    y = inputs.uniform_f32(1_000_000, -1.0, 1.0, 1234, inputs.S_SDOT_Y)
    r = rate(lambda: orc.dot_f64(x, y))
    out["sdot_n1000000"] = {"GB/s": round(8e6 * r / 1e9, 3), "cores": 1, "sample": "full n = 1e6, FP64 dot"}
    n = 512
    A = inputs.normal_f64(n * n, 1234, inputs.S_DGEMM_A).reshape(n, n)
    B = inputs.normal_f64(n * n, 1234, inputs.S_DGEMM_B).reshape(n, n)
    r = rate(lambda: orc.dgemm(A, B, threads=threads))
    out["dgemm"] = {"GFLOP/s": round(2 * n ** 3 * r / 1e9, 2), "cores": threads,
                    "sample": "512^3 of the 4096^3 configuration (N(0,1)), blocked FP64"}
    r = rate(lambda: orc.ozaki_dgemm(A, B, 8, threads=threads))
    out["ozaki8"] = {"GFLOP/s_fp64_equiv": round(2 * n ** 3 * r / 1e9, 2), "cores": threads,
                     "sample": "512^3, 8 slices (split, exact level sums, FP64 recombination)"}
    n = 1024
    Ah = inputs.uniform_f16(n * n, -1, 1, 1234, 31).reshape(n, n)
    Bh = inputs.uniform_f16(n * n, -1, 1, 1234, 32).reshape(n, n)
    r = rate(lambda: orc.hgemm(Ah, Bh, threads=threads))
    out["hgemm"] = {"GFLOP/s": round(2 * n ** 3 * r / 1e9, 2), "cores": threads,
                    "sample": "1024^3, FP32 accumulation of FP16 products"}
    Aw, q, sc, z = inputs.w4a16_problem(16, 11008, 4096, 128)
    packed, qz = inputs.pack_w4(q), inputs.pack_zeros(z)
    Wdq = orc.w4_dequant(packed, sc, qz, 128)
    r = rate(lambda: orc.w4a16_rows(Aw, Wdq, threads=threads))
    out["w4a16_M16"] = {"GFLOP/s": round(2 * 16 * 11008 * 4096 * r / 1e9, 2), "cores": threads,
                        "sample": "M = 16 configuration on the dequantised weights (FP32 k-ascending)"}
    r = rate(lambda: orc.w4_dequant(packed, sc, qz, 128))
    out["w4_dequant"] = {"weights/s": round(4096 * 11008 * r, 0), "cores": 1,
                         "sample": "4096 x 11008 int4 weights -> FP16"}
    Wq = Wdq[:1024]
    r = rate(lambda: orc.w4_quantize_sym(Wq, 128))
    out["w4_quantize_sym"] = {"weights/s": round(1024 * 11008 * r, 0), "cores": 1,
                              "sample": "1024 x 11008 fp16 weights -> int4 + scales (quantize.cpp:8-37)"}
    return out


def bench_routines(torch, R, timer, dev, quick):
    import numpy as np
    from libthom_amd import inputs
    out = {}
    flush_buf = torch.empty(512 * 1024 * 1024 // 4, dtype=torch.float32, device=dev)
    flush_src = torch.empty_like(flush_buf)

    def flush():
        R.copy_f32(flush_buf, flush_src)

    it = 5 if quick else 20
    # sgemv on the configured 16384 x 8192 matrix stored column-major (the
    # reference's col-major layout, sgemv.cpp:180-181): column slices + a
    # slice-order sum; back-to-back launches over the 537 MB stream
    m, n = 16384, 8192
    Ac = torch.empty(n, m, dtype=torch.float32, device=dev)
    R.fill_uniform(Ac, -1.0, 1.0, 1234, inputs.S_SGEMV_A)
    xc = torch.empty(n, dtype=torch.float32, device=dev)
    R.fill_uniform(xc, -1.0, 1.0, 1234, inputs.S_SGEMV_X)
    yc = torch.zeros(m, dtype=torch.float32, device=dev)
    one = torch.ones(1, dtype=torch.float32, device=dev)
    zero = torch.zeros(1, dtype=torch.float32, device=dev)
    col = lambda: R.sgemv(Ac, xc, yc, row_major=False, alpha_dev=one, beta_dev=zero)
    col()
    ms = timer.per_launch_ms(col, it)
    gbs = 4 * (m * n + n + m) / ms / 1e6
    rows, _ = sample_entries(m, 1, 2048)
    err = dot2_err(yc.cpu().numpy()[rows], Ac.cpu().numpy().T, xc.cpu().numpy()[:, None], rows, 0 * rows)
    out["sgemv_col_16384x8192"] = {"ms": round(ms, 5), "GB/s": round(gbs, 1), "frac_hbm": round(gbs / HBM_PEAK_GBS, 4),
                                   "max_rel_err_vs_fp64": err}
    del Ac
    torch.cuda.empty_cache()
    # reduction
    for n in (1_000_000, 500_000_000):
        x = torch.empty(n, dtype=torch.float32, device=dev)
        R.fill_uniform(x, 0.0, 1.0, 1234, inputs.S_REDUCE)
        o = torch.empty(1, dtype=torch.float32, device=dev)
        R.reduce_sum(x, o)
        ms = timer.per_launch_ms(lambda: R.reduce_sum(x, o), it, flush if n < 100_000_000 else None)
        gbs = 4 * n / ms / 1e6
        ref = _orc().sum_f64(x.cpu().numpy())   # the oracle's FP64 sum
        out[f"reduction_n{n}"] = {"ms": round(ms, 5), "GB/s": round(gbs, 1), "frac_hbm": round(gbs / HBM_PEAK_GBS, 4),
                                  "GFLOP/s": round((n - 1) / ms / 1e6, 1),
                                  "max_rel_err_vs_fp64": abs(o.item() - ref) / abs(ref)}
        del x
    # sdot
    for n in (1_000_000, 250_000_000):
        xs = torch.empty(n, dtype=torch.float32, device=dev)
        ys = torch.empty(n, dtype=torch.float32, device=dev)
        R.fill_uniform(xs, -1.0, 1.0, 1234, inputs.S_SDOT_X)
        R.fill_uniform(ys, -1.0, 1.0, 1234, inputs.S_SDOT_Y)
        o = torch.empty(1, dtype=torch.float32, device=dev)
        R.sdot(xs, ys, o)
        ms = timer.per_launch_ms(lambda: R.sdot(xs, ys, o), it, flush if n < 100_000_000 else None)
        gbs = 8 * n / ms / 1e6
        xh, yh = xs.cpu().numpy(), ys.cpu().numpy()   # the oracle's FP64 dot
        err = abs(o.item() - _orc().dot_f64(xh, yh)) / _orc().dot_f64(abs(xh), abs(yh))
        out[f"sdot_n{n}"] = {"ms": round(ms, 5), "GB/s": round(gbs, 1), "frac_hbm": round(gbs / HBM_PEAK_GBS, 4),
                             "max_rel_err_vs_fp64": err}
        del xs, ys, xh, yh
    torch.cuda.empty_cache()
    # dgemm: native FP64 MFMA vs split-precision on FP8 MFMA, errors against
    # the compensated (Dot2) product, relative to (|A||B|)_ij, on both inputs
    n = 4096
    A0 = inputs.normal_f64(n * n, 1234, inputs.S_DGEMM_A).reshape(n, n)
    B0 = inputs.normal_f64(n * n, 1234, inputs.S_DGEMM_B).reshape(n, n)
    rs = 2.0 ** inputs.int_uniform(n, -20, 20, 1234, inputs.S_DGEMM_RS)
    cs = 2.0 ** inputs.int_uniform(n, -20, 20, 1234, inputs.S_DGEMM_CS)
    study = {}
    for label, (Ah, Bh) in (("a_normal", (A0, B0)), ("b_scaled", (A0 * rs[:, None], B0 * cs[None, :]))):
        A = torch.from_numpy(Ah).to(dev)
        B = torch.from_numpy(Bh).to(dev)
        hi, lo = R.dgemm_compensated(A, B)
        scale = R.dgemm(A.abs(), B.abs())
        C = torch.empty(n, n, dtype=torch.float64, device=dev)
        R.dgemm(A, B, C)
        err = lambda Cx: float(((Cx - hi - lo).abs() / scale).max().item())
        ms = timer.steady_ms(lambda: R.dgemm(A, B, C), max(3, it // 4))
        tf = 2 * n ** 3 / ms / 1e9
        row = {"native": {"ms": round(ms, 4), "TFLOP/s": round(tf, 2), "frac_fp64_mfma": round(tf / FP64_PEAK, 4),
                          "max_rel_err": err(C)}}
        for S in OZ_SLICES:
            Co = torch.empty_like(C)
            R.dgemm_ozaki(A, B, S, Co)
            mso = timer.steady_ms(lambda: R.dgemm_ozaki(A, B, S, Co), max(3, it // 4))
            tfo = 2 * n ** 3 / mso / 1e9
            pairs = S * (S + 1) // 2
            row[f"ozaki{S}"] = {"ms": round(mso, 4), "TFLOP/s_fp64_equiv": round(tfo, 2),
                                "fp8_TFLOP/s": round(tfo * pairs, 1),
                                "frac_fp8_mfma": round(tfo * pairs / FP8_PEAK, 4), "max_rel_err": err(Co)}
        eq = [S for S in OZ_SLICES if row[f"ozaki{S}"]["max_rel_err"] <= row["native"]["max_rel_err"]]
        row["equal_error"] = ({"slices": eq[0], "emulated_TFLOP/s": row[f"ozaki{eq[0]}"]["TFLOP/s_fp64_equiv"],
                               "native_TFLOP/s": row["native"]["TFLOP/s"]} if eq else
                              {"slices": None,
                               "note": f"no slice count up to {OZ_SLICES[-1]} reaches the native error"})
        study[label] = row
        del A, B, C, hi, lo, scale
        torch.cuda.empty_cache()
    out["dgemm_4096"] = study
    # hgemm
    for n in (4096, 8192):
        A = torch.empty(n, n, dtype=torch.float16, device=dev)
        B = torch.empty(n, n, dtype=torch.float16, device=dev)
        R.fill_uniform(A, -1, 1, 1234, 31)
        R.fill_uniform(B, -1, 1, 1234, 32)
        C = torch.empty(n, n, dtype=torch.float16, device=dev)
        R.hgemm(A, B, C)
        ms = timer.steady_ms(lambda: R.hgemm(A, B, C), max(3, it // 4))
        tf = 2 * n ** 3 / ms / 1e9
        rows, cols = sample_entries(n, n, 2048)
        err = dot2_err(C.cpu().numpy()[rows, cols], A.cpu().numpy(), B.cpu().numpy(), rows, cols)
        out[f"hgemm_{n}"] = {"ms": round(ms, 4), "TFLOP/s": round(tf, 1), "frac_fp16_mfma": round(tf / FP16_PEAK, 4),
                             "config": repr(R.hgemm_default_config(n, n, n)), "max_rel_err_vs_fp64": err}
        del A, B, C
    # batched hgemm: 256 uneven problems, one persistent launch and its hipGraph
    from libthom_amd import tuner
    dims = R.batched_problem_dims()
    bcfg, policy = tuner.best_batched(len(dims))
    probs, _ = tuner.batched_problems(dims, dev, policy, bcfg)
    flops = sum(2 * M * N * K for M, N, K in dims)
    plan = R.GroupedHgemm(probs, bcfg)
    side = torch.cuda.Stream(device=dev)
    plan.run(side)
    plan.capture(side)
    with torch.cuda.stream(side):
        plan.replay(side)
        ms = timer.steady_ms(lambda: plan.replay(side), max(3, it // 2))
    tf = flops / ms / 1e9
    out["hgemm_batched_256"] = {"ms": round(ms, 4), "TFLOP/s": round(tf, 1), "frac_fp16_mfma": round(tf / FP16_PEAK, 4),
                                "flop": flops, "launch": "one persistent kernel in a hipGraph",
                                "config": repr(bcfg), "split_policy": policy, **plan.info()}
    plan.close()
    del probs
    torch.cuda.empty_cache()
    # w4a16
    from libthom_amd import W4A16Config
    Nw, Kw, G = 11008, 4096, 128
    tuned = {}
    try:
        w4_model = tuner.CostModel.load(os.path.join(tuner.TUNING_DIR, "w4a16_mlp.json"), device=str(dev))
    except Exception:   # no trained model: the tuned rows are left out
        w4_model = None
    for M in (1, 16, 128, 2048):
        Aw, q, s, z = inputs.w4a16_problem(M, Nw, Kw, G)
        At = torch.from_numpy(Aw).to(dev)
        Wt = torch.from_numpy(inputs.pack_w4(q).view(np.int32)).to(dev)
        St = torch.from_numpy(s).to(dev)
        Zt = torch.from_numpy(inputs.pack_zeros(z).view(np.int32)).to(dev)
        Ct = torch.empty(M, Nw, dtype=torch.float16, device=dev)
        R.w4a16(At, Wt, St, Zt, G, Ct)
        # one call on cold caches, HIP events around it (includes the launch)
        ms_call = timer.per_launch_ms(lambda: R.w4a16(At, Wt, St, Zt, G, Ct), it, flush)
        # M > 256, compute-bound: back-to-back launches after the clocks
        # settle.  M <= 256: back-to-back launches over 24 copies of the
        # weights (560 MB, more than the 256 MB Infinity Cache), replayed from
        # a hipGraph so the host does not pace the ~10 us kernels: the kernel's
        # own rate with the weights streamed from HBM
        ws = [(Wt, St, Zt)] + ([(Wt.clone(), St.clone(), Zt.clone()) for _ in range(23)] if M <= 256 else [])

        def time_cfg(cfg, warm_ms=100.0):
            if M > 256:
                return timer.steady_ms(lambda: R.w4a16(At, Wt, St, Zt, G, Ct, config=cfg), it, warm_ms)
            side = torch.cuda.Stream(device=dev)
            side.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(side):
                for w_, s_, z_ in ws:
                    R.w4a16(At, w_, s_, z_, G, Ct, config=cfg)
                graph = torch.cuda.CUDAGraph()
                with torch.cuda.graph(graph, stream=side):
                    for w_, s_, z_ in ws:
                        R.w4a16(At, w_, s_, z_, G, Ct, config=cfg)
                graph.replay()
                t = timer.steady_ms(graph.replay, max(4, it // 4), warm_ms) / len(ws)
            torch.cuda.current_stream().wait_stream(side)
            torch.cuda.synchronize()
            del graph
            return t

        ms = time_cfg(None)
        # the learned tuner's configuration for this shape (its MLP evaluated by
        # lthip_mlp_forward) against the exhaustive optimum over the same space
        if w4_model is not None:
            choice = w4_model.choose([(M, Nw, Kw, G)])[0]
            space = [c for c in tuner.W4A16_SPACE if tuner.valid(c, M, Kw, Nw, G)]
            times = {}
            for c in space:
                try:
                    times[c] = time_cfg(W4A16Config(*c), 20.0)
                except Exception:   # a configuration the library turns down for this shape
                    pass
            best = min(times, key=times.get) if times else None
            ok = best is not None and choice in times   # (null fields: the library turned the choice down)
            tuned[f"M{M}"] = {"mlp_config": list(choice), "us": round(times[choice] * 1e3, 2) if ok else None,
                              "exhaustive_best": list(best) if best else None,
                              "best_us": round(times[best] * 1e3, 2) if best else None,
                              "regret": round(times[choice] / times[best] - 1.0, 4) if ok else None,
                              "auto_us": round(ms * 1e3, 2), "candidates": len(space), "timed": len(times)}
        R.w4a16(At, Wt, St, Zt, G, Ct)
        del ws
        byts = 2 * M * Kw + Kw // 8 * Nw * 4 + Kw // G * Nw * 2 + Kw // G * Nw // 2 + 2 * M * Nw
        flops = 2 * M * Nw * Kw
        Wd = R.w4_dequant(Wt, St, Zt, G).cpu().numpy()   # bit-exact dequantisation
        rows, cols = sample_entries(M, Nw, 2048)
        err = dot2_err(Ct.cpu().numpy()[rows, cols], Aw, Wd, rows, cols)
        out[f"w4a16_M{M}"] = {"us": round(ms * 1e3, 2), "us_one_call_cold": round(ms_call * 1e3, 2),
                              "GB/s": round(byts / ms / 1e6, 1),
                              "frac_hbm": round(byts / ms / 1e6 / HBM_PEAK_GBS, 4),
                              "TFLOP/s": round(flops / ms / 1e9, 2),
                              "frac_fp16_mfma": round(flops / ms / 1e9 / FP16_PEAK, 4), "max_rel_err_vs_fp64": err,
                              "kernel": w4a16_kind_name(M, Nw, Kw)}
        del Wd
    if tuned:
        out["w4a16_tuned"] = tuned
    # grouped w4a16: many problems, each with its own weights, in one
    # persistent launch; against the same problems as separate automatic
    # launches, both replayed from a hipGraph (copies of the problem set rotate
    # the weights through HBM)
    out.update(bench_w4a16_grouped(torch, R, inputs, np, timer, dev, it))
    # the int4 quantizer (the producer of the w4a16 weights): W 4096 x 11008
    # fp16 -> packed int4, scales (and zero-points), over 4 copies of W (360 MB,
    # more than the Infinity Cache) replayed from a hipGraph
    Wq = torch.empty(Kw, Nw, dtype=torch.float16, device=dev)
    R.fill_uniform(Wq, -1.0, 1.0, 1234, inputs.S_W4_Q)
    copies = 4
    srcs = [Wq] + [Wq.clone() for _ in range(copies - 1)]
    for asym in (False, True):
        outs = [R.w4_quantize(w_, G, asym) for w_ in srcs]
        side = torch.cuda.Stream(device=dev)
        side.wait_stream(torch.cuda.current_stream())
        K8 = Kw // 8
        with torch.cuda.stream(side):
            graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(graph, stream=side):
                for w_, (p_, s_, z_) in zip(srcs, outs):
                    R.call("lthip_w4_quantize", w_.data_ptr(), p_.data_ptr(), s_.data_ptr(), R._ptr(z_), Kw, Nw, G,
                           1 if asym else 0, R._stream(dev))
            graph.replay()
            ms = timer.steady_ms(graph.replay, max(4, it // 4)) / copies
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        byts = 2 * Kw * Nw + 4 * K8 * Nw + 2 * (Kw // G) * Nw + (4 * (Kw // G) * (Nw // 8) if asym else 0)
        out["w4_quantize_" + ("asym" if asym else "sym")] = {
            "us": round(ms * 1e3, 2), "GB/s": round(byts / ms / 1e6, 1), "frac_hbm": round(byts / ms / 1e6 / HBM_PEAK_GBS, 4),
            "weights/s": round(Kw * Nw / ms * 1e3, 1), "shape": f"{Kw}x{Nw} fp16, group {G}"}
        del outs, graph
    del srcs, Wq
    return out


def w4a16_grouped_sets(np):
    """The grouped w4a16 workloads: (name, [(M, N, K, group)], copies)."""
    rng = np.random.default_rng(1234)
    experts = np.maximum(1, np.round(np.exp(rng.uniform(0.0, np.log(512), 32)))).astype(int)
    return [("w4a16_grouped_configured", [(M, 11008, 4096, 128) for M in (1, 16, 128, 2048)], 3),
            ("w4a16_grouped_32_experts", [(int(M), 2048, 4096, 128) for M in experts], 4)]


def bench_w4a16_grouped(torch, R, inputs, np, timer, dev, it):
    out = {}
    for name, dims, copies in w4a16_grouped_sets(np):
        sets = []
        for c in range(copies):
            probs = []
            for i, (M, N, K, G) in enumerate(dims):
                A, q, s, z = inputs.w4a16_problem(M, N, K, G, seed=1234 + 97 * c + i)
                probs.append((torch.from_numpy(A).to(dev), torch.from_numpy(inputs.pack_w4(q).view(np.int32)).to(dev),
                              torch.from_numpy(s).to(dev),
                              torch.from_numpy(inputs.pack_zeros(z).view(np.int32)).to(dev),
                              torch.empty(M, N, dtype=torch.float16, device=dev), G))
            sets.append(probs)
        plans = [R.GroupedW4A16(p) for p in sets]
        separate = [lambda a=a, w=w, s=s, z=z, C=C, G=G: R.w4a16(a, w, s, z, G, C)
                    for p in sets for (a, w, s, z, C, G) in p]

        def graph_us(fns):
            side = torch.cuda.Stream(device=dev)
            side.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(side):
                for f in fns:
                    f()
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g, stream=side):
                    for f in fns:
                        f()
                g.replay()
                ms = timer.steady_ms(g.replay, max(4, it // 4))
            torch.cuda.current_stream().wait_stream(side)
            torch.cuda.synchronize()
            return ms * 1e3 / copies

        us_g = graph_us([p.run for p in plans])
        us_s = graph_us(separate)
        # the grouped result against an FP64 product of sampled rows of each problem
        for p in plans:
            p.run()
        torch.cuda.synchronize()
        err = 0.0
        for (a, w, s, z, C, G) in sets[0]:
            rows, cols = sample_entries(a.shape[0], C.shape[1], 512)
            err = max(err, dot2_err(C.cpu().numpy()[rows, cols], a.cpu().numpy(),
                                    R.w4_dequant(w, s, z, G).cpu().numpy(), rows, cols))
        flops = sum(2 * M * N * K for M, N, K, _ in dims)
        out[name] = {"us": round(us_g, 1), "TFLOP/s": round(flops / us_g / 1e6, 1),
                     "frac_fp16_mfma": round(flops / us_g / 1e6 / FP16_PEAK, 4),
                     "separate_launches_us": round(us_s, 1), "speedup_vs_separate": round(us_s / us_g, 2),
                     "problems": len(dims), "M": [d[0] for d in dims], "N": dims[0][1], "K": dims[0][2],
                     "launch": "one persistent kernel in a hipGraph", **plans[0].info(),
                     "max_rel_err_vs_fp64": err}
        for p in plans:
            p.close()
        del sets, plans
        torch.cuda.empty_cache()
    return out


def w4a16_kind_name(M, N, K):
    # the automatic choice of lthip_w4a16_gemm for this shape (group 128)
    if M <= 2:
        return "decode (kind 4)"
    if M <= 64:
        return "skinny MFMA (kind 5)"
    if N % 32 != 0:
        return "tiled MFMA (kind 2)"
    if M <= 256 and ((M + 127) // 128) * ((N + 127) // 128) * (K // 64) >= 8 * 256:
        return "stream-K pipelined tile (kind 6, 128x128)"
    return "pipelined MFMA tile (kind 3, " + ("64x128)" if M <= 256 else "128x128 or 256x256)")


def add_measured_fractions(routines, peaks):
    """frac_measured next to each datasheet fraction: GB/s over the measured
    HBM read rate, TFLOP/s over the measured MFMA rate of the routine's type."""
    def walk(name, d):
        if name == "vendor_reference":
            return
        if not isinstance(d, dict):
            return
        if "GB/s" in d:
            d["frac_measured_hbm"] = round(d["GB/s"] / peaks["hbm_read_gbs"], 4)
        if "frac_fp16_mfma" in d:
            d["frac_measured_fp16"] = round(d["TFLOP/s"] / peaks["fp16_tflops"], 4)
        if "frac_fp64_mfma" in d:
            d["frac_measured_fp64"] = round(d["TFLOP/s"] / peaks["fp64_tflops"], 4)
        if "frac_fp8_mfma" in d:
            d["frac_measured_fp8"] = round(d["fp8_TFLOP/s"] / peaks["fp8_tflops"], 4)
        for k, v in d.items():
            walk(k, v)
    for k, v in routines.items():
        walk(k, v)


def vendor_reference(torch, timer, dev):
    """torch.matmul (hipBLASLt / rocBLAS) on the hgemm and dgemm shapes, and the
    FP16 GEMM on dequantised weights that the reference's w4a16 compares with."""
    out = {"note": "vendor library through torch.matmul, for context; not this library"}
    for n in (4096, 8192):
        a = torch.randn(n, n, device=dev, dtype=torch.float16)
        b = torch.randn(n, n, device=dev, dtype=torch.float16)
        torch.matmul(a, b)
        ms = timer.steady_ms(lambda: torch.matmul(a, b), 10)
        out[f"fp16_matmul_{n}"] = {"ms": round(ms, 4), "TFLOP/s": round(2 * n ** 3 / ms / 1e9, 1)}
        del a, b
    n = 4096
    a = torch.randn(n, n, device=dev, dtype=torch.float64)
    b = torch.randn(n, n, device=dev, dtype=torch.float64)
    torch.matmul(a, b)
    ms = timer.steady_ms(lambda: torch.matmul(a, b), 5)
    out[f"fp64_matmul_{n}"] = {"ms": round(ms, 4), "TFLOP/s": round(2 * n ** 3 / ms / 1e9, 2)}
    del a, b
    # the reference's w4a16 baseline (w4a16.cpp:306): an FP16 GEMM on the
    # dequantised weights, K = 4096, N = 11008; M <= 256 timed like the w4a16
    # rows (back-to-back launches over 4 copies of the 90 MB fp16 weights,
    # replayed from a hipGraph)
    try:
        vendor_w4a16_baseline(torch, timer, dev, out)
    except Exception as e:   # a context row, never worth losing the run for
        out["fp16_matmul_w4a16"] = {"error": f"{type(e).__name__}: {e}"[:200]}
    torch.cuda.empty_cache()
    return out


def vendor_w4a16_baseline(torch, timer, dev, out):
    K, N = 4096, 11008
    w = [torch.randn(K, N, device=dev, dtype=torch.float16) for _ in range(4)]
    for M in (1, 16, 128, 2048):
        a = torch.randn(M, K, device=dev, dtype=torch.float16)
        c = torch.empty(M, N, device=dev, dtype=torch.float16)
        for w_ in w:
            torch.matmul(a, w_, out=c)
        if M <= 256:
            side = torch.cuda.Stream(device=dev)
            side.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(side):
                graph = torch.cuda.CUDAGraph()
                with torch.cuda.graph(graph, stream=side):
                    for w_ in w:
                        torch.matmul(a, w_, out=c)
                graph.replay()
                ms = timer.steady_ms(graph.replay, 10) / len(w)
            torch.cuda.current_stream().wait_stream(side)
            torch.cuda.synchronize()
            del graph
        else:
            ms = timer.steady_ms(lambda: torch.matmul(a, w[0], out=c), 10)
        out[f"fp16_matmul_w4a16_M{M}"] = {"us": round(ms * 1e3, 2), "TFLOP/s": round(2 * M * N * K / ms / 1e9, 2),
                                          "GB/s": round((2 * K * N + 2 * M * K + 2 * M * N) / ms / 1e6, 1)}
        del a, c
    del w


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--quick", action="store_true", help="fewer iterations for the per-routine table")
    ap.add_argument("--no-routines", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-sweeps", action="store_true", help="skip the reference's SIZES sweeps (libthom bench)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run 5 plain 256 MiB copies (known traffic, to calibrate HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path (each routine would shard its "
              "columns or rows across GPUs independently); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from libthom_amd import inputs, routines as R

    dev = torch.device("cuda:0")
    timer = Timer(torch)
    log("allocating the sgemv workload (512 MiB matrix, filled on the GPU)")
    A = torch.empty(M_SG, N_SG, dtype=torch.float32, device=dev)
    R.fill_uniform(A, -1.0, 1.0, 1234, inputs.S_SGEMV_A)
    x = torch.empty(N_SG, dtype=torch.float32, device=dev)
    R.fill_uniform(x, -1.0, 1.0, 1234, inputs.S_SGEMV_X)
    y = torch.zeros(M_SG, dtype=torch.float32, device=dev)
    one = torch.ones(1, dtype=torch.float32, device=dev)
    zero = torch.zeros(1, dtype=torch.float32, device=dev)

    def step():
        R.sgemv(A, x, y, alpha_dev=one, beta_dev=zero)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(args.steps):
        step()
    end.record()
    torch.cuda.synchronize()
    ms = start.elapsed_time(end) / args.steps
    gbs = SG_BYTES / ms / 1e6
    log(f"sgemv 16384x8192: {ms * 1e3:.1f} us/step, {gbs:.0f} GB/s")
    rows, _ = sample_entries(M_SG, 1, 4096)
    sg_err = dot2_err(y.cpu().numpy()[rows], A.cpu().numpy(), x.cpu().numpy()[:, None], rows, 0 * rows)
    torch.cuda.empty_cache()

    if args.copy_probe:
        src = torch.empty(64 * 1024 * 1024, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        R.fill_uniform(src, 0.0, 1.0, 1234, 99)
        for _ in range(5):
            R.copy_f32(dst, src)
        torch.cuda.synchronize()
        log("copy probe: 5 x 256 MiB copied")
        del src, dst

    routines = None
    peaks = None
    if not args.no_routines:
        del A
        torch.cuda.empty_cache()
        log("measured peaks (MFMA probe loops, HBM read stream, copy)")
        peaks = R.measured_peaks(dev)
        log(f"measured peaks: {peaks}")
        log("per-routine table")
        routines = bench_routines(torch, R, timer, dev, args.quick)
        routines["vendor_reference"] = vendor_reference(torch, timer, dev)
        add_measured_fractions(routines, peaks)
    sweeps = None
    if not args.no_sweeps:
        log("the reference's SIZES sweeps (libthom bench <routine> hip)")
        sweeps = cli_sizes_sweeps()
    cpu = None
    if not args.no_cpu:
        log("cpu baseline (oracle sgemv, all cores and 1 thread)")
        cpu = cpu_baseline(args.cpu_seconds)

    line = {
        "metric": "% of HBM/MFMA roofline per routine; emulated vs native DGEMM TFLOP/s @ equal err",
        "value": round(100.0 * gbs / HBM_PEAK_GBS, 2),
        "unit": "% of HBM peak (sgemv 16384x8192 FP32, 8 TB/s)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (SplitMix64 seed 1234, U[-1,1))",
        "config": {"workload": "sgemv row-major y = A x, A 16384 x 8192 FP32", "M": M_SG, "N": N_SG,
                   "bytes_per_step": SG_BYTES},
        "max_rel_err_vs_fp64": sg_err,
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles("sgemv_16384x8192")},
        "cpu_baseline": cpu,
        "measured_peaks": peaks,
        "routines": routines,
        "sizes_sweeps": sweeps,
    }
    if peaks:
        line["roofline"]["measured_peak"] = peaks["hbm_read_gbs"]
        line["roofline"]["frac_measured"] = round(gbs / peaks["hbm_read_gbs"], 4)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
