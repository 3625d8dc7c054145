"""The libthom CLI (libthom_amd/bin/libthom): the reference's command line
(main.cpp, routines.h Handlers) over the gfx950 library.

CPU tests cover dispatch and argument validation, which happen before any HIP
call; GPU tests run each routine's test mode, whose CPU check makes the process
exit non-zero on a mismatch.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "libthom_amd", "bin", "libthom")


def run(*args, timeout=300, env=None, cwd=None):
    if not os.path.exists(EXE):
        pytest.fail("libthom CLI not built (make -C libthom_amd)")
    return subprocess.run([EXE, *map(str, args)], capture_output=True, text=True, timeout=timeout, env=env, cwd=cwd)


def test_usage_and_unknown_routine():
    r = run()
    assert r.returncode == 1 and "at least" in r.stderr
    r = run("test", "nosuch")
    assert r.returncode != 0 and "Invalid mode or routine" in r.stderr
    for routine in ("reduction", "sdot", "sgemv", "dgemm", "hgemm", "w4a16"):
        assert routine in r.stderr
    r = run("explode", "reduction")
    assert r.returncode != 0


def test_missing_arguments_list_implementations():
    r = run("test", "dgemm")
    assert r.returncode == 1 and "<size3> <nb_iter>" in r.stderr
    assert "hip" in r.stderr and "ozaki8" in r.stderr
    r = run("test", "sgemv", "nope", 4, 4, 1)
    assert r.returncode == 1 and "Invalid implementation" in r.stderr
    r = run("test", "reduction", "hip", 100, 0)
    assert r.returncode == 1 and "nb_iter" in r.stderr
    r = run("tune", "hgemm", 1, 2)
    assert r.returncode == 1 and "Usage" in r.stderr
    # the batched routines check their arguments before any HIP call too
    for args in (("test", "w4a16_grouped"), ("bench", "w4a16_grouped", "bogus"), ("tune", "w4a16_grouped", "graph"),
                 ("test", "w4a16_grouped", "graph", 4, 520), ("test", "w4a16_grouped", "graph", 0)):
        r = run(*args)
        assert r.returncode == 1 and ("Usage" in r.stderr or "need positive" in r.stderr), args
    r = run("test", "hgemm_batched", "bogus")
    assert r.returncode == 1 and "Usage" in r.stderr
    # every routine registers the host implementations beside the HIP ones
    for routine in ("reduction", "sdot", "sgemv", "dgemm", "hgemm", "w4a16"):
        r = run("test", routine)
        assert r.returncode == 1 and " cpu " in r.stderr + " " and "cpu_mt" in r.stderr, (routine, r.stderr)


@pytest.mark.gpu
@pytest.mark.parametrize("args", [
    ("reduction", "hip", 1_000_000, 3),
    ("reduction", "hip", 1_000_003, 2),
    ("sdot", "hip", 1_000_000, 3),
    ("sgemv", "hip", 2048, 1024, 2, 1, "row"),
    ("sgemv", "hip", 2000, 1000, 2, 1, "col"),
    ("dgemm", "hip", 512, 384, 256, 2),
    ("dgemm", "ozaki8", 256, 256, 512, 2),
    ("dgemm", "ozaki10", 256, 256, 512, 2, 1, "scaled"),
    ("dgemm", "ozaki11", 256, 128, 512, 2),
    ("hgemm", "hip", 1024, 512, 768, 2),
    ("w4a16", "hip", 16, 1024, 1024, 2),
    ("w4a16", "tiled", 192, 512, 1024, 2),
    ("w4a16", "stream", 1, 11008, 4096, 2, 1, "sym"),
    ("w4a16", "decode", 1, 11008, 4096, 2),
    ("w4a16", "decode", 3, 4096, 1024, 2, 1, "sym"),
    ("w4a16", "skinny", 16, 11008, 4096, 2),
    ("w4a16", "skinny", 40, 520, 1024, 2, 1, "sym"),
    ("w4a16", "pipelined", 300, 512, 1024, 2),
    ("w4a16", "pipelined", 2048, 11008, 4096, 2, 1, "sym"),
])
def test_cli_test_mode_passes_checks(args):
    r = run("test", *args)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "# GFLOP/s" in r.stdout


@pytest.mark.gpu
def test_cli_w4a16_grouped():
    # every problem bit-identical to the single-problem tile run alone
    for impl in ("persistent", "graph"):
        r = run("test", "w4a16_grouped", impl, 12, 544, 1024, 2)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        assert "ERROR" not in r.stdout and "# GFLOP/s" in r.stdout
    r = run("bench", "w4a16_grouped", "bogus")
    assert r.returncode != 0


@pytest.mark.gpu
def test_cli_batched_and_tune(tmp_path):
    r = run("test", "hgemm_batched", "persistent", 16, 2)
    assert r.returncode == 0, r.stdout + r.stderr
    r = run("test", "hgemm_batched", "graph", 16, 2)
    assert r.returncode == 0, r.stdout + r.stderr
    env = dict(os.environ, LTHIP_TUNING_DIR=str(tmp_path))
    r = run("tune", "hgemm", 256, 256, 512, 3, env=env)
    assert r.returncode == 0 and "[best]" in r.stdout, r.stdout[-2000:] + r.stderr
    rows = open(tmp_path / "hgemm.csv").read().splitlines()
    assert rows[0] == "M,N,K,bm,bn,bk,split_k,stages,ms" and len(rows) > 10
    r = run("test", "hgemm", "tuned", 256, 256, 512, 2, env=env)
    assert r.returncode == 0


@pytest.mark.gpu
@pytest.mark.parametrize("impl", ["cpu", "cpu_mt"])
@pytest.mark.parametrize("args", [
    ("reduction", 1_000_003, 2),
    ("sdot", 1_000_000, 2),
    ("sgemv", 1024, 512, 2, 1, "row"),
    ("sgemv", 600, 300, 2, 1, "col"),
    ("dgemm", 256, 192, 128, 2),
    ("hgemm", 256, 136, 192, 2),
    ("w4a16", 16, 512, 1024, 2),
    ("w4a16", 3, 256, 512, 2, 1, "sym"),
])
def test_cli_host_implementations(args, impl):
    # the reference's CPU path registered as an implementation and run through
    # the routine's own runner: its result passes the same check
    routine, *rest = args
    r = run("test", routine, impl, *rest)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "# host implementation: " in r.stdout and "host cores" in r.stdout and "CPU:" in r.stdout
    assert "# GFLOP/s" in r.stdout


@pytest.mark.gpu
def test_cli_bench_mode_writes_reference_csv(tmp_path):
    # bench mode runs the routine's SIZES and writes bench_<routine>_<impl>.csv
    # (routines.h:157-158) with the reference's columns (reduction.cpp:132,
    # w4a16.cpp:343)
    cases = {("reduction", "hip"): ("size,gflops", 8),
             ("w4a16", "hip"): ("M,N,K,time_us,gflops,gbps,pct_peak_bw,cublas_time_us,speedup_vs_cublas", 13)}
    for (routine, impl), (header, rows) in cases.items():
        print(f"libthom bench {routine} {impl}", flush=True)
        r = run("bench", routine, impl, timeout=600, cwd=str(tmp_path))
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        lines = (tmp_path / f"bench_{routine}_{impl}.csv").read_text().splitlines()
        assert lines[0] == header and len(lines) == rows + 1, lines[:3]
        for ln in lines[1:]:
            vals = [float(v) for v in ln.split(",")]
            assert len(vals) == len(header.split(",")) and all(v > 0 for v in vals), ln
