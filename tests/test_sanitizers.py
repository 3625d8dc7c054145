"""Host sanitizer runs (no GPU): the oracle under AddressSanitizer +
UndefinedBehaviorSanitizer over every function (oracle/sanitize_check.c), and
the CLI's argument handling under the same sanitizers (libthom_amd/bin/
libthom_san, the CLI's host code built with -Xarch_host -fsanitize=...)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_oracle_under_asan_ubsan():
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "check-sanitize"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0 and "sanitize ok" in r.stdout, r.stdout[-3000:] + r.stderr[-3000:]
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr


def test_cli_argument_handling_under_asan_ubsan():
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "libthom_amd"), "san"], capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    exe = os.path.join(ROOT, "libthom_amd", "bin", "libthom_san")
    # leak reports would come from the HIP runtime's own allocations: off
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    cases = [
        ((), 1, "at least"),
        (("test", "nosuch"), 1, "Invalid mode or routine"),
        (("explode", "reduction"), 1, None),
        (("test", "dgemm"), 1, "<size3> <nb_iter>"),
        (("test", "sgemv", "nope", 4, 4, 1), 1, "Invalid implementation"),
        (("test", "reduction", "hip", 100, 0), 1, "nb_iter"),
        (("test", "reduction", "hip", -5, 2), 1, "non-negative"),
        (("tune", "hgemm", 1, 2), 1, "Usage"),
        (("bench", "w4a16_grouped", "bogus"), 1, "Usage"),
        (("test", "w4a16_grouped", "graph", 0), 1, None),
        (("test", "hgemm_batched", "bogus"), 1, "Usage"),
        (("bench", "sdot"), 1, "Missing arguments"),
    ]
    for args, rc, text in cases:
        p = subprocess.run([exe, *map(str, args)], capture_output=True, text=True, timeout=60, env=env)
        out = p.stdout + p.stderr
        assert "runtime error" not in out and "AddressSanitizer" not in out, (args, out[-2000:])
        assert p.returncode == rc, (args, p.returncode, out[-1000:])
        if text:
            assert text in out, (args, out[-1000:])
