"""Host code under AddressSanitizer + UndefinedBehaviorSanitizer (no GPU):
the oracle's C restatement and the product's host-only sources (template
split and index tables, cost models, R-MAT source, CSC build, hole rule,
color stream, text and binary loaders) run a randomized self-test that also
checks its results against brute force and against each other."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_host_code_under_asan_ubsan(tmp_path):
    out = tmp_path / "build"
    build = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "sanitize"), f"OUT={out}"],
                           capture_output=True, text=True, timeout=300)
    assert build.returncode == 0, build.stderr[-2000:]
    env = dict(os.environ, OMP_NUM_THREADS="2", PGBSC_SANITIZE_TMP=str(tmp_path),
               ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:verify_asan_link_order=0", UBSAN_OPTIONS="print_stacktrace=1")
    run = subprocess.run([str(out / "host_selftest")], capture_output=True, text=True, env=env, timeout=300)
    log = run.stdout + run.stderr
    assert "AddressSanitizer" not in log and "runtime error" not in log, log[-4000:]
    assert run.returncode == 0, log[-4000:]
    assert "host self-test: passed" in run.stdout
