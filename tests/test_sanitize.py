"""The oracle under AddressSanitizer and UndefinedBehaviorSanitizer: `make -C oracle
sanitize` builds every oracle source with -fsanitize=address,undefined into a
standalone driver (oracle/sanitize_check.cpp) that runs each restated stage
(PCG with no, Jacobi and p-MG preconditioning, the non-blocking solvers, the AMG
base level, the learned smoother's loss and gradient) on small boxes. Any
memory error, leak or undefined behaviour ends it with a nonzero status."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_oracle_clean_under_asan_and_ubsan():
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "-s", "sanitize"], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1",
               OMP_NUM_THREADS="2")
    r = subprocess.run([os.path.join(ROOT, "oracle", "_build", "sanitize_check")], env=env, capture_output=True,
                       text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "SANITIZE CHECK OK" in r.stdout
    assert "runtime error" not in r.stderr
