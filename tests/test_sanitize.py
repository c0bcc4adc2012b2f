"""AddressSanitizer + UndefinedBehaviorSanitizer builds (CPU only): the
oracle driven through every stage (oracle/sanitize_main.cpp) and the C ABI's
host code driven through its device-free paths
(ek-pnp_amd/csrc/ekpnp_hostcheck.cpp).  Either binary aborts on a sanitizer
finding."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("where,binary,ok", [("oracle", "_san/oracle_san", "oracle sanitize run ok"),
                                             ("ek-pnp_amd", "_san/ekpnp_hostcheck", "hostcheck ok")])
def test_sanitizer_build_runs_clean(where, binary, ok):
    d = os.path.join(ROOT, where)
    subprocess.run(["make", "-s", "-C", d, "sanitize"], check=True, capture_output=True, timeout=900)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([os.path.join(d, binary)], capture_output=True, text=True, timeout=600, env=env)
    print(r.stdout, r.stderr[-2000:])
    assert r.returncode == 0, r.stderr[-2000:]
    assert ok in r.stdout
