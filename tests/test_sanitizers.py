"""AddressSanitizer + UndefinedBehaviorSanitizer run of the CPU build of the
decode kernels (ldpc_kernels.hip, ldpc_bp.hip under tests/kernel_emu), the
host helpers (ldpc_host.cpp) and the oracle (oracle/ldpc_oracle.c), all in
one executable built with -fsanitize=address,undefined
-fno-sanitize-recover=all (tests/kernel_emu/sanitize_main.cpp).  The
instrumented build takes minutes, so this test runs the executable when it
has been built (make -C tests/kernel_emu sanitize_check) and skips
otherwise."""
import os
import subprocess

import pytest

from conftest import ROOT

EXE = os.path.join(ROOT, "tests", "kernel_emu", "sanitize_check")


@pytest.mark.skipif(not os.path.exists(EXE),
                    reason="build first: make -C tests/kernel_emu sanitize_check (minutes)")
def test_kernels_host_and_oracle_clean_under_asan_ubsan():
    r = subprocess.run(["make", "-s", "-q", "-C", os.path.dirname(EXE), "sanitize_check"])
    if r.returncode != 0:
        pytest.skip("sanitize_check is older than its sources: rebuild it")
    out = subprocess.run([EXE], capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stderr[-4000:]
    assert "sanitize ok" in out.stdout
