"""AddressSanitizer + UndefinedBehaviorSanitizer runs (no GPU): the CPU oracle driven over
every shape, mode, marching cubes, bake, material and post-processing entry point, and
libupf's host-side ABI code (spec validation, layer set-up, flop tables, status strings)
driven with valid and corrupted specs.  Both are built by build() (tests/sanitize/
Makefile); any invalid access or undefined operation aborts the run."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "sanitize", "_build")


@pytest.mark.parametrize("prog", ["oracle_check", "abi_check"])
def test_sanitized_run_is_clean(prog):
    exe = os.path.join(BUILD, prog)
    if not os.path.exists(exe):
        pytest.fail(f"{exe} missing: run build() (make -C tests/sanitize)")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"{prog}: ok" in r.stdout
