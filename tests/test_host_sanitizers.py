"""The level-0 parsers and the FP64 host geometry built with AddressSanitizer and UndefinedBehaviorSanitizer
(host code only) and driven with seeded truncated, random and adversarial inputs (asar-focus_amd/sanitize/
host_fuzz.cpp): packets whose length fields point past their buffers, product headers with unbalanced quotes and
overlong digit strings, data sets with huge record counts, degenerate orbits. Any out-of-bounds access, signed
overflow or crash fails the run; inputs may be rejected with an error status."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "asar-focus_amd")


def test_host_code_under_address_and_undefined_behaviour_sanitizers():
    r = subprocess.run(["make", "-s", "-C", PKG, "build/sanitize/host_fuzz"], capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([os.path.join(PKG, "build", "sanitize", "host_fuzz"), "5"], capture_output=True, text=True,
                       timeout=600, env=env)
    assert r.returncode == 0 and "host_fuzz: done" in r.stdout, (r.stdout[-1000:], r.stderr[-3000:])
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-3000:]
