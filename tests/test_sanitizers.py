"""The CPU oracle and the product's host-compiled traversal (BVH builder, closest / any-hit
traversal with stack spilling, guiding-model math, slot order) built with AddressSanitizer and
UndefinedBehaviorSanitizer and run on small scenes (tests/sanitize/): no memory error, leak or
undefined behaviour is reported, and the drivers' own checks pass."""
import os
import subprocess

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sanitize")
ENV = dict(os.environ, ASAN_OPTIONS="halt_on_error=1:detect_leaks=1",
           UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")


def _run(exe):
    subprocess.run(["make", "-s", "-C", HERE], check=True, capture_output=True)
    r = subprocess.run([os.path.join(HERE, "_build", exe)], capture_output=True, text=True, timeout=600, env=ENV)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-2000:]
    assert "AddressSanitizer" not in out and "runtime error" not in out and "LeakSanitizer" not in out, out[-2000:]
    return r.stdout


def test_oracle_under_sanitizers():
    assert "sanitized oracle: ok" in _run("san_oracle")


def test_host_traversal_under_sanitizers():
    assert "sanitized host traversal: ok" in _run("san_hostemu")
