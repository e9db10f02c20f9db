"""The oracle under AddressSanitizer and UndefinedBehaviorSanitizer (host
only): oracle/sanitize/driver.cpp drives every constraint family, collider
shape, attachment, press, reset/launch, deformed shapes, self-collision and
both Gauss-Seidel colourings, the volume-cluster rules and the argument
checks of the C ABI on small scenes; any memory or undefined-behaviour error
aborts the run."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "3dgs-isaac-sim-xpbd_amd"))
import build_lib  # noqa: E402


@pytest.mark.timeout(600)
def test_oracle_sanitized_run():
    exe = build_lib.build_oracle_asan()
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=500, env=env)
    report = r.stdout + r.stderr
    assert "AddressSanitizer" not in report and "runtime error" not in report, report[-4000:]
    assert r.returncode == 0, report[-4000:]
    assert "oracle sanitizer run: ok" in r.stdout
    line = next(l for l in r.stdout.splitlines() if l.startswith("full scene"))
    cand = int(line.split("candidates")[1].split()[0])
    assert cand > 0, line
    print(line)
