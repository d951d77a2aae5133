"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer: its
closed-form and dense-matrix pins (test_oracle_recurrences.py,
test_oracle_math.py) run in a child process against the sanitizer build
(oracle/Makefile target `sanitize`) with the sanitizer runtimes preloaded;
any out-of-bounds access, use after free or undefined behaviour aborts the
run."""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

import pytest

from conftest import ORACLE_DIR, REPO


def _runtime(name):
    if not shutil.which("gcc"):
        return None
    path = subprocess.run(["gcc", f"-print-file-name={name}"], stdout=subprocess.PIPE, text=True).stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


@pytest.mark.skipif(_runtime("libasan.so") is None or _runtime("libubsan.so") is None,
                    reason="gcc sanitizer runtimes absent")
def test_oracle_pins_under_sanitizers(tmp_path):
    subprocess.run(["make", "-s", "-C", str(ORACLE_DIR), "sanitize"], check=True)
    env = dict(os.environ)
    env.update({
        "LD_PRELOAD": f"{_runtime('libasan.so')} {_runtime('libubsan.so')}",
        "ASAN_OPTIONS": "detect_leaks=0:abort_on_error=1",
        "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1",
        "TEA_ORACLE_LIB": str(ORACLE_DIR / "_san" / "liboracle.so"),
        "OMP_NUM_THREADS": "2",
    })
    proc = subprocess.run([sys.executable, "-m", "pytest", "-q", "-m", "not gpu", "-p", "no:cacheprovider",
                           "tests/test_oracle_recurrences.py", "tests/test_oracle_math.py"],
                          cwd=REPO, env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                          timeout=900)
    tail = proc.stdout[-3000:]
    assert proc.returncode == 0, tail
    assert "AddressSanitizer" not in proc.stdout and "runtime error" not in proc.stdout, tail
    assert " passed" in tail
