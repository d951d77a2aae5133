"""Shared fixtures.  GPU tests carry @pytest.mark.gpu; everything else runs on the CPU.

The CPU oracle (oracle/liboracle.so) is the checker; the HIP engine
(tealeaf_cuda_amd/libtea_engine.so) is the product under test.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(REPO))

GOLDEN = REPO / "tests" / "golden"
ORACLE_DIR = REPO / "oracle"
# TEA_ORACLE_LIB names another build of the oracle (the sanitizer build of
# oracle/Makefile's `sanitize` target is run this way)
ORACLE_SO = Path(os.environ["TEA_ORACLE_LIB"]) if os.environ.get("TEA_ORACLE_LIB") else ORACLE_DIR / "liboracle.so"
REF_ORACLE = ORACLE_DIR / "_ref" / "tea_leaf_ref_oracle"
REF_HIP = ORACLE_DIR / "_ref" / "tea_leaf_ref_hip"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


@pytest.fixture(scope="session")
def oracle_so() -> Path:
    """The oracle library; built from oracle/tea_oracle.c if it is not there yet."""
    if not ORACLE_SO.exists():
        subprocess.run(["make", "-s", "-C", str(ORACLE_DIR), "oracle"], check=True)
    return ORACLE_SO


@pytest.fixture(scope="session")
def driver_built():
    from tealeaf_cuda_amd import build

    if not build.DRIVER_SO.exists():
        build.build_driver()
    return build.DRIVER_SO


@pytest.fixture
def oracle_env(monkeypatch):
    """Returns a setter for the oracle's reduction mode (TEA_ORACLE_DOT)."""

    def set_mode(mode: str, threads: int | None = None):
        monkeypatch.setenv("TEA_ORACLE_DOT", mode)
        if threads is not None:
            monkeypatch.setenv("OMP_NUM_THREADS", str(threads))

    return set_mode


def run_in_dir(exe: Path, deck_text: str, workdir: Path, env=None, timeout=600):
    """Runs a reference-driver executable in workdir with tea.in = deck_text."""
    workdir.mkdir(parents=True, exist_ok=True)
    (workdir / "tea.in").write_text(deck_text)
    e = dict(os.environ)
    if env:
        e.update(env)
    return subprocess.run([str(exe)], cwd=workdir, env=e, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                          text=True, timeout=timeout)
