"""The C-ABI libraries load (no GPU needed) and export every entry point that
include/tea_leaf_engine.h declares; the driver binds exactly the 29 kernel
entry points of the reference's Fortran interface."""
from __future__ import annotations

import ctypes
import re

import pytest

from conftest import REPO
from tealeaf_cuda_amd import ABI_SYMBOLS, EXTENSION_SYMBOLS
from tealeaf_cuda_amd.runtime import DEVICE_SYMBOLS
from tealeaf_cuda_amd.build import DRIVER_SO, ENGINE_SO

HEADER = REPO / "include" / "tea_leaf_engine.h"


def declared():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.findall(r"^\s*(?:void|int|void\s*\*)\s*\*?\s*(\w+)\s*\(", text, flags=re.M)


def test_header_declares_the_reference_abi():
    names = declared()
    fortran = [n for n in names if n.endswith("_")]
    assert len(fortran) == 29
    assert set(fortran) == set(ABI_SYMBOLS)
    assert set(names) - set(fortran) == set(EXTENSION_SYMBOLS) | set(DEVICE_SYMBOLS)


def test_driver_binds_the_same_29_symbols():
    inc = (REPO / "tealeaf_cuda_amd" / "csrc" / "tea_abi_list.inc").read_text()
    assert set(re.findall(r"^TEA_FN\((\w+),", inc, flags=re.M)) == set(ABI_SYMBOLS)


@pytest.mark.parametrize("lib", ["engine", "oracle"])
def test_library_exports_every_declared_symbol(lib, oracle_so):
    path = ENGINE_SO if lib == "engine" else oracle_so
    if lib == "engine" and not path.exists():
        pytest.fail("libtea_engine.so has not been built (run __graft_entry__.build())")
    so = ctypes.CDLL(str(path))
    # the oracle works on host memory: entry points on caller-owned device
    # memory are the HIP library's alone
    want = [n for n in declared() if lib == "engine" or n not in DEVICE_SYMBOLS]
    missing = [n for n in want if not hasattr(so, n)]
    assert not missing


def test_driver_library_loads(driver_built):
    so = ctypes.CDLL(str(DRIVER_SO))
    assert hasattr(so, "tea_driver_run") and hasattr(so, "tea_driver_set_step_callback")


WS_HEADER = REPO / "include" / "tea_warmstart.h"


def test_engine_exports_the_warm_start_interface():
    from tealeaf_cuda_amd.warmstart import WARMSTART_SYMBOLS

    text = re.sub(r"/\*.*?\*/", "", WS_HEADER.read_text(), flags=re.S)
    names = re.findall(r"^\s*(?:const\s+char\s*\*|void\s*\*|void|int)\s*(\w+)\s*\(", text, flags=re.M)
    assert set(names) == set(WARMSTART_SYMBOLS)
    if not ENGINE_SO.exists():
        pytest.fail("libtea_engine.so has not been built (run __graft_entry__.build())")
    so = ctypes.CDLL(str(ENGINE_SO))
    assert not [n for n in names if not hasattr(so, n)]


def test_inner_pass_plan_covers_the_steps():
    """tea_engine_ppcg_inner_plan (no GPU state involved): every plan takes
    1..5 steps per pass, covers the run exactly, and 10 FP64 steps go as
    4, 3, 3 (the cheapest grouping by the measured pass times)."""
    if not ENGINE_SO.exists():
        pytest.fail("libtea_engine.so has not been built (run __graft_entry__.build())")
    so = ctypes.CDLL(str(ENGINE_SO))
    so.tea_engine_ppcg_inner_plan.argtypes = [ctypes.c_int]
    so.tea_engine_ppcg_inner_plan.restype = ctypes.c_int

    def plan(total):
        out, left = [], total
        while left > 0:
            n = so.tea_engine_ppcg_inner_plan(left)
            assert 1 <= n <= min(5, left)
            out.append(n)
            left -= n
        return out

    for total in range(1, 40):
        assert sum(plan(total)) == total
    assert plan(10) == [4, 3, 3]
    assert so.tea_engine_ppcg_inner_plan(0) == 0
