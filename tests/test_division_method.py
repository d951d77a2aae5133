"""The FP32 block-Jacobi sweep of the streaming inner pass forms its first
quotient dp0 = rhs0 / diag0 (tea_block_jacobi.cuknl:34) as
(float)(rhs0 * rd) with rd = 1 / (double)diag0 rounded to double once per
strip (strip_apply_rd in tealeaf_cuda_amd/csrc/engine_solver.hpp).  That is
the IEEE float quotient bit for bit: the exact quotient of two floats is at
least about 2^-49 (relative) from every midpoint between floats unless it is
a float itself, and rhs0 * rd is within 2^-52 of it.  Checked here against
numpy's float32 division on random operands over the whole range the sweep
sees (diagonals >= 1, residuals of any sign and magnitude) and on operands
built to put the quotient next to a midpoint."""
from __future__ import annotations

import numpy as np


def _check(a, b):
    ref = a / b
    got = (a.astype(np.float64) * (1.0 / b.astype(np.float64))).astype(np.float32)
    same = (ref.view(np.int32) == got.view(np.int32)) | (np.isnan(ref) & np.isnan(got))
    return int((~same).sum())


def test_reciprocal_quotient_is_the_ieee_quotient():
    rng = np.random.default_rng(7)
    n = 2_000_000
    bad = 0
    for _ in range(4):
        b = (1.0 + 10.0 ** rng.uniform(-8, 7, n)).astype(np.float32)
        a = (rng.choice([-1.0, 1.0], n) * 10.0 ** rng.uniform(-44, 30, n)).astype(np.float32)
        bad += _check(a, b)
        # quotients next to midpoints: a = fl(b * (m + 1/2) 2^e) for 24-bit m
        m = rng.integers(1 << 23, 1 << 24, n).astype(np.float64)
        a = (b.astype(np.float64) * (m + 0.5) * 2.0 ** rng.integers(-60, 10, n)).astype(np.float32)
        bad += _check(a, b)
    assert bad == 0


def test_exact_quotients_and_signs():
    b = np.array([1.0, 2.0, 3.0, 1.5, 4096.0, 1e6], dtype=np.float32)
    a = np.array([3.0, -7.0, 9.0, -4.5, 1.0, -0.0], dtype=np.float32)
    assert _check(a, b) == 0
    assert _check(np.float32([1e-45, -1e-40]), np.float32([1.0, 3.0])) == 0   # subnormal quotients
