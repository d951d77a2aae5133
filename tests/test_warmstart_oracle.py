"""Pins the warm-start network's CPU oracle (oracle/cnn_oracle.py):

* the vectorised float64 forward pass equals a direct per-pixel loop;
* its loss gradient equals central finite differences in every parameter
  block (W1, b1, each hidden W and b, W6, b6);
* the bfloat16 rounding is round-to-nearest-even on known values;
* the parameter layout round-trips and has the documented size.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

from conftest import REPO

sys.path.insert(0, str(REPO / "oracle"))
import cnn_oracle as co  # noqa: E402


def _random(seed, B=2, H=5, W=6, scale=0.3):
    rng = np.random.default_rng(seed)
    p = rng.standard_normal(co.NPARAM) * scale
    s = rng.standard_normal((B, 6, H, W))
    s[:, 4:] = rng.uniform(0.1, 3.0, (B, 2, H, W))      # face coefficients
    return p, s


def test_param_layout():
    assert co.NPARAM == 38177
    p, _ = _random(0)
    assert np.array_equal(co.pack(*co.unpack(p)), p)
    ip = co.init_params(3)
    assert ip.shape == (co.NPARAM,) and ip.dtype == np.float32
    assert not ip[co.OFF_W6:].any()


def test_forward_equals_direct_loops():
    p, s = _random(1)
    out, _ = co.forward(p, s)
    ref = co.forward_loop(p, s)
    np.testing.assert_allclose(out, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


def _blocks():
    b = [("w1", co.OFF_W1, co.OFF_B1), ("b1", co.OFF_B1, co.OFF_HID)]
    for l in range(4):
        o = co.OFF_HID + l * co.HID
        b += [(f"w{l + 2}", o, o + co.TAPS * co.CH * co.CH), (f"b{l + 2}", o + co.TAPS * co.CH * co.CH, o + co.HID)]
    b += [("w6", co.OFF_W6, co.OFF_B6), ("b6", co.OFF_B6, co.NPARAM)]
    return b


@pytest.mark.parametrize("loss_kind", ["mse", "energy"])
@pytest.mark.parametrize("name,lo,hi", _blocks())
def test_gradient_equals_finite_differences(name, lo, hi, loss_kind):
    p, s = _random(2, B=2, H=6, W=7)
    loss, grad = co.loss_and_grad(p, s, margin=1, loss=loss_kind)
    rng = np.random.default_rng(lo)
    idx = rng.choice(np.arange(lo, hi), size=min(6, hi - lo), replace=False)
    h = 1e-6
    for i in idx:
        q = p.copy()
        q[i] += h
        lp, _ = co.loss_and_grad(q, s, margin=1, loss=loss_kind)
        q[i] -= 2 * h
        lm, _ = co.loss_and_grad(q, s, margin=1, loss=loss_kind)
        fd = (lp - lm) / (2 * h)
        assert grad[i] == pytest.approx(fd, rel=1e-5, abs=1e-9 * max(1.0, abs(loss))), (name, i)


def test_bf16_rounds_to_nearest_even():
    vals = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -8 + 2 ** -20, -2.5, 3.0e-3, 0.0], dtype=np.float32)
    expect = [1.0, 1.0, 1 + 2 ** -6, 1 + 2 ** -7, -2.5, 0.0030059814453125, 0.0]
    assert co.bf16(vals).tolist() == expect


def test_bf16_mode_is_close_to_exact():
    p, s = _random(4, scale=0.1)
    out, _ = co.forward(p, s)
    outb, _ = co.forward(p, s, mode="bf16")
    assert np.linalg.norm(outb - out) <= 0.05 * np.linalg.norm(out)


def test_energy_is_the_operator_norm_by_brute_force():
    """e (A_M e) summed equals sum e^2 + sum over interior faces k (e_i - e_j)^2,
    enumerated face by face, and A_M is symmetric."""
    rng = np.random.default_rng(9)
    B, H, W, m = 1, 7, 8, 1
    e = rng.standard_normal((B, H, W))
    kx, ky = rng.uniform(0.1, 2.0, (2, B, H, W))
    inside = lambda y, x: m <= y < H - m and m <= x < W - m
    brute = 0.0
    for y in range(H):
        for x in range(W):
            if not inside(y, x):
                continue
            brute += e[0, y, x] ** 2
            if inside(y, x - 1):
                brute += kx[0, y, x] * (e[0, y, x] - e[0, y, x - 1]) ** 2
            if inside(y - 1, x):
                brute += ky[0, y, x] * (e[0, y, x] - e[0, y - 1, x]) ** 2
    assert float((e * co.masked_operator(e, kx, ky, m)).sum()) == pytest.approx(brute, rel=1e-12)
    n = H * W
    A = np.zeros((n, n))
    for j in range(n):
        v = np.zeros((B, H, W))
        v.flat[j] = 1.0
        A[:, j] = co.masked_operator(v, kx, ky, m).ravel()
    np.testing.assert_allclose(A, A.T, atol=1e-14)
