"""How much any predicted direction can add to the warm start's history span.

The warm start (tea_engine_warm_start_apply) combines the differences of the
last 8 steps' solutions, u0 + sum c_a v_a, with the coefficients that
minimise the residual's 2-norm; the trained network adds one more direction
to that span.  This test measures, on the CPU, what the best directions a
predictor could hand over are worth on the same problem, in iterations of
Jacobi-preconditioned CG (scipy) to the solver's stop: ||u0 - A u|| below
1e-10 of the residual of the u = u^n guess.

The problem is the 1024^2 learned-warm-start configuration's physics on a
256^2 mesh: held-out random-rectangle layouts (seeds 10000 and 10001) with
dt scaled by (1024/256)^2 = 16 so that dt K / dx^2, the operator's coupling per
cell, is that of 1024^2 at dt 0.004.  The solves that make the history run
through the oracle (plain FP64 PPCG, block Jacobi, 10 inner steps).

Directions added to the 8-step history span, at each step from step 10 on:
* the exact two-level correction P (P^T A P)^-1 P^T r_g over 4x4, 8x8 and
  16x16 aggregates, with two damped Jacobi sweeps (r_g the residual of the
  history guess): the A-orthogonal projection of the guess's error onto a
  coarse space, the best that a predictor whose output lives on a coarse
  grid (an encoder-decoder's coarse levels) can give;
* the exact remaining error u* - u_g with independent per-cell relative
  noise of 0.1%, 0.4% (about a BF16 rounding) and 3%: a fine-grid
  predictor of that accuracy.

Measured (totals over steps 10-16, both seeds): the coarse directions save
1-6% of the history guess's iterations and the 3%-accurate prediction 3%,
while 0.4% saves 28% and 0.1% 52%.  A direction is worth iterations only
when it predicts the guess's remaining error to well under 1% per cell.
"""
from __future__ import annotations

import os
import tempfile

import numpy as np
import pytest

sp = pytest.importorskip("scipy.sparse")
spl = pytest.importorskip("scipy.sparse.linalg")

from tealeaf_cuda_amd import Engine, run_deck  # noqa: E402
from tealeaf_cuda_amd import runtime as rt  # noqa: E402
from tealeaf_cuda_amd.warmstart import solver_deck  # noqa: E402

N_CELLS = 256
LIKE = 1024
HISTORY = 8
STEPS = 16


def _history(oracle_so, seed):
    """u after each step of the oracle's plain FP64 PPCG solve, and the
    operator's face coefficients."""
    dt = 0.004 * (LIKE / N_CELLS) ** 2
    eng = Engine(str(oracle_so))
    caps, caps_k = [], []

    def on_step(step):
        eng.refresh()
        h = eng.h
        caps.append(eng.get_field(rt.FIELD_U)[h:h + N_CELLS, h:h + N_CELLS].copy())
        if step == 1:
            caps_k.append((eng.get_field(rt.XFIELD_KX), eng.get_field(rt.XFIELD_KY), h))

    with tempfile.TemporaryDirectory() as d:
        deck = solver_deck(N_CELLS, seed, STEPS, eps=1e-12, initial_timestep=dt).write(os.path.join(d, "tea.in"))
        run_deck(deck, backend=str(oracle_so), on_step=on_step)
    return caps, caps_k[0], dt


def _operator(kxf, kyf, h):
    n = N_CELLS
    kx = kxf[h:h + n, h:h + n + 1]
    ky = kyf[h:h + n + 1, h:h + n]
    diag = 1.0 + kx[:, :-1] + kx[:, 1:] + ky[:-1, :] + ky[1:, :]
    ix = np.arange(n * n).reshape(n, n)
    rows = [ix.ravel(), ix[:, 1:].ravel(), ix[:, :-1].ravel(), ix[1:, :].ravel(), ix[:-1, :].ravel()]
    cols = [ix.ravel(), ix[:, :-1].ravel(), ix[:, 1:].ravel(), ix[:-1, :].ravel(), ix[1:, :].ravel()]
    vals = [diag.ravel(), -kx[:, 1:-1].ravel(), -kx[:, 1:-1].ravel(), -ky[1:-1, :].ravel(), -ky[1:-1, :].ravel()]
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n * n, n * n))


def _ceiling(oracle_so, seed):
    us, (kxf, kyf, h), dt = _history(oracle_so, seed)
    n2 = N_CELLS * N_CELLS
    A = _operator(kxf, kyf, h)
    dinv = sp.diags(1.0 / A.diagonal())

    def iterations(b, x0, tol):
        count = [0]

        def cb(_):
            count[0] += 1
        spl.cg(A, b, x0=x0, rtol=0, atol=tol, M=dinv, maxiter=50000, callback=cb)
        return count[0]

    def least_squares(x0, b, dirs):
        D = np.stack(dirs, 1)
        c = np.linalg.lstsq(A @ D, b - A @ x0, rcond=1e-13)[0]
        return x0 + D @ c

    def aggregates(f):
        m = N_CELLS // f
        k = np.arange(n2)
        return sp.csr_matrix((np.ones(n2), (k, ((k // N_CELLS) // f) * m + (k % N_CELLS) // f)), shape=(n2, m * m))

    coarse = {f: aggregates(f) for f in (4, 8, 16)}
    lus = {f: spl.splu((P.T @ A @ P).tocsc()) for f, P in coarse.items()}
    flat = [u.ravel() for u in us]   # flat[k]: u after step k + 1
    total = {}
    rng = np.random.default_rng(seed)
    for k in range(HISTORY + 1, len(flat)):
        b = flat[k - 1]                 # step k + 1 solves A u = u^n, u^n = u after step k
        tol = 1e-10 * np.linalg.norm(b - A @ b)
        hist = [flat[k - 1 - j] - flat[k - 2 - j] for j in range(HISTORY)]
        g = least_squares(b, b, hist)
        row = {"u^n": iterations(b, b, tol), "history": iterations(b, g, tol)}
        rg = b - A @ g
        for f, P in coarse.items():
            ec = P @ lus[f].solve(P.T @ rg)
            for _ in range(2):
                ec = ec + 0.7 * (dinv @ (rg - A @ ec))
            row[f"coarse{f}"] = iterations(b, least_squares(b, b, hist + [ec]), tol)
        err = spl.spsolve(A.tocsc(), b) - g
        for lvl in (0.001, 0.004, 0.03):
            noisy = err * (1.0 + lvl * rng.standard_normal(n2))
            row[f"error{lvl}"] = iterations(b, least_squares(b, b, hist + [noisy]), tol)
        for key, v in row.items():
            total[key] = total.get(key, 0) + v
    return total, dt


def test_predicted_direction_ceiling(oracle_so, oracle_env):
    oracle_env("plain")
    totals = {}
    for seed in (10000, 10001):
        t, dt = _ceiling(oracle_so, seed)
        for key, v in t.items():
            totals[key] = totals.get(key, 0) + v
    gain = {key: round(1.0 - v / totals["history"], 3) for key, v in totals.items()}
    print(f"[warm start ceiling] dt {dt:g}, iterations {totals}, saving over the history guess {gain}", flush=True)
    assert totals["history"] < 0.9 * totals["u^n"]          # the history span itself pays
    for f in (4, 8, 16):
        assert gain[f"coarse{f}"] < 0.10, gain              # no coarse-grid predictor reaches 10%
    assert gain["error0.03"] < 0.10, gain                   # nor a fine one accurate to 3% per cell
    assert gain["error0.001"] > 0.25, gain                  # while the measure does show a real gain
