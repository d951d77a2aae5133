"""Full-size runs (4096^2 and 8192^2) checked through properties that do not
depend on the mesh size; the oracle is too slow at these sizes, and the
small-grid parity tests pin the arithmetic bit for bit.

* Every solver reaches the same solution: CG with the diagonal Jacobi
  preconditioner and Chebyshev without a preconditioner (20 CG pre-steps,
  tl_ch_cg_epslim at its default 1.0, so the switch waits for r.r <= 1) agree
  with plain FP64 PPCG (block Jacobi, stopped on the FP64 true residual) on
  the reference's rectangle layout at 4096^2.
* Zero-flux boundaries: A = I + L with L's columns summing to zero, so
  sum(u) is preserved up to the solve's residual; volume and mass depend on
  the layout only.
* FP32 refinement with the learned warm start on a seeded random layout at
  4096^2 and 8192^2 reaches ||u0 - A u||_2 < 1e-10 ||u0 - A u^n||_2 in every
  step and agrees with the FP64 solve of the same deck.

Tolerances: sum(u) changes by sum(r), r = u0 - A u the final residual, so the
internal-energy drift is allowed 100 times the solve's true relative residual
(CG jac_diag stops on its preconditioned recurrence at a true residual of
~2e-8 and drifts ~2e-7); the temperature sum weights the solution error by
density and is held to 1e-6 between solvers.
"""
from __future__ import annotations

import pytest
import torch  # noqa: F401  (first: torch's HIP runtime must be the process's one)

from conftest import REPO
from tealeaf_cuda_amd import Deck, random_layout, run_deck

pytestmark = pytest.mark.gpu

WEIGHTS_4096 = REPO / "tealeaf_cuda_amd" / "weights" / "warmstart_4096.tlws"


def _deck(n, steps, **kw):
    base = dict(x_cells=n, y_cells=n, end_step=steps, eps=1e-10, max_iters=10000, halo_depth=1,
                summary_frequency=1, check_result=True)
    base.update(kw)
    return Deck(**base)


def _rel(a, b):
    return abs(a - b) / abs(b)


def _close(a, b, rel):
    return _rel(a, b) <= rel


@pytest.fixture(scope="module")
def ppcg_4096(tmp_path_factory):
    print("[gpu fullsize] 4096^2 rectangles, PPCG jac_block FP64, true-residual stop", flush=True)
    d = _deck(4096, 1, solver="ppcg", preconditioner="jac_block", ppcg_inner_steps=10, inner_precision="fp64")
    return run_deck(d.write(tmp_path_factory.mktemp("ppcg") / "tea.in"))


@pytest.mark.parametrize("solver,prec", [("cg", "jac_diag"), ("chebyshev", "none")])
def test_4096_solvers_agree(solver, prec, ppcg_4096, tmp_path):
    print(f"[gpu fullsize] 4096^2 rectangles, {solver} {prec}", flush=True)
    kw = dict(ch_cg_presteps=20) if solver == "chebyshev" else {}
    res = run_deck(_deck(4096, 1, solver=solver, preconditioner=prec, **kw).write(tmp_path / "tea.in"))
    ref = ppcg_4096
    assert 0 < res.iterations[0] < 10000
    assert 0.0 < res.exact_error[0] < 1e-7
    assert ref.exact_error[0] < 1e-10
    s0, s1, r1 = res.summaries[0], res.summaries[-1], ref.summaries[-1]
    print(f"[gpu fullsize] iterations {res.iterations} residual {res.exact_error} energy drift "
          f"{_rel(s1.energy, s0.energy):.2e}, vs PPCG: energy {_rel(s1.energy, r1.energy):.2e} "
          f"temperature {_rel(s1.u, r1.u):.2e}", flush=True)
    assert s1.volume == r1.volume and s1.mass == r1.mass
    assert _close(s1.energy, s0.energy, 100 * res.exact_error[0])      # conservation
    assert _close(s1.energy, r1.energy, 100 * res.exact_error[0])
    assert _close(r1.energy, s0.energy, 100 * ref.exact_error[0])
    assert _close(s1.u, r1.u, 1e-6)


@pytest.mark.parametrize("n,seed", [(4096, 4096), (8192, 8192)])
def test_fp32_learned_warm_start_reaches_fp64_target(n, seed, tmp_path):
    steps = 3
    layout = random_layout(seed, cells=n)
    common = dict(solver="ppcg", preconditioner="jac_block", ppcg_inner_steps=10, states=layout)
    print(f"[gpu fullsize] {n}^2 layout {seed}, FP32 + learned warm start, {steps} steps", flush=True)
    fast = run_deck(_deck(n, steps, inner_precision="fp32", warm_start=str(WEIGHTS_4096), **common)
                    .write(tmp_path / "fast.in"))
    print(f"[gpu fullsize] {n}^2 layout {seed}, FP64 from u^n", flush=True)
    plain = run_deck(_deck(n, steps, inner_precision="fp64", **common).write(tmp_path / "plain.in"))
    for r in (fast, plain):
        print(f"[gpu fullsize] iterations {r.iterations} residual {r.exact_error} energy drift "
              f"{_rel(r.summaries[-1].energy, r.summaries[0].energy):.2e}", flush=True)
    print("[gpu fullsize] fast vs plain: energy", [f"{_rel(a.energy, b.energy):.2e}" for a, b in
                                                 zip(fast.summaries, plain.summaries)],
          "temperature", [f"{_rel(a.u, b.u):.2e}" for a, b in zip(fast.summaries, plain.summaries)], flush=True)
    assert fast.warm_start == [0] + [1] * (steps - 1)
    for r in (fast, plain):
        assert all(0.0 < e < 1e-10 for e in r.exact_error), r.exact_error
        assert all(0 < i < 10000 for i in r.iterations)
        assert _close(r.summaries[-1].energy, r.summaries[0].energy, 100 * steps * 1e-10)
    for a, b in zip(fast.summaries, plain.summaries):
        assert a.volume == b.volume and a.mass == b.mass
        assert _close(a.energy, b.energy, 100 * steps * 1e-10)
        assert _close(a.u, b.u, 1e-6)
