"""The host driver's Chebyshev iterations in runs (tea_engine_cheby_iterate_n
between norm checks) against one tea_leaf_cheby_iterate_kernel_cuda_ call per
iteration (TEA_CHEBY_PASSES=0), over the oracle: the same iteration counts,
field summaries, conduction errors and, with verbose_on, the same residual
after every iteration (tea_solve.f90:409-413).  Between norm checks the loop
only takes the square root of its error again (tea_solve.f90:407-416); a run
ends at the iteration after which that test stops the loop, which the decks
with a loose tl_eps exercise.  On the GPU the runs are streaming passes of
several iterations (tests/test_gpu_parity.py compares them with the oracle).
"""
from __future__ import annotations

import pytest

from tealeaf_cuda_amd import Deck, parse_tea_out, random_layout, run_deck

DECKS = {
    "none": dict(preconditioner="none"),
    "diag": dict(preconditioner="jac_diag"),
    "block": dict(preconditioner="jac_block"),
    "block_ragged": dict(preconditioner="jac_block", y_cells=37),
    "loose_eps": dict(preconditioner="jac_diag", eps=0.05),
    "looser_eps": dict(preconditioner="none", eps=0.5),
    "random_layout": dict(preconditioner="jac_diag", states=random_layout(11, cells=48)),
    "fp32_restart": dict(preconditioner="jac_diag", inner_precision="fp32", inner_eps=1e-4),
    "few_iters": dict(preconditioner="none", max_iters=37),
    # long runs of Chebyshev iterations between the norm checks (est_itc large)
    "long_none": dict(preconditioner="none", x_cells=96, y_cells=96, eps=1e-13, ch_cg_presteps=8, end_step=2),
    "long_block": dict(preconditioner="jac_block", x_cells=96, y_cells=88, eps=1e-13, ch_cg_presteps=8,
                       end_step=2),
}


def _deck(**kw):
    base = dict(x_cells=48, y_cells=40, end_step=3, eps=1e-10, solver="chebyshev", ch_cg_presteps=20,
                check_result=True, summary_frequency=1, verbose=True)
    base.update(kw)
    return Deck(**base)


@pytest.mark.parametrize("name", sorted(DECKS))
def test_runs_of_iterations_give_the_same_output(name, oracle_so, driver_built, oracle_env, monkeypatch, tmp_path):
    oracle_env("dot2")
    path = _deck(**DECKS[name]).write(tmp_path / "tea.in")
    monkeypatch.setenv("TEA_CHEBY_PASSES", "0")
    one = run_deck(path, backend=oracle_so, out_path=tmp_path / "one.out")
    monkeypatch.delenv("TEA_CHEBY_PASSES")
    runs = run_deck(path, backend=oracle_so, out_path=tmp_path / "runs.out")
    a, b = parse_tea_out(tmp_path / "one.out"), parse_tea_out(tmp_path / "runs.out")
    assert runs.iterations == one.iterations
    assert [s.values() for s in runs.summaries] == [s.values() for s in one.summaries]
    assert b["summaries"] == a["summaries"]
    assert b["conduction_error"] == a["conduction_error"]
    assert b["residuals"] == a["residuals"] and len(a["residuals"]) > 0
