"""Pins the C++ host driver against the reference's own Fortran driver.

oracle/_ref/tea_leaf_ref_oracle is the reference program (tea_leaf.f90,
diffuse.f90, tea_solve.f90, read_input.f90, ... compiled unchanged from the
reference sources by oracle/Makefile) linked against the oracle kernel
library.  Running the same deck through it and through
tealeaf_cuda_amd/csrc/tea_driver.cpp over the same oracle must give the same
tea.out numbers to every printed digit: any difference in the solver control,
the eigenvalue estimate, the Chebyshev/PPCG coefficients or the call sequence
would change them.

Skipped where the reference build is absent (it needs the reference sources).
"""
from __future__ import annotations

import pytest

from conftest import REF_ORACLE, run_in_dir
from tealeaf_cuda_amd import Deck, parse_tea_out, random_layout, run_deck

pytestmark = pytest.mark.skipif(not REF_ORACLE.exists(), reason="reference driver build (oracle/_ref) absent")

DECKS = {
    "cg_none": dict(solver="cg", preconditioner="none", eps=1e-15),
    "cg_diag": dict(solver="cg", preconditioner="jac_diag", eps=1e-12),
    "cg_block": dict(solver="cg", preconditioner="jac_block", eps=1e-12),
    "ppcg_none": dict(solver="ppcg", preconditioner="none", ppcg_inner_steps=10, ch_cg_presteps=20),
    "ppcg_diag_h2": dict(solver="ppcg", preconditioner="jac_diag", ppcg_inner_steps=7, halo_depth=2),
    "ppcg_block": dict(solver="ppcg", preconditioner="jac_block", ppcg_inner_steps=10, ch_cg_presteps=20),
    "ppcg_none_h3": dict(solver="ppcg", preconditioner="none", ppcg_inner_steps=8, halo_depth=3),
    "jacobi": dict(solver="jacobi", preconditioner="none", eps=1e-6, max_iters=400),
    "cg_reflective": dict(solver="cg", preconditioner="jac_diag", reflective_boundary=True),
    "ppcg_reflective": dict(solver="ppcg", preconditioner="jac_diag", reflective_boundary=True,
                            ppcg_inner_steps=6),
    "cg_inverse_density": dict(solver="cg", preconditioner="none", inverse_density=True),
    "ppcg_random_layout": dict(solver="ppcg", preconditioner="jac_block", ppcg_inner_steps=10,
                               states=random_layout(7, cells=48)),
    # verbose_on: the residual after every iteration (tea_solve.f90:409-413)
    "ppcg_block_verbose": dict(solver="ppcg", preconditioner="jac_block", ppcg_inner_steps=10, verbose=True),
    "cg_diag_verbose": dict(solver="cg", preconditioner="jac_diag", eps=1e-12, verbose=True),
    "ppcg_none_h2_verbose": dict(solver="ppcg", preconditioner="none", ppcg_inner_steps=6, halo_depth=2,
                                 verbose=True),
}


def _deck(**kw):
    base = dict(x_cells=48, y_cells=40, end_step=4, eps=1e-10, check_result=True, summary_frequency=2)
    base.update(kw)
    return Deck(**base)


@pytest.mark.parametrize("name", sorted(DECKS))
def test_same_output_as_reference_driver(name, oracle_so, driver_built, tmp_path):
    deck = _deck(**DECKS[name])
    ref = run_in_dir(REF_ORACLE, deck.text(), tmp_path / "ref", env={"TEA_ORACLE_DOT": "dot2"})
    assert ref.returncode == 0, ref.stderr[-2000:]
    ref_out = parse_tea_out(tmp_path / "ref" / "tea.out")

    deck_path = deck.write(tmp_path / "tea.in")
    ours = run_deck(deck_path, backend=oracle_so, out_path=tmp_path / "ours.out")
    our_out = parse_tea_out(tmp_path / "ours.out")

    assert ours.iterations == ref_out["iterations"]
    assert len(our_out["summaries"]) == len(ref_out["summaries"])
    for (s1, v1), (s2, v2) in zip(our_out["summaries"], ref_out["summaries"]):
        assert s1 == s2
        assert v1 == v2  # 17 significant digits, bit-for-bit after printing
    assert our_out["conduction_error"] == ref_out["conduction_error"]
    assert our_out["residuals"] == ref_out["residuals"]
    assert (len(ref_out["residuals"]) > 0) == bool(deck.verbose)


def test_reference_driver_chebyshev_aborts(oracle_so, tmp_path):
    """tea_solve.f90:224-247 runs the PPCG re-initialisation for Chebyshev too;
    the kernel layer registers those kernels for PPCG only, so the reference
    program dies there.  The oracle reproduces that error."""
    deck = _deck(solver="chebyshev", preconditioner="none", ch_cg_presteps=10)
    ref = run_in_dir(REF_ORACLE, deck.text(), tmp_path / "ref")
    assert ref.returncode != 0
    assert "kernel not registered" in ref.stderr
