"""GPU parity: the HIP engine against the oracle on the same inputs.

Whole decks go through the same host driver over both kernel libraries.  The
engine's reductions are compensated (order-independent), as are the oracle's
in its default mode, and both evaluate every pointwise expression in the
reference's order without fused multiply-adds, so the bar is bit-for-bit:
identical iteration counts, field summaries, conduction errors and exact
residuals.  (The floating-point tolerance the project states for parity with
the reference, 1e-10 relative, is checked against the golden files in
test_gpu_golden.py.)

Every ABI entry point is also called one at a time on random fields in both
libraries, comparing every field afterwards.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from tealeaf_cuda_amd import Deck, Engine, StateSpec, TeaError, parse_tea_out, random_layout, run_deck
from tealeaf_cuda_amd import runtime as rt

pytestmark = pytest.mark.gpu

LAYOUT = [StateSpec(100.0, 0.0001), StateSpec(0.1, 25.0, "rectangle", 0.0, 1.0, 1.0, 2.0),
          StateSpec(0.1, 0.1, "rectangle", 1.0, 6.0, 1.0, 2.0), StateSpec(0.1, 0.1, "rectangle", 5.0, 6.0, 1.0, 8.0),
          StateSpec(0.1, 0.1, "rectangle", 5.0, 10.0, 7.0, 8.0),
          StateSpec(5.0, 2.0, "circle", 3.0, 0.0, 6.0, 0.0, 1.5), StateSpec(1.0, 9.0, "point", 2.5, 0.0, 2.5, 0.0)]

DECKS = {
    "cg_none_eps15": dict(solver="cg", preconditioner="none", eps=1e-15),
    "cg_diag": dict(solver="cg", preconditioner="jac_diag"),
    "cg_block": dict(solver="cg", preconditioner="jac_block"),
    "cg_block_ragged": dict(solver="cg", preconditioner="jac_block", x_cells=131, y_cells=77),
    "cheby_none": dict(solver="chebyshev", preconditioner="none", ch_cg_presteps=20),
    "cheby_diag": dict(solver="chebyshev", preconditioner="jac_diag", ch_cg_presteps=20),
    "cheby_block": dict(solver="chebyshev", preconditioner="jac_block", ch_cg_presteps=20),
    "ppcg_none": dict(solver="ppcg", preconditioner="none", ppcg_inner_steps=10),
    "ppcg_diag_h2": dict(solver="ppcg", preconditioner="jac_diag", ppcg_inner_steps=9, halo_depth=2),
    "ppcg_block": dict(solver="ppcg", preconditioner="jac_block", ppcg_inner_steps=10),
    "ppcg_block_ragged": dict(solver="ppcg", preconditioner="jac_block", ppcg_inner_steps=10,
                              x_cells=97, y_cells=66),
    "ppcg_none_h4": dict(solver="ppcg", preconditioner="none", ppcg_inner_steps=10, halo_depth=4),
    "jacobi": dict(solver="jacobi", preconditioner="none", eps=1e-7, max_iters=3000),
    "reflective_ppcg": dict(solver="ppcg", preconditioner="jac_diag", reflective_boundary=True,
                            ppcg_inner_steps=8),
    "reflective_cg_h2": dict(solver="cg", preconditioner="jac_block", reflective_boundary=True, halo_depth=2),
    "inverse_density": dict(solver="cg", preconditioner="jac_diag", inverse_density=True),
    "random_layout_ppcg": dict(solver="ppcg", preconditioner="jac_block", ppcg_inner_steps=10,
                               states=random_layout(11, cells=128), x_cells=128, y_cells=128),
    "tiny_3x5": dict(solver="ppcg", preconditioner="jac_block", ppcg_inner_steps=4, x_cells=3, y_cells=5),
    "mixed_ppcg_block": dict(solver="ppcg", preconditioner="jac_block", ppcg_inner_steps=10,
                             inner_precision="fp32"),
    "mixed_ppcg_diag_h2": dict(solver="ppcg", preconditioner="jac_diag", ppcg_inner_steps=9, halo_depth=2,
                               inner_precision="fp32", reliable_delta=0.01),
    "mixed_ppcg_none_ragged_h4": dict(solver="ppcg", preconditioner="none", ppcg_inner_steps=10, halo_depth=4,
                                      x_cells=97, y_cells=66, inner_precision="fp32"),
    "mixed_cg_block": dict(solver="cg", preconditioner="jac_block", inner_precision="fp32", reliable_delta=0.05),
    "mixed_restart_ppcg_block": dict(solver="ppcg", preconditioner="jac_block", ppcg_inner_steps=10,
                                     inner_precision="fp32", mixed_refinement="restart", inner_eps=1e-5),
    "mixed_restart_cg_diag": dict(solver="cg", preconditioner="jac_diag", inner_precision="fp32",
                                  mixed_refinement="restart", inner_eps=1e-4),
    "refined_fp64_ppcg_block": dict(solver="ppcg", preconditioner="jac_block", ppcg_inner_steps=10,
                                    inner_precision="fp64"),
    "mixed_cheby_block": dict(solver="chebyshev", preconditioner="jac_block", ch_cg_presteps=20,
                              inner_precision="fp32", inner_eps=1e-5),
    "mixed_cheby_diag": dict(solver="chebyshev", preconditioner="jac_diag", ch_cg_presteps=20,
                             inner_precision="fp32", inner_eps=1e-4),
    "mixed_random_layout": dict(solver="ppcg", preconditioner="jac_block", ppcg_inner_steps=10,
                                states=random_layout(12, cells=128), x_cells=128, y_cells=128,
                                inner_precision="fp32"),
    # verbose_on prints every iteration's residual, and keeps the refinement
    # loop's outer iteration undeferred (the end is issued on its own)
    "verbose_ppcg_block": dict(solver="ppcg", preconditioner="jac_block", ppcg_inner_steps=10, verbose=True),
    "verbose_mixed_ppcg_block": dict(solver="ppcg", preconditioner="jac_block", ppcg_inner_steps=10,
                                     inner_precision="fp32", verbose=True),
    "verbose_cg_diag": dict(solver="cg", preconditioner="jac_diag", verbose=True),
}


def _deck(**kw):
    base = dict(x_cells=96, y_cells=80, end_step=3, eps=1e-10, check_result=True, summary_frequency=1,
                states=LAYOUT)
    base.update(kw)
    return Deck(**base)


@pytest.mark.parametrize("name", sorted(DECKS))
def test_deck_bitwise(name, oracle_so, driver_built, oracle_env, tmp_path):
    oracle_env("dot2")
    print(f"[gpu parity] deck {name}", flush=True)
    path = _deck(**DECKS[name]).write(tmp_path / "tea.in")
    gpu = run_deck(path, out_path=tmp_path / "gpu.out")
    cpu = run_deck(path, backend=oracle_so, out_path=tmp_path / "cpu.out")
    res_g, res_c = (parse_tea_out(tmp_path / f)["residuals"] for f in ("gpu.out", "cpu.out"))
    assert res_g == res_c and (len(res_g) > 0) == bool(DECKS[name].get("verbose"))
    assert gpu.iterations == cpu.iterations
    assert gpu.ppcg_inner == cpu.ppcg_inner
    assert gpu.refinements == cpu.refinements
    assert gpu.conduction_error == cpu.conduction_error
    assert gpu.exact_error == cpu.exact_error
    assert len(gpu.summaries) == len(cpu.summaries)
    for a, b in zip(gpu.summaries, cpu.summaries):
        assert a.values() == b.values(), (a, b)


@pytest.mark.parametrize("switch", ["TEA_FUSED_OUTER", "TEA_INNER_PASSES", "TEA_INNER_MARCH", "TEA_SPECULATE",
                                    "TEA_W_FROM_P", "TEA_FUSED_CG", "TEA_BEGIN_MARCH", "TEA_FUSED_CHEBY",
                                    "TEA_PPCG_ITERATION", "TEA_FUSED_END", "TEA_CHEBY_PASSES", "TEA_SD_FROM_R",
                                    "TEA_END_STORE_R", "TEA_CG_Z_FROM_R"])
@pytest.mark.parametrize("name", ["ppcg_block", "ppcg_diag_h2", "mixed_ppcg_block", "random_layout_ppcg", "cg_diag",
                                  "cg_block_ragged", "cheby_none", "mixed_cheby_diag"])
def test_deck_bitwise_unfused_outer(name, switch, oracle_so, driver_built, oracle_env, monkeypatch, tmp_path):
    """The driver's separate-call PPCG path (TEA_FUSED_OUTER=0), the fused
    outer iteration with one inner step per pass (TEA_INNER_PASSES=0), the
    inner passes on the tile kernels (TEA_INNER_MARCH=0), the refinement
    loop without inner passes queued ahead of the replacement test
    (TEA_SPECULATE=0), the FP32 fused pair with w written by begin and read
    by ur (TEA_W_FROM_P=0), CG iterations through the separate calls
    instead of the fused one (TEA_FUSED_CG=0), the fused begins on the tile
    kernels (TEA_BEGIN_MARCH=0), Chebyshev iterations on the separate
    calc_p / calc_u kernels (TEA_FUSED_CHEBY=0), and PPCG outer iterations
    through begin_ur, the passes and end as separate calls instead of one
    (TEA_PPCG_ITERATION=0), and the end of an outer iteration as its own
    launch instead of in the last inner pass (TEA_FUSED_END=0), and
    Chebyshev iterations one call each instead of in runs of streaming passes
    (TEA_CHEBY_PASSES=0), and a fused ur that stores z and sd itself instead of
    leaving sd to the first inner pass (TEA_SD_FROM_R=0), and the fused begin
    storing r and forming r.z itself instead of the end before it
    (TEA_END_STORE_R=0), and the fused CG iteration with jac_diag storing z
    instead of leaving it to the next begin (TEA_CG_Z_FROM_R=0), on the GPU
    give the same results as the oracle."""
    oracle_env("dot2")
    monkeypatch.setenv(switch, "0")
    print(f"[gpu parity] unfused deck {name}", flush=True)
    path = _deck(**DECKS[name]).write(tmp_path / "tea.in")
    gpu = run_deck(path)
    cpu = run_deck(path, backend=oracle_so)
    assert gpu.iterations == cpu.iterations and gpu.refinements == cpu.refinements
    assert gpu.exact_error == cpu.exact_error
    for a, b in zip(gpu.summaries, cpu.summaries):
        assert a.values() == b.values(), (a, b)


def test_deck_bitwise_inner_passes_always_queued(oracle_so, driver_built, oracle_env, monkeypatch, tmp_path):
    """TEA_SPECULATE=2: the refinement loop queues the inner passes ahead of
    every replacement test, so every replacement overwrites queued passes;
    the GPU still gives the oracle's results."""
    oracle_env("dot2")
    monkeypatch.setenv("TEA_SPECULATE", "2")
    print("[gpu parity] deck mixed_ppcg_block, inner passes always queued", flush=True)
    path = _deck(**DECKS["mixed_ppcg_block"]).write(tmp_path / "tea.in")
    gpu = run_deck(path)
    cpu = run_deck(path, backend=oracle_so)
    assert gpu.iterations == cpu.iterations and gpu.refinements == cpu.refinements
    assert gpu.exact_error == cpu.exact_error
    for a, b in zip(gpu.summaries, cpu.summaries):
        assert a.values() == b.values(), (a, b)


# ------------------------------------------------------------------ entry points

def _prepare(eng, tmp_path, solver, prec, nx=70, ny=45, halo=2):
    deck = Deck(x_cells=nx, y_cells=ny, solver=solver, preconditioner=prec, halo_depth=halo, states=LAYOUT)
    eng.initialise(deck.write(tmp_path / f"ep_{solver}_{prec}.in"), nx, ny)
    eng.initialise_chunk(0.0, 0.0, 10.0 / nx, 10.0 / ny)
    st = LAYOUT
    eng.generate_chunk([s.density for s in st], [s.energy for s in st], [s.xmin for s in st],
                       [s.xmax for s in st], [s.ymin for s in st], [s.ymax for s in st], [s.radius for s in st],
                       [{"rectangle": 1, "circle": 2, "point": 3}.get(s.geometry or "", 0) for s in st])
    eng.set_field_kernel()
    eng.common_init(rt.CONDUCTIVITY, 0.004)


def _randomise(engines, seed):
    rng = np.random.default_rng(seed)
    shape = engines[0].shape()
    for fid in (rt.FIELD_U, rt.FIELD_P, rt.FIELD_SD, rt.FIELD_R, rt.XFIELD_U0, rt.XFIELD_W, rt.XFIELD_Z,
                rt.XFIELD_RTEMP, rt.XFIELD_UTEMP, rt.XFIELD_RSTORE):
        a = rng.standard_normal(shape)
        for e in engines:
            e.set_field(fid, a)


def _compare(g, c, label):
    for fid in range(1, 20):
        a, b = g.get_field(fid), c.get_field(fid)
        assert np.array_equal(a, b, equal_nan=True), f"{label}: field {fid} differs"


CALLS = {
    "calc_residual": lambda e: e.calc_residual(),
    "calc_2norm_0": lambda e: e.calc_2norm(0),
    "calc_2norm_1": lambda e: e.calc_2norm(1),
    "calc_2norm_2": lambda e: e.calc_2norm(2),
    "finalise": lambda e: e.finalise(),
    "cg_init": lambda e: e.cg_init(),
    "cg_calc_w": lambda e: e.cg_calc_w(),
    "cg_calc_ur": lambda e: e.cg_calc_ur(0.37),
    "cg_calc_p": lambda e: e.cg_calc_p(0.61),
    "field_summary": lambda e: e.field_summary(),
    "fused_cg_begin_ur_p": lambda e: e.cg_begin_ur(True, 0.61, 0.8),
    "fused_cg_begin_ur": lambda e: e.cg_begin_ur(False, 0.0, 0.8),
    # the w owed by the fused iteration is written out before the next call reads it
    "fused_cg_begin_ur_then_norm_then_w": lambda e: (e.cg_begin_ur(True, 0.61, 0.8), e.calc_2norm(1),
                                                     e.cg_calc_w()),
    "fused_cg_begin_ur_twice": lambda e: (e.cg_begin_ur(False, 0.0, 0.8), e.cg_begin_ur(True, 0.3, 1.1)),
    # with jac_diag z = Mi r is owed as well: the next fused begin forms it from r, a separate call
    # (calc_p reads z; the norm of r.z) has it written out first
    "fused_cg_begin_ur_three": lambda e: (e.cg_begin_ur(True, 0.61, 0.8), e.cg_begin_ur(True, 0.3, 1.1),
                                          e.cg_begin_ur(True, 0.45, 0.9)),
    "fused_cg_begin_ur_then_calc_p": lambda e: (e.cg_begin_ur(True, 0.61, 0.8), e.cg_calc_p(0.4)),
    "fused_cg_begin_ur_then_norm_rz": lambda e: (e.cg_begin_ur(True, 0.61, 0.8), e.calc_2norm(2),
                                                 e.cg_begin_ur(True, 0.3, 1.1)),
}
def _tile_kernels(fn):
    """Runs fn with the engine's inner-step passes on the tile kernels
    (TEA_INNER_MARCH=0; the oracle ignores the switch)."""
    def call(e):
        os.environ["TEA_INNER_MARCH"] = "0"
        try:
            return fn(e)
        finally:
            del os.environ["TEA_INNER_MARCH"]
    return call


A10 = [0.3, 0.5, 0.7, 0.2, 0.6, 0.4, 0.35, 0.55, 0.25, 0.45]
B10 = [1.1, 0.9, 0.8, 1.2, 0.7, 1.0, 0.95, 0.85, 1.05, 0.75]
PPCG_CALLS = {
    "ppcg_init_1": lambda e: e.ppcg_init(1),
    "ppcg_init_2": lambda e: e.ppcg_init(2),
    "ppcg_init_3": lambda e: e.ppcg_init(3),
    "ppcg_init_sd": lambda e: e.ppcg_init_sd(1.7),
    "ppcg_init_sd_new": lambda e: e.ppcg_init_sd_new(1.7),
    "ppcg_inner": lambda e: (e.ppcg_init_constants([0.3, 0.5, 0.7], [1.1, 0.9, 0.8]), e.ppcg_inner(2)),
    "ppcg_store_r": lambda e: e.ppcg_store_r(),
    "ppcg_update_z": lambda e: e.ppcg_update_z(),
    "ppcg_calc_rrn": lambda e: e.ppcg_calc_rrn(),
    "ppcg_calc_2norm": lambda e: e.ppcg_calc_2norm(),
    "ppcg_calc_p": lambda e: e.ppcg_calc_p(0.45),
    "fused_ppcg_begin_p": lambda e: e.ppcg_begin(True, 0.45),
    "fused_ppcg_begin": lambda e: e.ppcg_begin(False, 0.0),
    "fused_ppcg_ur": lambda e: e.ppcg_ur(0.37, 1.7),
    "fused_ppcg_end": lambda e: e.ppcg_end(),
    "fused_ppcg_begin_ur_p": lambda e: e.ppcg_begin_ur(True, 0.45, 1.7),
    "fused_ppcg_begin_ur": lambda e: e.ppcg_begin_ur(False, 0.0, 1.7),
    "fused_ppcg_end_begin_ur": lambda e: e.ppcg_end_begin_ur(0.8, 1.7),
    "fused_ppcg_end_then_begin_ur": lambda e: (e.ppcg_end(), e.ppcg_begin_ur(True, 0.45, 1.7)),
    "fused_ppcg_end_then_inner_n2": lambda e: (
        e.ppcg_init_constants([0.3, 0.5, 0.7, 0.2, 0.6, 0.4], [1.1, 0.9, 0.8, 1.2, 0.7, 1.0]),
        e.ppcg_end(), e.ppcg_inner_n(1, 2)),
    "fused_ppcg_end_begin_ur_inner6": lambda e: (
        e.ppcg_init_constants([0.3, 0.5, 0.7, 0.2, 0.6, 0.4], [1.1, 0.9, 0.8, 1.2, 0.7, 1.0]),
        e.ppcg_end_begin_ur_inner(0.8, 1.7, 6)),
    # the last pass of one call forms the next call's end (taken only when no other call comes between)
    "fused_ppcg_end_begin_ur_inner6_twice": lambda e: (
        e.ppcg_init_constants([0.3, 0.5, 0.7, 0.2, 0.6, 0.4], [1.1, 0.9, 0.8, 1.2, 0.7, 1.0]),
        e.ppcg_end_begin_ur_inner(0.8, 1.7, 6), e.ppcg_end_begin_ur_inner(0.6, 1.7, 6),
        e.ppcg_end_begin_ur(0.7, 1.7)),
    "fused_ppcg_end_begin_ur_inner6_norm_between": lambda e: (
        e.ppcg_init_constants([0.3, 0.5, 0.7, 0.2, 0.6, 0.4], [1.1, 0.9, 0.8, 1.2, 0.7, 1.0]),
        e.ppcg_end_begin_ur_inner(0.8, 1.7, 6), e.calc_2norm(1), e.ppcg_end_begin_ur_inner(0.6, 1.7, 6)),
    "fused_ppcg_end_begin_ur_inner6_then_end": lambda e: (
        e.ppcg_init_constants([0.3, 0.5, 0.7, 0.2, 0.6, 0.4], [1.1, 0.9, 0.8, 1.2, 0.7, 1.0]),
        e.ppcg_end_begin_ur_inner(0.8, 1.7, 6), e.ppcg_end()),
    "fused_ppcg_inner_n4_then_end_begin_ur": lambda e: (
        e.ppcg_init_constants([0.3, 0.5, 0.7, 0.2, 0.6, 0.4], [1.1, 0.9, 0.8, 1.2, 0.7, 1.0]),
        e.ppcg_begin_ur(False, 0.0, 1.7), e.ppcg_inner_n(1, 4), e.ppcg_end_begin_ur(0.8, 1.7)),
    # ten block-Jacobi steps: the first pass forms sd from r (ur stores neither
    # z nor sd); z = M^-1 r stays owed until the end replaces it or a reader
    # (the norm of r.z, the field comparison) has it written out
    "fused_ppcg_end_begin_ur_inner10_twice": lambda e: (
        e.ppcg_init_constants(A10, B10), e.ppcg_end_begin_ur_inner(0.8, 1.7, 10),
        e.ppcg_end_begin_ur_inner(0.6, 1.7, 10), e.ppcg_end_begin_ur(0.7, 1.7)),
    "fused_ppcg_end_begin_ur_inner10_norm_between": lambda e: (
        e.ppcg_init_constants(A10, B10), e.ppcg_end_begin_ur_inner(0.8, 1.7, 10), e.calc_2norm(2),
        e.ppcg_end_begin_ur_inner(0.6, 1.7, 10)),
    "fused_ppcg_end_begin_ur_inner10": lambda e: (
        e.ppcg_init_constants(A10, B10), e.ppcg_end_begin_ur_inner(0.8, 1.7, 10)),
    "fused_ppcg_iteration10_twice": lambda e: (
        e.ppcg_init_constants(A10, B10), e.ppcg_iteration(False, 0.0, 1.7, 10), e.ppcg_iteration(True, 0.45, 1.7, 10)),
    "fused_ppcg_begin_ur_then_inner_n4_n3_n3_end": lambda e: (
        e.ppcg_init_constants(A10, B10), e.ppcg_begin_ur(True, 0.45, 1.7), e.ppcg_inner_n(1, 4),
        e.ppcg_inner_n(5, 3), e.ppcg_inner_n(8, 3), e.ppcg_end()),
    "fused_ppcg_begin_ur_then_inner_n5_n5_end": lambda e: (
        e.ppcg_init_constants(A10, B10), e.ppcg_begin_ur(True, 0.45, 1.7), e.ppcg_inner_n(1, 5),
        e.ppcg_inner_n(6, 5), e.ppcg_end()),
    "fused_ppcg_begin_ur_then_begin_ur": lambda e: (
        e.ppcg_begin_ur(False, 0.0, 1.7), e.ppcg_begin_ur(True, 0.45, 1.7)),
    "fused_ppcg_inner2": lambda e: (e.ppcg_init_constants([0.3, 0.5, 0.7], [1.1, 0.9, 0.8]), e.ppcg_inner2(2)),
    # a whole outer iteration in one call, twice (the second with p owed), and with no inner steps
    "fused_ppcg_iteration_twice": lambda e: (
        e.ppcg_init_constants([0.3, 0.5, 0.7, 0.2, 0.6, 0.4], [1.1, 0.9, 0.8, 1.2, 0.7, 1.0]),
        e.ppcg_iteration(False, 0.0, 1.7, 6), e.ppcg_iteration(True, 0.45, 1.7, 6)),
    "fused_ppcg_end_then_iteration0": lambda e: (e.ppcg_end(), e.ppcg_iteration(True, 0.45, 1.7, 0)),
    **{f"fused_ppcg_inner_n{n}": (lambda n: lambda e: (
        e.ppcg_init_constants([0.3, 0.5, 0.7, 0.2, 0.6, 0.4], [1.1, 0.9, 0.8, 1.2, 0.7, 1.0]),
        e.ppcg_inner_n(7 - n, n)))(n) for n in range(1, 6)},
    # begin_ur leaves rtemp = r and utemp = sd to the next streaming pass; with
    # another call in between they are written out first
    "fused_ppcg_begin_ur_then_inner_n5": lambda e: (
        e.ppcg_init_constants([0.3, 0.5, 0.7, 0.2, 0.6, 0.4], [1.1, 0.9, 0.8, 1.2, 0.7, 1.0]),
        e.ppcg_begin_ur(True, 0.45, 1.7), e.ppcg_inner_n(1, 5)),
    "fused_ppcg_begin_ur_then_norm_then_inner_n3": lambda e: (
        e.ppcg_init_constants([0.3, 0.5, 0.7, 0.2, 0.6, 0.4], [1.1, 0.9, 0.8, 1.2, 0.7, 1.0]),
        e.ppcg_begin_ur(False, 0.0, 1.7), e.calc_2norm(1), e.ppcg_inner_n(2, 3)),
    # the same passes through the tile kernels instead of the streaming kernel
    **{f"fused_ppcg_inner_n{n}_tile": (lambda n: _tile_kernels(lambda e: (
        e.ppcg_init_constants([0.3, 0.5, 0.7, 0.2, 0.6, 0.4], [1.1, 0.9, 0.8, 1.2, 0.7, 1.0]),
        e.ppcg_inner_n(7 - n, n))))(n) for n in range(1, 6)},
}
CH_A7, CH_B7 = [0.2, 0.4, 0.6, 0.3, 0.5, 0.7, 0.25], [0.9, 0.8, 0.7, 0.85, 0.75, 0.65, 0.95]
CHEBY_CALLS = {
    "cheby_init": lambda e: e.cheby_init([0.2, 0.4, 0.6], [0.9, 0.8, 0.7], 1.3),
    "cheby_iterate": lambda e: (e.cheby_init([0.2, 0.4, 0.6], [0.9, 0.8, 0.7], 1.3), e.cheby_iterate(3)),
    # the fused iteration leaves w and r owed: iterations in a row, a norm between them, a halo update
    "cheby_iterate_three": lambda e: (e.cheby_init([0.2, 0.4, 0.6], [0.9, 0.8, 0.7], 1.3), e.cheby_iterate(1),
                                      e.cheby_iterate(2), e.cheby_iterate(3)),
    "cheby_iterate_norm_iterate": lambda e: (e.cheby_init([0.2, 0.4, 0.6], [0.9, 0.8, 0.7], 1.3), e.cheby_iterate(1),
                                             e.calc_2norm(1), e.cheby_iterate(2)),
    "cheby_iterate_halo_iterate": lambda e: (e.cheby_init([0.2, 0.4, 0.6], [0.9, 0.8, 0.7], 1.3), e.cheby_iterate(1),
                                             e.update_halo([0, 0, 0, 1, 0, 0, 0], 2), e.cheby_iterate(2)),
    # runs of iterations in streaming passes (tea_engine_cheby_iterate_n): one pass, several with a norm
    # after them, and after a fused single iteration
    "cheby_iterate_n3": lambda e: (e.cheby_init(CH_A7, CH_B7, 1.3), e.cheby_iterate_n(1, 3)),
    "cheby_iterate_n7_norm": lambda e: (e.cheby_init(CH_A7, CH_B7, 1.3), e.cheby_iterate_n(1, 7), e.calc_2norm(1)),
    "cheby_iterate_then_n5": lambda e: (e.cheby_init(CH_A7, CH_B7, 1.3), e.cheby_iterate(1),
                                        e.cheby_iterate_n(2, 5), e.cheby_iterate(7)),
}


@pytest.mark.parametrize("prec", ["none", "jac_diag", "jac_block"])
@pytest.mark.parametrize("solver,calls", [("cg", CALLS), ("ppcg", PPCG_CALLS), ("chebyshev", CHEBY_CALLS)])
def test_entry_points_bitwise(solver, calls, prec, oracle_so, oracle_env, tmp_path):
    oracle_env("dot2")
    g, c = Engine(), Engine(oracle_so)
    try:
        for k, (name, fn) in enumerate(sorted(calls.items())):
            print(f"[gpu parity] entry point {solver}/{prec}/{name}", flush=True)
            for e in (g, c):
                _prepare(e, tmp_path, solver, prec)
            _randomise((g, c), seed=k)
            for e in (g, c):
                e.cg_init()                 # preconditioner set-up (Mi or the block factors)
            _randomise((g, c), seed=100 + k)
            out_g, out_c = fn(g), fn(c)
            assert out_g == out_c, name
            _compare(g, c, name)
    finally:
        g.finalize()
        c.finalize()


@pytest.mark.parametrize("depth", [1, 2])
def test_halo_entry_points_bitwise(depth, oracle_so, oracle_env, tmp_path):
    oracle_env("dot2")
    g, c = Engine(), Engine(oracle_so)
    try:
        for e in (g, c):
            _prepare(e, tmp_path, "cg", "none", nx=37, ny=29, halo=2)
        _randomise((g, c), seed=99)
        fields = [1, 1, 1, 1, 1, 1, 1]
        for e in (g, c):
            e.update_halo(fields, depth)
        _compare(g, c, "update_halo")
        for face in (rt.CHUNK_LEFT, rt.CHUNK_RIGHT, rt.CHUNK_BOTTOM, rt.CHUNK_TOP):
            side = depth * ((29 if face <= rt.CHUNK_RIGHT else 37) + 2 * depth)
            off = [k * side for k in range(7)]
            bufs = [e.pack(fields, off, depth, face, 7 * side) for e in (g, c)]
            assert np.array_equal(bufs[0], bufs[1])
            rnd = np.random.default_rng(face).standard_normal(bufs[0].shape)
            for e in (g, c):
                e.unpack(fields, off, depth, face, rnd)
            _compare(g, c, f"unpack face {face}")
    finally:
        g.finalize()
        c.finalize()


# ------------------------------------------------------------------ FP32 mode

FP32_FIELDS = (rt.FIELD_U, rt.FIELD_P, rt.FIELD_SD, rt.FIELD_R, rt.XFIELD_U0, rt.XFIELD_W, rt.XFIELD_Z,
               rt.XFIELD_RTEMP, rt.XFIELD_UTEMP, rt.XFIELD_RSTORE)


def _randomise32(engines, seed):
    rng = np.random.default_rng(seed)
    shape = engines[0].shape()
    for fid in FP32_FIELDS:
        a = rng.standard_normal(shape).astype(np.float32)
        for e in engines:
            e.set_field32(fid, a)


@pytest.mark.parametrize("prec", ["none", "jac_diag", "jac_block"])
def test_fp32_entry_points_bitwise(prec, oracle_so, oracle_env, tmp_path):
    """FP32 mode: every CG/PPCG entry point on random float fields, then
    mp_begin / mp_end, compared bit for bit (FP32 fields and FP64 results)."""
    oracle_env("dot2")
    g, c = Engine(), Engine(oracle_so)
    calls = dict(CALLS)
    calls.update(PPCG_CALLS)
    for k in ("calc_residual", "calc_2norm_0", "calc_2norm_1", "calc_2norm_2", "finalise", "field_summary"):
        del calls[k]
    try:
        for k, (name, fn) in enumerate(sorted(calls.items())):
            print(f"[gpu parity] fp32 entry point {prec}/{name}", flush=True)
            for e in (g, c):
                _prepare(e, tmp_path, "ppcg", prec)
                e.set_precision(32)
                e.mp_setup()
            _randomise32((g, c), seed=k)
            for e in (g, c):
                e.cg_init()
            _randomise32((g, c), seed=200 + k)
            out_g, out_c = fn(g), fn(c)
            assert out_g == out_c, name
            for fid in range(1, 20):
                if fid in (rt.XFIELD_CP, rt.XFIELD_BFP):
                    continue  # FP32 mode keeps no Thomas factors: the solves recompute them in registers
                assert np.array_equal(g.get_field32(fid), c.get_field32(fid), equal_nan=True), f"{name}: {fid}"
        print("[gpu parity] fp32 mp_begin/mp_end", flush=True)
        rng = np.random.default_rng(7)
        r, u = rng.standard_normal(g.shape()), rng.standard_normal(g.shape())
        outs = []
        for e in (g, c):
            e.set_precision(64)
            e.set_field(rt.FIELD_R, r)
            e.set_field(rt.FIELD_U, u)
            e.set_precision(32)
            outs.append(e.mp_begin())
            e.set_field32(rt.FIELD_U, r.astype(np.float32) * np.float32(0.5))
            e.mp_end()
        assert outs[0] == outs[1]
        _compare(g, c, "mp_end")
        assert np.array_equal(g.get_field32(rt.FIELD_R), c.get_field32(rt.FIELD_R))
        # the fused replacement, twice (the second on the swapped u buffers),
        # with random u (halo included) and u32
        for rep in range(2):
            print(f"[gpu parity] fp32 mp_replace {prec} call {rep}", flush=True)
            u = rng.standard_normal(g.shape())
            u32 = (rng.standard_normal(g.shape()) * 0.01).astype(np.float32)
            for e in (g, c):
                e.set_precision(64)
                e.set_field(rt.FIELD_U, u)
                e.set_precision(32)
                e.set_field32(rt.FIELD_U, u32)
            outs = [e.mp_replace(1.7) for e in (g, c)]
            assert outs[0] == outs[1]
            for fid in range(1, 20):
                if fid in (rt.XFIELD_CP, rt.XFIELD_BFP):
                    continue
                assert np.array_equal(g.get_field32(fid), c.get_field32(fid), equal_nan=True), f"mp_replace: {fid}"
            for e in (g, c):
                e.set_precision(64)
            _compare(g, c, "mp_replace")
            for e in (g, c):
                e.set_precision(32)
        # a replacement that ends the solve, then mp_leave: the FP64 state and
        # the FP32 fields outside the inner solver (u32, r32, u0_32) as after
        # set_precision(64); the FP32 mode is off
        print(f"[gpu parity] fp32 mp_replace {prec} then mp_leave", flush=True)
        for e in (g, c):
            e.set_precision(32)
        outs = [e.mp_replace(1.3) for e in (g, c)]
        assert outs[0] == outs[1]
        for e in (g, c):
            e.mp_leave()
        _compare(g, c, "mp_leave")
        for fid in (rt.FIELD_U, rt.FIELD_R, rt.XFIELD_U0):
            assert np.array_equal(g.get_field32(fid), c.get_field32(fid), equal_nan=True), f"mp_leave: {fid}"
        for e in (g, c):
            with pytest.raises(TeaError, match="FP32 solver mode"):
                e.mp_replace(1.3)
    finally:
        g.finalize()
        c.finalize()


@pytest.mark.parametrize("prec", ["none", "jac_diag", "jac_block"])
def test_fp32_cheby_entry_points_bitwise(prec, oracle_so, oracle_env, tmp_path):
    """FP32 mode, Chebyshev: cheby_init / cheby_iterate and the 2-norms of the
    FP32 solver's fields on random float fields, compared bit for bit."""
    oracle_env("dot2")
    g, c = Engine(), Engine(oracle_so)
    calls = dict(CHEBY_CALLS)
    calls.update({k: CALLS[k] for k in ("calc_2norm_0", "calc_2norm_1", "calc_2norm_2")})
    try:
        for k, (name, fn) in enumerate(sorted(calls.items())):
            print(f"[gpu parity] fp32 chebyshev entry point {prec}/{name}", flush=True)
            for e in (g, c):
                _prepare(e, tmp_path, "chebyshev", prec)
                e.set_precision(32)
                e.mp_setup()
            _randomise32((g, c), seed=300 + k)
            for e in (g, c):
                e.cg_init()
            _randomise32((g, c), seed=400 + k)
            out_g, out_c = fn(g), fn(c)
            assert out_g == out_c, name
            for fid in range(1, 20):
                if fid in (rt.XFIELD_CP, rt.XFIELD_BFP):
                    continue  # FP32 mode keeps no Thomas factors: the solves recompute them in registers
                assert np.array_equal(g.get_field32(fid), c.get_field32(fid), equal_nan=True), f"{name}: {fid}"
    finally:
        g.finalize()
        c.finalize()
