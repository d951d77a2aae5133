"""GPU runs of the measured configurations at their own shapes, against the
oracle where the oracle finishes in seconds.

* C1: plain FP64 CG without a preconditioner, 256^2 cells of the reference's
  4-rectangle problem, tl_eps 1e-15, halo depth 2, 10 steps: the HIP engine
  equals the oracle bit for bit (per-step iteration counts, summaries,
  errors), through this repository's driver and through the reference's own
  Fortran driver linked to each library;
* the production shape of the streaming inner-step passes: 1024^2, 2 steps,
  bitwise against the oracle in FP64 (passes of 4+3+3 steps, chunks of 5
  strips) and FP32 (5+5, chunks of 3 strips); and the small parity decks
  again with chunks of several strips forced (TEA_MARCH_CHUNK), so that chunk
  boundaries inside the mesh are compared at every precision and
  preconditioner;
* C4: the learned warm start at 1024^2 on held-out seeds: iterations to the
  FP64 true residual 1e-10 from u^n, from the Galerkin combination of the
  previous steps, and from that combination with the trained network.
"""
from __future__ import annotations

import json

import pytest

from conftest import REF_HIP, REF_ORACLE, REPO, run_in_dir
from tealeaf_cuda_amd import Deck, StateSpec, parse_tea_out, random_layout, run_deck, standard_states

pytestmark = pytest.mark.gpu


def _same(gpu, cpu):
    assert gpu.iterations == cpu.iterations
    assert gpu.ppcg_inner == cpu.ppcg_inner
    assert gpu.refinements == cpu.refinements
    assert gpu.conduction_error == cpu.conduction_error
    assert gpu.exact_error == cpu.exact_error
    assert len(gpu.summaries) == len(cpu.summaries)
    for a, b in zip(gpu.summaries, cpu.summaries):
        assert a.values() == b.values(), (a, b)


def c1_deck() -> Deck:
    return Deck(x_cells=256, y_cells=256, end_step=10, solver="cg", preconditioner="none", eps=1e-15,
                halo_depth=2, max_iters=10000, summary_frequency=1, check_result=True, states=standard_states())


def test_c1_bitwise(oracle_so, driver_built, oracle_env, tmp_path):
    oracle_env("dot2")
    print("[gpu configs] C1: 256^2 CG none, eps 1e-15, halo 2, 10 steps", flush=True)
    path = c1_deck().write(tmp_path / "tea.in")
    gpu = run_deck(path)
    cpu = run_deck(path, backend=oracle_so)
    _same(gpu, cpu)
    assert len(gpu.iterations) == 10 and min(gpu.iterations) > 100
    ie = [s.energy for s in gpu.summaries]
    assert max(ie) - min(ie) <= 1e-12 * ie[0]        # conserved: zero-flux boundaries


@pytest.mark.skipif(not (REF_HIP.exists() and REF_ORACLE.exists()), reason="reference driver build absent")
def test_c1_reference_program_on_both_libraries(oracle_so, tmp_path):
    print("[gpu configs] C1 through the reference Fortran driver", flush=True)
    text = c1_deck().text()
    cpu = run_in_dir(REF_ORACLE, text, tmp_path / "cpu", env={"TEA_ORACLE_DOT": "dot2"})
    gpu = run_in_dir(REF_HIP, text, tmp_path / "gpu")
    assert cpu.returncode == 0, cpu.stderr[-2000:]
    assert gpu.returncode == 0, gpu.stderr[-2000:]
    a, b = parse_tea_out(tmp_path / "gpu" / "tea.out"), parse_tea_out(tmp_path / "cpu" / "tea.out")
    assert a["iterations"] == b["iterations"] and len(a["iterations"]) == 10
    assert a["summaries"] == b["summaries"]
    assert a["conduction_error"] == b["conduction_error"]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_production_shape_bitwise(precision, oracle_so, driver_built, oracle_env, tmp_path):
    """1024^2: the streaming passes run 18-19 bands of 64 lanes with chunks of
    3 (FP32) or 5 (FP64) strips, as at 4096^2 (chunks of 34-79)."""
    oracle_env("dot2")
    print(f"[gpu configs] production shape 1024^2 PPCG jac_block, {precision}, 2 steps", flush=True)
    kw = dict(inner_precision="fp32") if precision == "fp32" else {}
    deck = Deck(x_cells=1024, y_cells=1024, end_step=2, solver="ppcg", preconditioner="jac_block", eps=1e-10,
                ppcg_inner_steps=10, halo_depth=1, summary_frequency=1, check_result=True,
                states=random_layout(10000, cells=1024), **kw)
    path = deck.write(tmp_path / "tea.in")
    gpu = run_deck(path)
    cpu = run_deck(path, backend=oracle_so)
    _same(gpu, cpu)
    assert all(e < (1e-10 if precision == "fp32" else 1e-8) for e in gpu.exact_error)


LAYOUT = [StateSpec(100.0, 0.0001), StateSpec(0.1, 25.0, "rectangle", 0.0, 1.0, 1.0, 2.0),
          StateSpec(0.1, 0.1, "rectangle", 1.0, 6.0, 1.0, 2.0), StateSpec(0.1, 0.1, "rectangle", 5.0, 6.0, 1.0, 8.0),
          StateSpec(5.0, 2.0, "circle", 3.0, 0.0, 6.0, 0.0, 1.5)]
CHUNKED = {
    "ppcg_block": dict(preconditioner="jac_block"),
    "ppcg_diag_h2": dict(preconditioner="jac_diag", halo_depth=2, ppcg_inner_steps=9),
    "ppcg_none_ragged": dict(preconditioner="none", x_cells=97, y_cells=66),
    "mixed_ppcg_block": dict(preconditioner="jac_block", inner_precision="fp32"),
    "mixed_ppcg_diag": dict(preconditioner="jac_diag", inner_precision="fp32", reliable_delta=0.01),
    "mixed_ppcg_none_ragged": dict(preconditioner="none", inner_precision="fp32", x_cells=131, y_cells=77),
}


@pytest.mark.parametrize("chunk", [2, 7])
@pytest.mark.parametrize("name", sorted(CHUNKED))
def test_multi_strip_chunks_bitwise(name, chunk, oracle_so, driver_built, oracle_env, monkeypatch, tmp_path):
    oracle_env("dot2")
    monkeypatch.setenv("TEA_MARCH_CHUNK", str(chunk))
    print(f"[gpu configs] deck {name} with chunks of {chunk} strips", flush=True)
    kw = dict(x_cells=96, y_cells=120, end_step=3, solver="ppcg", ppcg_inner_steps=10, eps=1e-10,
              check_result=True, summary_frequency=1, states=LAYOUT)
    kw.update(CHUNKED[name])
    path = Deck(**kw).write(tmp_path / "tea.in")
    gpu = run_deck(path)
    cpu = run_deck(path, backend=oracle_so)
    _same(gpu, cpu)


@pytest.mark.parametrize("form", ["0", "1", "2"])
@pytest.mark.parametrize("name", ["ppcg_block", "ppcg_diag_h2", "mixed_ppcg_block", "mixed_ppcg_diag",
                                  "mixed_ppcg_none_ragged"])
def test_pass_forms_bitwise(name, form, oracle_so, driver_built, oracle_env, monkeypatch, tmp_path):
    """Every form of the streaming inner-step pass (TEA_INNER_RING: 0 the
    shifting pipeline, 1 the ring, 2 the ring with its coefficients in LDS)
    with chunks of 3 strips gives the oracle's results."""
    oracle_env("dot2")
    monkeypatch.setenv("TEA_MARCH_CHUNK", "3")
    monkeypatch.setenv("TEA_INNER_RING", form)
    print(f"[gpu configs] deck {name}, pass form {form}", flush=True)
    kw = dict(x_cells=96, y_cells=120, end_step=3, solver="ppcg", ppcg_inner_steps=10, eps=1e-10,
              check_result=True, summary_frequency=1, states=LAYOUT)
    kw.update(CHUNKED[name])
    path = Deck(**kw).write(tmp_path / "tea.in")
    gpu = run_deck(path)
    cpu = run_deck(path, backend=oracle_so)
    _same(gpu, cpu)


def test_c4_heldout_warm_start(tmp_path):
    """C4 on held-out seeds 10000-10002, 5 steps each: iterations to the FP64
    true residual 1e-10 (of the u^n guess) for the three guesses.  Every run
    reaches the target; both warm starts save iterations over u^n."""
    from tealeaf_cuda_amd import warmstart as wsm

    weights = REPO / "tealeaf_cuda_amd" / "weights" / "warmstart_1024.tlws"
    variants = {"plain": None, "galerkin": "extrapolate", "network": str(weights)}
    print("[gpu configs] C4: 1024^2 learned warm start, held-out seeds 10000-10002 x 5 steps", flush=True)
    rows = wsm.evaluate(range(10000, 10003), 1024, 5, variants, tmp_path, inner_precision="fp64", log=None)
    summary = wsm.summarise(rows, variants)
    print("[gpu configs] C4 " + json.dumps(summary), flush=True)
    for r in rows:
        for tag in variants:
            assert all(0 < e < 1e-10 for e in r[tag]["exact"]), (r["seed"], tag, r[tag]["exact"])
    assert summary["galerkin"]["iterations"] < summary["plain"]["iterations"]
    assert summary["network"]["iterations"] < summary["plain"]["iterations"]


@pytest.mark.skipif(not REF_HIP.exists(), reason="reference driver build absent")
def test_profiler_on_table(tmp_path):
    """profiler_on: the reference program on the HIP library prints the
    kernel layer's per-kernel table (cuda_errors.cu:91-106) when it ends."""
    print("[gpu configs] profiler_on table through the reference Fortran driver", flush=True)
    deck = Deck(x_cells=64, y_cells=64, end_step=2, solver="ppcg", preconditioner="jac_block", eps=1e-10,
                ppcg_inner_steps=6, profiler_on=True)
    run = run_in_dir(REF_HIP, deck.text(), tmp_path)
    assert run.returncode == 0, run.stderr[-2000:]
    lines = run.stdout.splitlines()
    assert "@@@@@ PROFILING @@@@@" in lines
    rows = lines[lines.index("@@@@@ PROFILING @@@@@") + 1:]
    table = {}
    for line in rows:
        name, sep, ms = line.rpartition(" : ")
        if not sep:
            break
        table[name.strip()] = float(ms)
    assert {"cg_calc_w", "calc_residual", "field_summary"} <= set(table), sorted(table)
    assert all(v >= 0.0 for v in table.values()) and sum(table.values()) > 0.0
    assert list(table) == sorted(table)
    assert "Profiler Output" in (tmp_path / "tea.out").read_text()   # the driver's own table
