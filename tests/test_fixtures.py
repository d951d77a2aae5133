"""The committed text fixtures of the seeded-layout generator and of the
tea.out format (tests/golden/make_fixtures.py writes them).

* the layout generator reproduces the committed deck of the 4096^2
  measurement and the held-out 1024^2 layouts byte for byte, and bench.py
  measures exactly that deck;
* the reference's own Fortran driver reads the committed deck's states as the
  generator meant them (step-0 field summary against the closed form of the
  painted rectangles, on a coarse copy of the deck);
* the host driver over the oracle rewrites the committed tea.out byte for
  byte; parse_tea_out reads it; the reference's Fortran driver prints the same
  numbers; the HIP engine's log has the same iteration counts, residuals and
  field summaries (gpu).
"""
from __future__ import annotations

import re
import sys
from pathlib import Path

import numpy as np
import pytest

from conftest import GOLDEN, REF_ORACLE, REPO, run_in_dir
from tealeaf_cuda_amd import Deck, parse_tea_out, random_layout

sys.path.insert(0, str(GOLDEN))
import make_fixtures as mf  # noqa: E402


def test_measurement_deck_fixture():
    assert mf.measurement_deck().text() == (GOLDEN / "deck_random_layout_4096.in").read_text()


def test_bench_measures_the_fixture_deck(tmp_path):
    sys.path.insert(0, str(REPO))
    import bench

    path = bench.make_deck(tmp_path / "tea.in", 4096, 20, **bench.variants(4096)["fp64_reference"])
    assert Path(path).read_text() == (GOLDEN / "deck_random_layout_4096.in").read_text()


def test_heldout_layout_fixture():
    text = (GOLDEN / "layouts_1024_heldout.in").read_text()
    assert mf.heldout_text() == text
    assert text.count("! seed") == len(mf.HELDOUT_SEEDS)


def _painted_mass(states, n):
    """Closed-form mass of a layout on an n x n mesh of [0,10]^2: cell (i, j)
    takes the last state whose rectangle, shrunk by a hundredth of a cell
    (read_input.f90:320-327), overlaps it (generate_chunk_kernel.cuknl)."""
    dx = 10.0 / n
    x = np.arange(n) * dx
    dens = np.full((n, n), states[0].density)
    for s in states[1:]:
        xa, xb, ya, yb = s.xmin + dx / 100, s.xmax - dx / 100, s.ymin + dx / 100, s.ymax - dx / 100
        cx = (x + dx >= xa) & (x < xb)
        cy = (x + dx >= ya) & (x < yb)
        dens[np.ix_(cy, cx)] = s.density
    return float(dens.sum() * dx * dx)


@pytest.mark.skipif(not REF_ORACLE.exists(), reason="reference driver build (oracle/_ref) absent")
def test_reference_driver_reads_the_fixture_states(oracle_so, tmp_path):
    text = (GOLDEN / "deck_random_layout_4096.in").read_text()
    coarse = re.sub(r"^x_cells=4096$", "x_cells=256", text, flags=re.M)
    coarse = re.sub(r"^y_cells=4096$", "y_cells=256", coarse, flags=re.M)
    coarse = re.sub(r"^end_step=20$", "end_step=1", coarse, flags=re.M)
    ref = run_in_dir(REF_ORACLE, coarse, tmp_path, env={"TEA_ORACLE_DOT": "dot2"})
    assert ref.returncode == 0, ref.stderr[-2000:]
    step0 = parse_tea_out(tmp_path / "tea.out")["summaries"][0]
    assert step0[0] == 0
    vol, mass = step0[1][0], step0[1][1]
    assert vol == pytest.approx(100.0, rel=1e-14)
    assert mass == pytest.approx(_painted_mass(random_layout(4096, cells=4096), 256), rel=1e-12)


def test_tea_out_fixture_rewritten(oracle_so, driver_built, oracle_env, tmp_path):
    oracle_env("dot2")
    mf.write_tea_out(tmp_path / "tea.out", oracle_so)
    assert (tmp_path / "tea.out").read_text() == (GOLDEN / "random_layout_128_ppcg.out").read_text()


def test_tea_out_fixture_parses():
    out = parse_tea_out(GOLDEN / "random_layout_128_ppcg.out")
    assert len(out["iterations"]) == 3 and all(k > 0 for k in out["iterations"])
    assert [s for s, _ in out["summaries"]] == [0, 1, 2, 3, 3]
    assert len(out["residuals"]) >= sum(out["iterations"])
    vol = [v[0] for _, v in out["summaries"]]
    mass = [v[1] for _, v in out["summaries"]]
    assert vol == [100.0] * 5 and len(set(mass)) == 1       # volume and mass do not change
    ie = [v[3] for _, v in out["summaries"]]
    assert max(ie) - min(ie) <= 1e-10 * ie[0]               # energy is conserved to the solver tolerance


@pytest.mark.skipif(not REF_ORACLE.exists(), reason="reference driver build (oracle/_ref) absent")
def test_reference_driver_prints_the_fixture(oracle_so, tmp_path):
    ref = run_in_dir(REF_ORACLE, mf.tea_out_deck().text(), tmp_path, env={"TEA_ORACLE_DOT": "dot2"})
    assert ref.returncode == 0, ref.stderr[-2000:]
    a, b = parse_tea_out(tmp_path / "tea.out"), parse_tea_out(GOLDEN / "random_layout_128_ppcg.out")
    assert a["iterations"] == b["iterations"]
    assert a["residuals"] == b["residuals"]
    assert a["conduction_error"] == b["conduction_error"]
    assert a["summaries"] == b["summaries"]


@pytest.mark.gpu
def test_hip_engine_prints_the_fixture(tmp_path):
    import tempfile

    from tealeaf_cuda_amd import run_deck

    print("[gpu fixtures] 128^2 random-layout PPCG deck, verbose tea.out", flush=True)
    with tempfile.TemporaryDirectory() as d:
        deck = mf.tea_out_deck().write(Path(d) / "tea.in")
        run_deck(deck, out_path=tmp_path / "tea.out")
    got = mf.strip_timing((tmp_path / "tea.out").read_text())
    assert got == (GOLDEN / "random_layout_128_ppcg.out").read_text()
