"""Pins the oracle and the host driver against the reference's own golden
outputs: Benchmarks/tea_bm_{1..4}.out (CG, no preconditioner, tl_eps 1e-15,
10 steps, one task), copied as data into tests/golden.

With the oracle's serial (Fortran-order) reductions the per-step iteration
counts must be identical and every summary value must agree to 1e-10
relative; the step-0 volume, mass and density, which only depend on the set-up
arithmetic and the summation order, must agree to all 17 printed digits.
"""
from __future__ import annotations

import pytest

from conftest import GOLDEN
from tealeaf_cuda_amd import parse_tea_out, run_deck

CASES = [1, 2, 3, 4]


@pytest.mark.parametrize("bm", CASES)
def test_benchmark_golden_serial(bm, oracle_so, driver_built, oracle_env, tmp_path):
    oracle_env("serial")
    golden = parse_tea_out(GOLDEN / f"tea_bm_{bm}.out")
    res = run_deck(GOLDEN / f"tea_bm_{bm}.in", backend=oracle_so, out_path=tmp_path / "tea.out")

    assert res.iterations == golden["iterations"]

    got = [(s.step, s.values()) for s in res.summaries]
    assert [s for s, _ in got] == [s for s, _ in golden["summaries"]]
    for (_, g), (_, ref) in zip(got, golden["summaries"]):
        for a, b in zip(g, ref):
            assert abs(a - b) <= 1e-10 * abs(b), (a, b)

    # set-up only: identical to the printed 17 significant digits
    step0 = res.summaries[0]
    for a, b in zip((step0.volume, step0.mass, step0.density), golden["summaries"][0][1][:3]):
        assert float(f"{a:.16e}") == float(f"{b:.16e}"), (a, b)

    # the reference's own QA criterion (field_summary.f90:96-115): U within 0.001 %
    qa = {1: 157.55084183279294, 2: 106.27221178646569, 3: 99.955877498324000, 4: 97.277332050749976}[bm]
    assert abs(100.0 * res.final.u / qa - 100.0) < 0.001

    # the driver's tea.out log parses back to the same numbers
    log = parse_tea_out(tmp_path / "tea.out")
    assert log["iterations"] == res.iterations
    assert [v for _, v in log["summaries"]][0] == pytest.approx(step0.values(), rel=1e-15)


@pytest.mark.parametrize("bm", [1, 2])
def test_benchmark_golden_compensated(bm, oracle_so, driver_built, oracle_env):
    """The compensated (order-independent) reductions the GPU uses land within
    1e-10 of the same goldens, with the same iteration counts."""
    oracle_env("dot2")
    golden = parse_tea_out(GOLDEN / f"tea_bm_{bm}.out")
    res = run_deck(GOLDEN / f"tea_bm_{bm}.in", backend=oracle_so)
    assert res.iterations == golden["iterations"]
    for s, (_, ref) in zip(res.summaries, golden["summaries"]):
        for a, b in zip(s.values(), ref):
            assert abs(a - b) <= 1e-10 * abs(b)
