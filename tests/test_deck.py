"""Deck writer, seeded layout generator and tea.out parser."""
from __future__ import annotations

import math

import pytest

from conftest import GOLDEN
from tealeaf_cuda_amd import Deck, parse_tea_out, random_layout, run_deck


def test_parse_golden_outputs():
    out = parse_tea_out(GOLDEN / "tea_bm_2.out")
    assert out["iterations"] == [228, 242, 244, 238, 231, 222, 215, 208, 202, 197]
    assert [s for s, _ in out["summaries"]] == [0, 10, 10]
    assert out["summaries"][-1][1][4] == 106.27221178646569
    assert len(out["conduction_error"]) == 10
    out5 = parse_tea_out(GOLDEN / "tea_bm_5.out")
    assert out5["summaries"][-1][1][4] == 95.462351583362249
    assert len(out5["iterations"]) == 10


@pytest.mark.parametrize("seed", [0, 1, 2, 17, 123])
def test_random_layout_properties(seed):
    st = random_layout(seed)
    assert st == random_layout(seed)                  # deterministic
    assert st[0].density == 100.0 and st[0].energy == 1e-4 and st[0].geometry is None
    assert 3 <= len(st) - 1 <= 8
    for s in st[1:]:
        assert s.geometry == "rectangle"
        assert 0.1 <= s.density <= 100.0
        assert 1e-4 <= s.energy <= 25.0
        assert 0.0 <= s.xmin < s.xmax <= 10.0 and 0.0 <= s.ymin < s.ymax <= 10.0
        assert s.xmax - s.xmin >= 2 * 10.0 / 4096 and s.ymax - s.ymin >= 2 * 10.0 / 4096
    dens = [math.log10(s.density) for seed_ in range(200) for s in random_layout(seed_)[1:]]
    assert min(dens) < -0.8 and max(dens) > 1.8     # spans the log-uniform range


def test_deck_round_trip_through_the_driver(oracle_so, driver_built, tmp_path):
    d = Deck(x_cells=16, y_cells=12, end_step=2, solver="cg", eps=1e-12, states=random_layout(3, cells=16),
             summary_frequency=1, ymax=7.5)
    text = d.text()
    assert "*tea" in text and "*endtea" in text and "tl_use_cg" in text
    for line in text.splitlines():  # every real literal carries a '.', as read_input.f90 needs
        for tok in line.replace("=", " ").split():
            if any(ch.isdigit() for ch in tok) and ("e" in tok and "." not in tok.split("e")[0]):
                pytest.fail(f"real literal without '.': {tok}")
    assert max(len(line) for line in text.splitlines()) <= 100   # read_input.f90 reads (a100)
    res = run_deck(d.write(tmp_path / "tea.in"), backend=oracle_so, out_path=tmp_path / "tea.out")
    assert res.summaries[0].volume == pytest.approx(10.0 * 7.5, rel=1e-13)
    assert res.summaries[0].mass < 100.0 * 75.0 * 0.99               # the rectangles were painted
    assert [s.step for s in res.summaries] == [0, 1, 2, 2]
    assert parse_tea_out(tmp_path / "tea.out")["iterations"] == res.iterations
