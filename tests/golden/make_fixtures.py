"""Writes the text fixtures of the seeded-layout generator and of the tea.out
format under tests/golden/ (run from the repository root:
``python tests/golden/make_fixtures.py``).

* ``deck_random_layout_4096.in`` -- the plain FP64 deck of the 4096^2
  measurement: random_layout(seed 4096) emitted as tea.in states 2..n+1
  (read_input.f90:237-289 reads them), PPCG with block Jacobi and 10 inner
  steps, tl_eps 1e-10, halo depth 1, 20 steps;
* ``layouts_1024_heldout.in`` -- the state lines of the first ten held-out
  seeds (10000..10009) of the 1024^2 learned warm start, one deck body each,
  separated by a line ``! seed <n>``;
* ``random_layout_128_ppcg.out`` -- the tea.out log (verbose_on: the residual
  of every iteration) of a 128^2 random-layout PPCG deck, 3 steps, written by
  the host driver over the CPU oracle with compensated reductions, without
  its 'Wall clock' timing lines; the
  reference's own Fortran driver writes the same numbers
  (tests/test_fixtures.py checks both, and the HIP engine against it).

The fixtures are data: regenerating them must reproduce them byte for byte.
"""
from __future__ import annotations

import os
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
REPO = HERE.parent.parent
sys.path.insert(0, str(REPO))

from tealeaf_cuda_amd import Deck, random_layout  # noqa: E402

HELDOUT_SEEDS = tuple(range(10000, 10010))


def measurement_deck() -> Deck:
    return Deck(x_cells=4096, y_cells=4096, end_step=20, solver="ppcg", preconditioner="jac_block", eps=1e-10,
                max_iters=10000, halo_depth=1, ppcg_inner_steps=10, summary_frequency=0, check_result=True,
                states=random_layout(4096, cells=4096))


def heldout_text() -> str:
    out = []
    for seed in HELDOUT_SEEDS:
        out.append(f"! seed {seed}")
        out += [s.line(k + 1) for k, s in enumerate(random_layout(seed, cells=1024))]
    return "\n".join(out) + "\n"


def tea_out_deck() -> Deck:
    return Deck(x_cells=128, y_cells=128, end_step=3, solver="ppcg", preconditioner="jac_block", eps=1e-10,
                ppcg_inner_steps=10, summary_frequency=1, check_result=True, verbose=True,
                states=random_layout(21, cells=128))


def strip_timing(text: str) -> str:
    """The log without its 'Wall clock' lines (run times, not data)."""
    return "".join(l for l in text.splitlines(keepends=True) if not l.strip().startswith("Wall clock"))


def write_tea_out(dest: Path, backend: Path) -> None:
    """Runs tea_out_deck() through the host driver over `backend` and writes
    its tea.out, timing lines removed, to dest."""
    import tempfile

    from tealeaf_cuda_amd import run_deck

    with tempfile.TemporaryDirectory() as d:
        deck = tea_out_deck().write(Path(d) / "tea.in")
        run_deck(deck, backend=backend, out_path=Path(d) / "tea.out")
        dest.write_text(strip_timing((Path(d) / "tea.out").read_text()))


def main() -> None:
    (HERE / "deck_random_layout_4096.in").write_text(measurement_deck().text())
    (HERE / "layouts_1024_heldout.in").write_text(heldout_text())
    os.environ["TEA_ORACLE_DOT"] = "dot2"
    write_tea_out(HERE / "random_layout_128_ppcg.out", REPO / "oracle" / "liboracle.so")
    print("wrote", *(p.name for p in sorted(HERE.glob("*")) if p.suffix in (".in", ".out") and "bm_" not in p.name))


if __name__ == "__main__":
    main()
