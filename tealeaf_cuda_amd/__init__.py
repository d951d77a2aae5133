"""TeaLeaf solver engine for AMD MI355X (gfx950).

The kernel layer is ``libtea_engine.so``: hand-written HIP kernels behind the
29-symbol Fortran C-ABI of the TeaLeaf CUDA reference (include/tea_leaf_engine.h).
``run_deck`` drives a whole tea.in deck through the host driver; ``Engine``
exposes the ABI entry points one by one.
"""
from .deck import Deck, StateSpec, parse_tea_out, random_layout, standard_states
from .runtime import (ABI_SYMBOLS, EXTENSION_SYMBOLS, Engine, RunResult, Summary, TeaError, driver_library,
                      engine_library, run_deck)

__all__ = [
    "ABI_SYMBOLS", "EXTENSION_SYMBOLS", "Deck", "Engine", "RunResult", "StateSpec", "Summary", "TeaError",
    "driver_library", "engine_library", "parse_tea_out", "random_layout", "run_deck", "standard_states",
]
