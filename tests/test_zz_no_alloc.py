"""tea_engine_calc_2norm_noalloc (the 2-norm without allocation) against
tea_leaf_calc_2norm_kernel_cuda_ on the same data, through the C-ABI: the
three norms (u0.u0, r.r, r.z) on a padded copy of the engine's fields in
caller-owned device memory, bit for bit, on two calls in a row, the second
on a smaller mesh; sizes beyond the limits are rejected with an error."""
from __future__ import annotations

import ctypes
import math

import numpy as np
import pytest
import torch  # first: torch's HIP runtime must be the process's one

from tealeaf_cuda_amd import Deck, Engine, StateSpec
from tealeaf_cuda_amd import runtime as rt

pytestmark = pytest.mark.gpu

LAYOUT = [StateSpec(100.0, 0.0001), StateSpec(0.1, 25.0, "rectangle", 0.0, 1.0, 1.0, 2.0),
          StateSpec(5.0, 2.0, "circle", 3.0, 0.0, 6.0, 0.0, 1.5)]


def _chunk(eng, tmp_path, nx, ny, halo):
    deck = Deck(x_cells=nx, y_cells=ny, solver="cg", preconditioner="jac_diag", halo_depth=halo, states=LAYOUT)
    eng.initialise(deck.write(tmp_path / f"noalloc_{nx}x{ny}.in"), nx, ny)
    eng.initialise_chunk(0.0, 0.0, 10.0 / nx, 10.0 / ny)
    st = LAYOUT
    eng.generate_chunk([s.density for s in st], [s.energy for s in st], [s.xmin for s in st],
                       [s.xmax for s in st], [s.ymin for s in st], [s.ymax for s in st], [s.radius for s in st],
                       [{"rectangle": 1, "circle": 2, "point": 3}.get(s.geometry or "", 0) for s in st])
    eng.set_field_kernel()
    eng.common_init(rt.CONDUCTIVITY, 0.004)


def _fn(eng):
    fn = eng.lib.tea_engine_calc_2norm_noalloc
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_long,
                   ctypes.c_void_p, ctypes.c_void_p]
    fn.restype = ctypes.c_int
    return fn


def test_noalloc_2norm_equals_calc_2norm(tmp_path):
    eng = Engine()
    fn = _fn(eng)
    stream = torch.cuda.current_stream()
    try:
        for call, (nx, ny, halo, pad, seed) in enumerate([(333, 270, 2, 5, 1), (70, 45, 1, 0, 2)]):
            print(f"[gpu no-alloc] call {call}: {nx} x {ny}, halo {halo}, pitch pad {pad}", flush=True)
            _chunk(eng, tmp_path, nx, ny, halo)
            rng = np.random.default_rng(seed)
            shape = eng.shape()
            fields = {fid: rng.standard_normal(shape) for fid in (rt.XFIELD_U0, rt.FIELD_R, rt.XFIELD_Z)}
            for fid, a in fields.items():
                eng.set_field(fid, a)
            expect = [eng.calc_2norm(k) for k in range(3)]
            pitch = shape[1] + pad
            dev = {}
            for fid, a in fields.items():
                t = torch.full((shape[0], pitch), math.nan, dtype=torch.float64, device="cuda")
                t[:, :shape[1]] = torch.from_numpy(a).cuda()    # the padding holds NaN: never read
                dev[fid] = t
            out = torch.full((3,), math.nan, dtype=torch.float64, device="cuda")
            pairs = [(rt.XFIELD_U0, rt.XFIELD_U0), (rt.FIELD_R, rt.FIELD_R), (rt.FIELD_R, rt.XFIELD_Z)]
            torch.cuda.synchronize()
            for k, (fa, fb) in enumerate(pairs):
                rc = fn(dev[fa].data_ptr(), dev[fb].data_ptr(), nx, ny, halo, pitch, out[k:].data_ptr(),
                        stream.cuda_stream)
                assert rc == 0
            torch.cuda.synchronize()
            got = out.cpu().tolist()
            assert got == expect, (got, expect)
        print("[gpu no-alloc] size limits", flush=True)
        t = torch.zeros(16, dtype=torch.float64, device="cuda")
        assert fn(t.data_ptr(), t.data_ptr(), 0, 4, 1, 8, t.data_ptr(), None) != 0
        assert fn(t.data_ptr(), t.data_ptr(), 1 << 17, 4, 1, (1 << 17) + 2, t.data_ptr(), None) != 0
        assert fn(t.data_ptr(), t.data_ptr(), 4, 4, 1, 5, t.data_ptr(), None) != 0      # pitch < nx + 2 halo
        buf = ctypes.create_string_buffer(1024)
        assert eng.lib.tea_engine_last_error(buf, 1024) == 1 and b"calc_2norm_noalloc" in buf.value
        eng.lib.tea_engine_clear_error()
    finally:
        eng.finalize()
