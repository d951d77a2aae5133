"""Mixed-precision refinement (tl_inner_precision=fp32) on the oracle.

The refinement loop has no counterpart in the reference; what pins it:

* the FP32 instantiation of every CG/PPCG/Chebyshev kernel equals the FP64 one exactly
  on inputs where float arithmetic is exact (small integers and dyadic
  scalars), so the FP32 kernels compute the same expressions as the FP64
  kernels that the reference's golden files pin; with the jac_diag / jac_block
  preconditioners (divisions) they agree to float rounding;
* mp_begin / mp_end are the roundings r32 = fl32(r) (also the FP32 right-hand
  side u0_32, which Chebyshev reads) and u += u32 exactly;
* both refinement schemes (restart: FP32 inner solves to tl_inner_eps, the
  scheme Chebyshev always uses; reliable: residual replacement keeping the
  search direction) reach the FP64
  true relative residual tl_eps on every step
  (tl_check_result measures ||u0 - A u|| / ||r0|| in FP64 after the solve), and
  its field summaries agree with a tightly converged FP64 run to the accuracy
  that residual implies;
* decks the refinement does not support are rejected with an error.
"""
from __future__ import annotations

import numpy as np
import pytest

from tealeaf_cuda_amd import Deck, Engine, StateSpec, TeaError, run_deck
from tealeaf_cuda_amd import runtime as rt

from test_oracle_math import interior, put_interior, setup_chunk

LAYOUT = [StateSpec(100.0, 0.0001), StateSpec(0.1, 25.0, "rectangle", 0.0, 1.0, 1.0, 2.0),
          StateSpec(0.1, 0.1, "rectangle", 1.0, 6.0, 1.0, 2.0), StateSpec(0.1, 0.1, "rectangle", 5.0, 6.0, 1.0, 8.0),
          StateSpec(0.1, 0.1, "rectangle", 5.0, 10.0, 7.0, 8.0)]

SOLVE_FIELDS = (rt.FIELD_U, rt.FIELD_P, rt.FIELD_SD, rt.FIELD_R, rt.XFIELD_U0, rt.XFIELD_W, rt.XFIELD_Z,
                rt.XFIELD_RTEMP, rt.XFIELD_UTEMP, rt.XFIELD_RSTORE)


def _integer_chunk(eng, tmp_path, solver, prec, seed, nx=13, ny=9):
    """A chunk whose Kx/Ky and solver fields hold small integers (exact in float)."""
    setup_chunk(eng, tmp_path, nx=nx, ny=ny, solver=solver, prec=prec, halo=2)
    rng = np.random.default_rng(seed)
    shape = eng.shape()
    for fid in (rt.XFIELD_KX, rt.XFIELD_KY):
        k = rng.integers(1, 5, size=shape).astype(np.float64)
        eng.set_field(fid, k)
    eng.set_precision(32)
    eng.mp_setup()
    vals = {}
    for fid in SOLVE_FIELDS:
        a = np.zeros(shape)
        h = eng.h
        a[h:h + ny, h:h + nx] = rng.integers(-8, 9, size=(ny, nx))
        vals[fid] = a
    return vals


def _load(eng, vals, bits):
    eng.set_precision(bits)
    for fid, a in vals.items():
        if bits == 64:
            eng.set_field(fid, a)
        else:
            eng.set_field32(fid, a.astype(np.float32))


def _read(eng, fid, bits):
    return interior(eng, eng.get_field(fid) if bits == 64 else eng.get_field32(fid).astype(np.float64))


EXACT_CALLS = {
    "cg_init": lambda e: e.cg_init(),
    "cg_calc_w": lambda e: e.cg_calc_w(),
    "cg_calc_ur": lambda e: e.cg_calc_ur(0.5),
    "cg_calc_p": lambda e: e.cg_calc_p(0.25),
    "ppcg_init_1": lambda e: e.ppcg_init(1),
    "ppcg_init_3": lambda e: e.ppcg_init(3),
    "ppcg_init_sd_new": lambda e: e.ppcg_init_sd_new(2.0),
    "ppcg_inner": lambda e: (e.ppcg_init_constants([0.5, 0.25, 0.125], [0.75, 0.5, 0.25]), e.ppcg_inner(2)),
    "ppcg_store_r": lambda e: e.ppcg_store_r(),
    "ppcg_calc_rrn": lambda e: e.ppcg_calc_rrn(),
    "ppcg_calc_2norm": lambda e: e.ppcg_calc_2norm(),
    "ppcg_calc_p": lambda e: e.ppcg_calc_p(0.5),
}
EXACT_CHEBY_CALLS = {
    "cheby_init": lambda e: e.cheby_init([0.5, 0.25], [0.75, 0.5], 2.0),
    "cheby_iterate": lambda e: (e.cheby_init([0.5, 0.25], [0.75, 0.5], 2.0), e.cheby_iterate(2)),
    "calc_2norm_0": lambda e: e.calc_2norm(0),
    "calc_2norm_1": lambda e: e.calc_2norm(1),
    "calc_2norm_2": lambda e: e.calc_2norm(2),
}


@pytest.mark.parametrize("solver,name", [("ppcg", n) for n in sorted(EXACT_CALLS)]
                         + [("chebyshev", n) for n in sorted(EXACT_CHEBY_CALLS)])
def test_fp32_kernels_equal_fp64_on_exact_inputs(solver, name, oracle_so, oracle_env, tmp_path):
    oracle_env("dot2")
    calls = EXACT_CALLS if solver == "ppcg" else EXACT_CHEBY_CALLS
    eng = Engine(oracle_so)
    try:
        vals = _integer_chunk(eng, tmp_path, solver, "none", seed=len(name))
        outs = {}
        fields = {}
        for bits in (64, 32):
            _load(eng, vals, bits)
            outs[bits] = calls[name](eng)
            fields[bits] = {fid: _read(eng, fid, bits) for fid in SOLVE_FIELDS}
        assert outs[32] == outs[64]
        for fid in SOLVE_FIELDS:
            assert np.array_equal(fields[32][fid], fields[64][fid]), f"{name}: field {fid}"
    finally:
        eng.finalize()


@pytest.mark.parametrize("prec", ["jac_diag", "jac_block"])
def test_fp32_preconditioners_agree_to_float_rounding(prec, oracle_so, oracle_env, tmp_path):
    oracle_env("dot2")
    eng = Engine(oracle_so)
    try:
        vals = _integer_chunk(eng, tmp_path, "ppcg", prec, seed=5)
        z = {}
        for bits in (64, 32):
            _load(eng, vals, bits)
            rro = eng.cg_init()
            z[bits] = (_read(eng, rt.XFIELD_Z, bits), rro)
        np.testing.assert_allclose(z[32][0], z[64][0], rtol=2e-6, atol=2e-6 * np.abs(z[64][0]).max())
        assert z[32][1] == pytest.approx(z[64][1], rel=1e-5)
    finally:
        eng.finalize()


def test_mp_begin_end_are_roundings(oracle_so, oracle_env, tmp_path):
    oracle_env("dot2")
    eng = Engine(oracle_so)
    try:
        setup_chunk(eng, tmp_path, nx=11, ny=7, solver="cg")
        rng = np.random.default_rng(3)
        r = rng.standard_normal(eng.shape())
        u = rng.standard_normal(eng.shape())
        eng.set_field(rt.FIELD_R, r)
        eng.set_field(rt.FIELD_U, u)
        eng.set_precision(32)
        rr = eng.mp_begin()
        r32 = interior(eng, eng.get_field32(rt.FIELD_R))
        assert np.array_equal(r32, interior(eng, r).astype(np.float32))
        assert np.array_equal(interior(eng, eng.get_field32(rt.XFIELD_U0)), r32)
        assert not interior(eng, eng.get_field32(rt.FIELD_U)).any()
        x = r32.astype(np.float64).ravel()
        assert rr == pytest.approx(float(np.dot(x, x)), rel=1e-15)
        d = rng.standard_normal(eng.shape()).astype(np.float32)
        eng.set_field32(rt.FIELD_U, d)
        eng.mp_end()
        assert np.array_equal(interior(eng, eng.get_field(rt.FIELD_U)),
                              interior(eng, u) + interior(eng, d).astype(np.float64))
    finally:
        eng.finalize()


def _run(tmp_path, oracle_so, name, **kw):
    base = dict(x_cells=96, y_cells=80, end_step=3, eps=1e-10, check_result=True, summary_frequency=1,
                states=LAYOUT)
    base.update(kw)
    return run_deck(Deck(**base).write(tmp_path / f"{name}.in"), backend=oracle_so,
                    out_path=tmp_path / f"{name}.out")


MIXED_CASES = [
    ("ppcg", "jac_block", dict(mixed_refinement="restart", inner_eps=1e-6)),
    ("ppcg", "jac_diag", dict(mixed_refinement="restart", inner_eps=1e-4)),
    ("cg", "jac_block", dict(mixed_refinement="restart", inner_eps=1e-5)),
    ("cg", "none", dict(mixed_refinement="restart", inner_eps=1e-3)),
    ("ppcg", "jac_block", dict()),
    ("ppcg", "none", dict(reliable_delta=0.01)),
    ("cg", "jac_diag", dict(reliable_delta=0.3)),
    ("cg", "jac_block", dict(mixed_refinement="reliable")),
    ("chebyshev", "jac_block", dict(inner_eps=1e-5)),
    ("chebyshev", "jac_diag", dict(inner_eps=1e-4)),
    ("chebyshev", "none", dict(mixed_refinement="reliable", inner_eps=1e-5)),   # restarts regardless
]


@pytest.mark.parametrize("solver,prec,kw", MIXED_CASES)
def test_mixed_reaches_fp64_true_residual(solver, prec, kw, oracle_so, oracle_env, tmp_path):
    oracle_env("dot2")
    inner = dict(ppcg_inner_steps=10) if solver == "ppcg" else (dict(ch_cg_presteps=20) if solver == "chebyshev" else {})
    mixed = _run(tmp_path, oracle_so, "mixed", solver=solver, preconditioner=prec, inner_precision="fp32",
                 **kw, **inner)
    ref = _run(tmp_path, oracle_so, "fp64", solver=solver, preconditioner=prec, eps=1e-14, **inner)
    assert all(k >= 2 for k in mixed.refinements)
    for k, e in enumerate(mixed.exact_error):
        assert 0 < e < 1e-10, (k, e)
        assert mixed.conduction_error[k] == e          # the refinement loop stops on the true residual
    assert len(mixed.summaries) == len(ref.summaries)
    for a, b in zip(mixed.summaries, ref.summaries):
        np.testing.assert_allclose(a.values(), b.values(), rtol=1e-9)
    text = (tmp_path / "mixed.out").read_text()
    assert "FP32 refinements" in text


@pytest.mark.parametrize("solver,prec,kw", [MIXED_CASES[0], MIXED_CASES[4], MIXED_CASES[6],
                                             ("ppcg", "jac_block", dict(inner_precision="fp64"))])
def test_exact_error_is_the_final_true_residual(solver, prec, kw, oracle_so, oracle_env, tmp_path):
    """tl_check_result after a refinement loop reports the loop's last true
    residual instead of computing it again: it must be ||u0 - A u|| / ||r0||
    for the u the step ends with, checked here against numpy on the fields
    (r0 = u0 - A u0: no warm start, so the solve starts from u = u0)."""
    oracle_env("dot2")
    eng = Engine(str(oracle_so))
    measured = []

    def on_step(step):
        eng.refresh()
        h, ny, nx = eng.h, 80, 96
        u = eng.get_field(rt.FIELD_U)
        u0 = eng.get_field(rt.XFIELD_U0)
        kx, ky = eng.get_field(rt.XFIELD_KX), eng.get_field(rt.XFIELD_KY)

        def residual(v):
            c = v[h:h + ny, h:h + nx]
            kl, kr = kx[h:h + ny, h:h + nx], kx[h:h + ny, h + 1:h + nx + 1]
            kb, kt = ky[h:h + ny, h:h + nx], ky[h + 1:h + ny + 1, h:h + nx]
            av = ((1.0 + (kt + kb) + (kr + kl)) * c - (kt * v[h + 1:h + ny + 1, h:h + nx] + kb * v[h - 1:h + ny - 1, h:h + nx])
                  - (kr * v[h:h + ny, h + 1:h + nx + 1] + kl * v[h:h + ny, h - 1:h + nx - 1]))
            return u0[h:h + ny, h:h + nx] - av
        measured.append(np.linalg.norm(residual(u)) / np.linalg.norm(residual(u0)))

    inner = dict(ppcg_inner_steps=10) if solver == "ppcg" else {}
    kw = dict(kw)
    kw.setdefault("inner_precision", "fp32")
    deck = Deck(x_cells=96, y_cells=80, end_step=3, eps=1e-10, check_result=True, summary_frequency=1, states=LAYOUT,
                solver=solver, preconditioner=prec, **kw, **inner)
    res = run_deck(deck.write(tmp_path / "tea.in"), backend=oracle_so, on_step=on_step)
    assert len(measured) == len(res.exact_error) == 3
    for got, want in zip(res.exact_error, measured):
        assert 0 < got < 1e-10
        assert got == pytest.approx(want, rel=1e-4), (res.exact_error, measured)


@pytest.mark.parametrize("solver", ["cg", "ppcg", "chebyshev"])
def test_fp64_restart_is_the_fp64_solve(solver, oracle_so, oracle_env, tmp_path):
    """tl_inner_precision fp64 with restarts runs its inner solves on the FP64
    fields: with tl_inner_eps equal to a plain FP64 run's tl_eps, its first
    inner solve is that run, bit for bit, and already meets tl_eps, so it is
    the only one."""
    oracle_env("dot2")
    extra = dict(ppcg_inner_steps=10) if solver == "ppcg" else (dict(ch_cg_presteps=20) if solver == "chebyshev" else {})
    plain = _run(tmp_path, oracle_so, "plain", solver=solver, preconditioner="jac_block", eps=1e-13, **extra)
    rst = _run(tmp_path, oracle_so, "rst", solver=solver, preconditioner="jac_block", inner_precision="fp64",
               mixed_refinement="restart", inner_eps=1e-13, eps=1e-10, **extra)
    assert rst.refinements == [1] * len(plain.iterations)
    assert rst.iterations == plain.iterations
    for a, b in zip(rst.summaries, plain.summaries):
        assert a.values() == b.values()


def test_reliable_keeps_the_krylov_process(oracle_so, oracle_env, tmp_path):
    """Residual replacement keeps p, so the FP32 loop needs about the
    iterations of the FP64 loop with the same true-residual stop (the restart
    scheme needs more), and both reach the target."""
    oracle_env("dot2")
    base = dict(solver="ppcg", preconditioner="jac_block", ppcg_inner_steps=10, x_cells=128, y_cells=128)
    r32 = _run(tmp_path, oracle_so, "r32", inner_precision="fp32", **base)
    r64 = _run(tmp_path, oracle_so, "r64", inner_precision="fp64", **base)
    rst = _run(tmp_path, oracle_so, "rst", inner_precision="fp32", mixed_refinement="restart", inner_eps=1e-4, **base)
    for a, b in zip(r32.iterations, r64.iterations):
        assert abs(a - b) <= max(2, 0.1 * b)
    assert sum(rst.iterations) > sum(r32.iterations)
    for r in (r32, r64, rst):
        assert max(r.exact_error) < 1e-10
    assert all(k >= 1 for k in r64.refinements)


@pytest.mark.parametrize("bits", ["fp32", "fp64"])
def test_queued_inner_passes_change_nothing(bits, oracle_so, oracle_env, monkeypatch, tmp_path):
    """The residual-replacement loop may queue an outer iteration's inner
    passes before its replacement test; a replacement after all overwrites what
    they wrote and runs them again.  Queued when no replacement is expected
    (default), never (TEA_SPECULATE=0) and always (TEA_SPECULATE=2, so every
    replacement undoes queued passes): the same iterations, replacements,
    residuals and summaries."""
    oracle_env("dot2")
    base = dict(solver="ppcg", preconditioner="jac_block", ppcg_inner_steps=10, inner_precision=bits)
    runs = {}
    for mode in ("default", "0", "2"):
        if mode == "default":
            monkeypatch.delenv("TEA_SPECULATE", raising=False)
        else:
            monkeypatch.setenv("TEA_SPECULATE", mode)
        runs[mode] = _run(tmp_path, oracle_so, f"spec{mode}", **base)
    ref = runs["0"]
    assert sum(ref.refinements) >= 2 * len(ref.refinements)
    for mode in ("default", "2"):
        r = runs[mode]
        assert r.iterations == ref.iterations and r.ppcg_inner == ref.ppcg_inner
        assert r.refinements == ref.refinements
        assert r.exact_error == ref.exact_error
        for a, b in zip(r.summaries, ref.summaries):
            assert a.values() == b.values()


@pytest.mark.parametrize("kw,msg", [
    (dict(solver="jacobi", preconditioner="none"), "tl_inner_precision"),
    (dict(solver="ppcg", preconditioner="none", reflective_boundary=True), "reflective"),
])
def test_mixed_rejects_unsupported_decks(kw, msg, oracle_so, tmp_path):
    with pytest.raises(TeaError, match=msg):
        _run(tmp_path, oracle_so, "bad", inner_precision="fp32", end_step=1, **kw)


def test_fp32_mode_rejects_jacobi_calls(oracle_so, tmp_path):
    eng = Engine(oracle_so)
    try:
        setup_chunk(eng, tmp_path, solver="jacobi")
        eng.set_precision(32)
        with pytest.raises(TeaError):
            eng.jacobi_solve()
        with pytest.raises(TeaError):
            eng.set_precision(16)
    finally:
        eng.finalize()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_unreachable_target_stops_on_stagnation(prec, oracle_so, oracle_env, tmp_path):
    """tl_eps below FP64's attainable accuracy: the refinement loop notices the
    true residual no longer follows the recurrence and stops long before
    tl_max_iters, at the best true residual it reached."""
    oracle_env("dot2")
    r = _run(tmp_path, oracle_so, "stag", solver="ppcg", preconditioner="jac_block", ppcg_inner_steps=10,
             inner_precision=prec, eps=1e-18, max_iters=5000, end_step=2)
    assert all(n < 1000 for n in r.iterations), r.iterations
    assert all(e < 1e-13 for e in r.exact_error), r.exact_error
    assert "stagnated" in (tmp_path / "stag.out").read_text()
