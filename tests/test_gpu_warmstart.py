"""GPU tests of the learned warm start: the HIP network kernels against the
CPU oracle (oracle/cnn_oracle.py, bfloat16 rounding emulated at the points
where the kernels store BF16), training on the GPU, the engine's inference
on its own fields, and whole decks that use it.

Tolerances: the GPU accumulates in FP32 in its own order, so a BF16 value
that sits at a rounding boundary can round the other way (one BF16 unit,
2^-8 relative); relative norm errors of 5e-3 (forward) and 2e-2 (gradient
blocks) leave room for a few such flips and catch any wrong term.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest
import torch  # noqa: F401  (first: torch's HIP runtime must be the process's one)

from conftest import REPO
from tealeaf_cuda_amd import Deck, Engine, run_deck
from tealeaf_cuda_amd import runtime as rt
from tealeaf_cuda_amd import warmstart as wsm

sys.path.insert(0, str(REPO / "oracle"))
import cnn_oracle as co  # noqa: E402

pytestmark = pytest.mark.gpu


def _params(seed, w6_scale=0.05):
    rng = np.random.default_rng(seed)
    p = co.init_params(seed).astype(np.float64)
    p[co.OFF_W6:co.OFF_B6] = rng.standard_normal(co.OFF_B6 - co.OFF_W6) * w6_scale
    p[co.OFF_B1:co.OFF_HID] = rng.standard_normal(co.CH) * 0.1
    p[co.OFF_B6] = 0.01
    return p.astype(np.float32)


def _samples(seed, B, H, W):
    rng = np.random.default_rng(seed)
    s = rng.standard_normal((B, 6, H, W)).astype(np.float32)
    s[:, 4:] = rng.uniform(0.1, 3.0, (B, 2, H, W))
    return s, torch.from_numpy(s).cuda()


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


def test_forward_matches_oracle():
    print("[gpu warmstart] forward vs oracle", flush=True)
    B, H, W = 2, 8, 64
    t = wsm.Trainer(B, H, W, margin=1)
    try:
        p = _params(1)
        t.set_params(p)
        s, sd = _samples(2, B, H, W)
        out = t.forward(sd)
        ref, _ = co.forward(p, s, mode="bf16")
        assert _rel(out, ref) < 5e-3
        exact, _ = co.forward(p, s)
        assert _rel(out, exact) < 5e-2
    finally:
        t.close()



@pytest.mark.parametrize("run_rows", [None, 48, 64, 96])
def test_marching_layers_match_windowed(run_rows, monkeypatch):
    """One image with a height a multiple of 16: the hidden and output layers
    march down runs of rows (k_conv32_march, k_conv_out_march; runs of several
    blocks forced with TEA_WS_MARCH_ROWS) and give the windowed kernels'
    output bit for bit (TEA_WS_MARCH=0), which the oracle also checks."""
    print(f"[gpu warmstart] marching layers, runs of {run_rows} rows", flush=True)
    B, H, W = 1, 96, 128
    t = wsm.Trainer(B, H, W, margin=1)
    try:
        p = _params(3)
        t.set_params(p)
        s, sd = _samples(4, B, H, W)
        monkeypatch.setenv("TEA_WS_MARCH", "0")
        windowed = t.forward(sd)
        monkeypatch.delenv("TEA_WS_MARCH")
        if run_rows:
            monkeypatch.setenv("TEA_WS_MARCH_ROWS", str(run_rows))
        marched = t.forward(sd)
        assert np.array_equal(marched, windowed)
        ref, _ = co.forward(p, s, mode="bf16")
        assert _rel(marched, ref) < 5e-3
    finally:
        t.close()

@pytest.mark.parametrize("loss_kind", ["mse", "energy"])
def test_gradient_matches_oracle(loss_kind):
    print(f"[gpu warmstart] gradient vs oracle ({loss_kind})", flush=True)
    B, H, W = 2, 8, 64
    t = wsm.Trainer(B, H, W, margin=1, loss=loss_kind)
    try:
        p = _params(3)
        t.set_params(p)
        s, sd = _samples(4, B, H, W)
        loss, g = t.gradient(sd)
        rl, rg = co.loss_and_grad(p, s, margin=1, mode="bf16", loss=loss_kind)
        assert loss == pytest.approx(rl, rel=1e-3)
        blocks = [("w1", co.OFF_W1, co.OFF_B1), ("b1", co.OFF_B1, co.OFF_HID)]
        for l in range(4):
            o = co.OFF_HID + l * co.HID
            blocks += [(f"w{l + 2}", o, o + co.TAPS * co.CH * co.CH), (f"b{l + 2}", o + co.TAPS * co.CH * co.CH,
                                                                      o + co.HID)]
        blocks += [("w6", co.OFF_W6, co.OFF_B6), ("b6", co.OFF_B6, co.NPARAM)]
        for name, lo, hi in blocks:
            assert _rel(g[lo:hi], rg[lo:hi]) < 2e-2, name
    finally:
        t.close()


def test_training_fits_a_learnable_target():
    """Target = a fixed linear stencil of the features: Adam on the GPU must
    cut the loss by a large factor."""
    print("[gpu warmstart] training", flush=True)
    B, H, W = 4, 16, 64
    rng = np.random.default_rng(5)
    s = rng.standard_normal((B, 6, H, W)).astype(np.float32)
    s[:, 4:] = 0.5
    f = s[:, :3]
    s[:, 3] = 0.5 * f[:, 1] - 0.25 * np.roll(f[:, 1], 1, axis=2) + 0.3 * f[:, 0] * (f[:, 2] > 0)
    sd = torch.from_numpy(s).cuda()
    t = wsm.Trainer(B, H, W, margin=1, loss="mse")
    try:
        t.set_params(co.init_params(6))
        first = t.step(sd, 1e-3, want_loss=True)
        for _ in range(300):
            t.step(sd, 1e-3)
        last, _ = t.gradient(sd)
        assert last < 0.05 * first, (first, last)
    finally:
        t.close()


def _operator(eng, h, nx, ny):
    """The interior 5-point operator I - dt div K grad from the engine's Kx/Ky."""
    kx, ky = eng.get_field(rt.XFIELD_KX), eng.get_field(rt.XFIELD_KY)

    def A(v):
        p = np.zeros(eng.shape())
        p[h:h + ny, h:h + nx] = v
        c = (slice(h, h + ny), slice(h, h + nx))
        kr, kl = kx[h:h + ny, h + 1:h + nx + 1], kx[c]
        ku, kd = ky[h + 1:h + ny + 1, h:h + nx], ky[c]
        return ((1 + ku + kd + kr + kl) * v - ku * p[h + 1:h + ny + 1, h:h + nx] - kd * p[h - 1:h + ny - 1, h:h + nx]
                - kr * p[h:h + ny, h + 1:h + nx + 1] - kl * p[h:h + ny, h - 1:h + nx - 1])
    return A


@pytest.mark.parametrize("with_net,history,norm", [(True, None, None), (False, None, None), (True, 4, "energy"),
                                                  (False, 3, None), (False, 1, "energy"), (False, 2, "residual"),
                                                  (True, 8, None), (False, 6, "energy")])
def test_engine_apply_matches_oracle(with_net, history, norm, tmp_path):
    """tea_engine_warm_start_apply on engine fields over a history of solves
    (each u0 the exact solution of A u = the previous u0, as the timesteps
    leave them): the guess u0 + sum c_a v_a over the differences of the last
    `history` steps (default 8) and the network's direction, with the
    coefficients that minimise the residual's 2-norm (default) or the error's
    energy norm, computed here by least squares from the oracle's network
    output and a dense restatement of the operator; the guess's residual
    (energy) is never above that of u = u0."""
    import ctypes

    print(f"[gpu warmstart] engine apply (network {with_net}, history {history}, norm {norm})", flush=True)
    from test_oracle_math import setup_chunk

    lib = wsm.library()
    eng = Engine()
    hist = history or 8
    try:
        nx, ny = 64, 8
        setup_chunk(eng, tmp_path, nx=nx, ny=ny, solver="cg", halo=1,
                    extra=([f"tl_warm_start_history={history}"] if history else []) +
                    ([f"tl_warm_start_norm={norm}"] if norm else []))
        p = _params(7, w6_scale=0.2)
        if with_net:
            assert lib.tea_engine_warm_start_load(p.ctypes.data_as(wsm._F), p.size) == 0
        else:
            assert lib.tea_engine_warm_start_load(None, 0) == 0
        rng = np.random.default_rng(8)
        h = eng.h
        shape = eng.shape()
        A = _operator(eng, h, nx, ny)
        eye = np.eye(nx * ny)
        Am = np.stack([A(eye[j].reshape(ny, nx)).ravel() for j in range(nx * ny)], axis=1)

        def put(fid, vals):
            a = np.zeros(shape)
            a[h:h + ny, h:h + nx] = vals
            eng.set_field(fid, a)

        rho = 10.0 ** rng.uniform(-1, 2, (ny, nx))
        put(rt.FIELD_DENSITY, rho)
        us = [rng.uniform(1.0, 3.0, (ny, nx))]
        for _ in range(hist + 2):
            us.append(np.linalg.solve(Am, us[-1].ravel()).reshape(ny, nx))
        applied = ctypes.c_int(-1)
        for k, u0 in enumerate(us):
            put(rt.XFIELD_U0, u0)
            put(rt.FIELD_U, u0)
            lib.tea_engine_warm_start_apply(ctypes.byref(applied))
            assert applied.value == (1 if k else 0)
            if k == 0:
                continue
            dirs = [us[k - j] - us[k - j - 1] for j in range(min(k, hist))]
            if with_net:
                d = dirs[0]
                sc, su = np.sqrt(np.mean(d * d)), np.sqrt(np.mean(u0 * u0))
                planes = np.stack([(np.log10(rho) - 0.5) / 1.5, d / sc, u0 / su]).astype(np.float32)[None]
                net, _ = co.forward(p, planes, mode="bf16")
                dirs.append(sc * net[0])
            V = np.stack([v.ravel() for v in dirs], axis=1)
            AV = Am @ V
            r0 = (u0 - A(u0)).ravel()
            if norm == "energy":
                c = np.linalg.lstsq(V.T @ AV, V.T @ r0, rcond=1e-14)[0]
            else:
                c = np.linalg.lstsq(AV, r0, rcond=1e-14)[0]
            expect = u0 + (V @ c).reshape(ny, nx)
            u = eng.get_field(rt.FIELD_U)[h:h + ny, h:h + nx]
            tol = (2e-2 if with_net else 1e-6) * np.abs(expect - u0).max()
            assert np.abs(u - expect).max() <= tol, (k, np.abs(u - expect).max(), tol)
            r0 = r0.reshape(ny, nx)
            r = u0 - A(u)
            if norm == "energy":   # ||e||_A^2 = r . A^-1 r; compare through e0 - e = u - u0
                d = u - u0
                assert np.sum(d * A(d)) - 2 * np.sum(d * r0) <= 1e-9 * np.sum(r0 * r0)
            else:
                assert np.sum(r * r) <= np.sum(r0 * r0) * (1 + 1e-12)
    finally:
        eng.finalize()


def test_warm_start_deck_reaches_the_same_target(tmp_path):
    """A deck with tl_warm_start: the guess is used from step 2 on and every
    step still reaches ||u0 - A u|| < 1e-10 ||u0 - A u^n||."""
    print("[gpu warmstart] deck", flush=True)
    wpath = wsm.save_weights(tmp_path / "w.tlws", _params(9, w6_scale=0.01))
    for prec, ws in (("fp64", str(wpath)), ("fp32", str(wpath)), ("fp64", "extrapolate")):
        deck = wsm.solver_deck(128, 21, 4, inner_precision=prec, check_result=True, warm_start=ws)
        r = run_deck(deck.write(tmp_path / f"ws_{prec}.in"), out_path=tmp_path / f"ws_{prec}.out")
        assert r.warm_start == [0, 1, 1, 1]
        assert all(0 < e < 1e-10 for e in r.exact_error), r.exact_error
        assert "Warm start" in (tmp_path / f"ws_{prec}.out").read_text()


def test_make_samples_matches_numpy():
    """Training crops: features, the target of the guess with coefficients
    (c0, c1), and the face coefficients; cells outside the mesh are zero."""
    import ctypes

    print("[gpu warmstart] make_samples", flush=True)
    lib = wsm.library()
    rng = np.random.default_rng(11)
    nx, ny, crop = 96, 80, 32
    f = {k: rng.uniform(0.5, 2.0, (ny, nx)) for k in ("rho", "p2", "p1", "u0", "un", "kx", "ky")}
    dev = {k: torch.from_numpy(v).cuda() for k, v in f.items()}
    origins = np.array([[0, 0], [40, 30], [80, 60]], dtype=np.int32)
    od = torch.from_numpy(origins.ravel()).cuda()
    out = torch.zeros((3, wsm.NPLANES, crop, crop), dtype=torch.float32, device="cuda")
    scales = (ctypes.c_double * 2)()
    c0, c1 = 0.9, 0.2
    rc = lib.tea_ws_make_samples(*(ctypes.c_void_p(dev[k].data_ptr()) for k in ("rho", "p2", "p1", "u0", "un",
                                                                                  "kx", "ky")),
                                 nx, ny, c0, c1, ctypes.c_void_p(od.data_ptr()), 3, crop,
                                 ctypes.c_void_p(out.data_ptr()), scales)
    assert rc == 0
    d = f["u0"] - f["p1"]
    s, su = np.sqrt(np.mean(d * d)), np.sqrt(np.mean(f["u0"] ** 2))
    assert scales[0] == pytest.approx(s, rel=1e-12) and scales[1] == pytest.approx(su, rel=1e-12)
    planes = [(np.log10(f["rho"]) - 0.5) / 1.5, d / s, f["u0"] / su,
              ((f["un"] - f["u0"]) - c0 * d - c1 * (f["p1"] - f["p2"])) / s, f["kx"], f["ky"]]
    got = out.cpu().numpy()
    for k, (ox, oy) in enumerate(origins):
        for q, pl in enumerate(planes):
            ref = np.zeros((crop, crop))
            w, h = min(crop, nx - ox), min(crop, ny - oy)
            ref[:h, :w] = pl[oy:oy + h, ox:ox + w]
            np.testing.assert_allclose(got[k, q], ref.astype(np.float32), rtol=2e-7, atol=1e-7, err_msg=f"{k} {q}")
