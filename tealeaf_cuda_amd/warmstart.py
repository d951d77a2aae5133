"""Learned warm start: weight files, the GPU trainer, training data from the
solver, and the iteration-count evaluation (include/tea_warmstart.h).

The network's forward and backward passes are the HIP kernels of
csrc/warmstart.hip (BF16 matrix-core convolutions, FP32 accumulation).  Torch
is used here only to hold device buffers and to index batches out of the
sample store.

torch is imported before the engine library is loaded: both link a HIP
runtime of the same soname, and the first one loaded serves the process.

Training data: each seed's random-rectangle layout is solved for `steps`
timesteps by the plain FP64 PPCG solver (jac_block, 10 inner steps, tl_eps
1e-10); after every step the density, u^n (the step's u0) and u^{n+1} (its
solution) are copied off the engine, and from the second step on, random
crops of the features (density, u^n, u^n - u^{n-1}) and the target
(u^{n+1} - u^n - (u^n - u^{n-1})) / s become training samples
(tea_ws_make_samples).
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np
import torch  # before the engine library: torch's HIP runtime is then the process's one

from .deck import Deck, random_layout
from .runtime import FIELD_DENSITY, FIELD_U, XFIELD_KX, XFIELD_KY, XFIELD_U0, TeaError, engine_library, run_deck

MAGIC = b"TLWS0001"
WEIGHTS_DIR = Path(__file__).resolve().parent / "weights"

_F = ctypes.POINTER(ctypes.c_float)
_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_P = ctypes.c_void_p

_WS = {
    "tea_ws_param_count": ([], ctypes.c_int),
    "tea_ws_last_error": ([], ctypes.c_char_p),
    "tea_engine_warm_start_load": ([_F, ctypes.c_int], ctypes.c_int),
    "tea_engine_warm_start_apply": ([_I], None),
    "tea_engine_warm_start_reset": ([], None),
    "tea_engine_warm_start_coefficients": ([_D], None),
    "tea_engine_copy_field": ([ctypes.c_int, _P], ctypes.c_int),
    "tea_ws_trainer_create": ([ctypes.c_int] * 5, _P),
    "tea_ws_trainer_destroy": ([_P], None),
    "tea_ws_trainer_set_params": ([_P, _F], ctypes.c_int),
    "tea_ws_trainer_get_params": ([_P, _F], ctypes.c_int),
    "tea_ws_trainer_forward": ([_P, _P, _F], ctypes.c_int),
    "tea_ws_trainer_gradient": ([_P, _P, _F, _D], ctypes.c_int),
    "tea_ws_trainer_step": ([_P, _P, ctypes.c_float, _D], ctypes.c_int),
    "tea_ws_trainer_sync": ([_P], ctypes.c_int),
    "tea_ws_make_samples": ([_P] * 7 + [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, _P,
                                        ctypes.c_int, ctypes.c_int, _P, _D], ctypes.c_int),
}
WARMSTART_SYMBOLS = tuple(_WS)
NPLANES = 6                 # training sample planes: f0, f1, f2, target, Kx, Ky
LOSSES = {"mse": 0, "energy": 1}
_lib_cache = {}


def library():
    """The engine library with the warm-start entry points bound."""
    path = str(engine_library())
    if path not in _lib_cache:
        lib = ctypes.CDLL(path)
        for name, (args, res) in _WS.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        _lib_cache[path] = lib
    return _lib_cache[path]


def param_count() -> int:
    return library().tea_ws_param_count()


def _check(rc):
    if rc != 0:
        raise TeaError(library().tea_ws_last_error().decode())


# ----------------------------------------------------------------------------- weight files

def save_weights(path, params) -> Path:
    p = np.ascontiguousarray(params, dtype="<f4")
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    with open(path, "wb") as f:
        f.write(MAGIC)
        f.write(np.int32(p.size).astype("<i4").tobytes())
        f.write(p.tobytes())
    return path


def load_weights(path) -> np.ndarray:
    data = Path(path).read_bytes()
    if data[:8] != MAGIC:
        raise TeaError(f"{path}: not a warm-start weight file")
    n = int(np.frombuffer(data[8:12], dtype="<i4")[0])
    p = np.frombuffer(data[12:], dtype="<f4")
    if p.size != n:
        raise TeaError(f"{path}: {p.size} parameters, header says {n}")
    return p.astype(np.float32)


def default_weights(n: int) -> Path:
    """The shipped weights for an n x n mesh."""
    return WEIGHTS_DIR / f"warmstart_{n}.tlws"


# ----------------------------------------------------------------------------- trainer

class Trainer:
    """GPU trainer for batches of `batch` crops of height x width cells;
    loss "mse" or "energy" (tea_warmstart.h)."""

    def __init__(self, batch: int, height: int, width: int, margin: int = 6, loss: str = "energy"):
        self.lib = library()
        self.shape = (batch, height, width)
        self.h = self.lib.tea_ws_trainer_create(batch, height, width, margin, LOSSES[loss])
        if not self.h:
            raise TeaError(self.lib.tea_ws_last_error().decode())

    def close(self):
        if self.h:
            self.lib.tea_ws_trainer_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        _check(self.lib.tea_ws_trainer_set_params(self.h, p.ctypes.data_as(_F)))

    def get_params(self) -> np.ndarray:
        p = np.empty(param_count(), dtype=np.float32)
        _check(self.lib.tea_ws_trainer_get_params(self.h, p.ctypes.data_as(_F)))
        return p

    def _ptr(self, samples):
        b, hh, ww = self.shape
        if tuple(samples.shape) != (b, NPLANES, hh, ww) or not samples.is_contiguous() or samples.dtype.itemsize != 4:
            raise TeaError(f"samples must be a contiguous float32 device tensor of shape {(b, NPLANES, hh, ww)}")
        return ctypes.c_void_p(samples.data_ptr())

    def forward(self, samples) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.float32)
        _check(self.lib.tea_ws_trainer_forward(self.h, self._ptr(samples), out.ctypes.data_as(_F)))
        return out

    def gradient(self, samples):
        g = np.empty(param_count(), dtype=np.float32)
        loss = ctypes.c_double()
        _check(self.lib.tea_ws_trainer_gradient(self.h, self._ptr(samples), g.ctypes.data_as(_F),
                                                ctypes.byref(loss)))
        return loss.value, g

    def step(self, samples, lr: float, want_loss: bool = False):
        loss = ctypes.c_double()
        _check(self.lib.tea_ws_trainer_step(self.h, self._ptr(samples), float(lr),
                                            ctypes.byref(loss) if want_loss else None))
        return loss.value if want_loss else None

    def sync(self):
        _check(self.lib.tea_ws_trainer_sync(self.h))


# ----------------------------------------------------------------------------- data

def solver_deck(n: int, seed: int, steps: int, **kw) -> Deck:
    """Plain FP64 PPCG on seed's random-rectangle layout (the data source)."""
    base = dict(x_cells=n, y_cells=n, end_step=steps, solver="ppcg", preconditioner="jac_block", eps=1e-10,
                ppcg_inner_steps=10, halo_depth=1, summary_frequency=0, states=random_layout(seed, cells=n))
    base.update(kw)
    return Deck(**base)


@dataclass
class SampleStore:
    """Device-resident training samples [capacity, 6, crop, crop]."""
    capacity: int
    crop: int
    device: str = "cuda"
    count: int = 0
    data: object = None

    def __post_init__(self):
        self.data = torch.empty((self.capacity, NPLANES, self.crop, self.crop), dtype=torch.float32,
                                device=self.device)

    def batch(self, idx):
        return self.data[idx].contiguous()


@dataclass
class GenStats:
    steps: int = 0
    samples: int = 0
    solve_seconds: float = 0.0
    iterations: list = field(default_factory=list)


def generate(seeds, n: int, steps: int, crops_per_step: int, store: SampleStore, tmpdir, rng=None,
             log=print, residual_of: str = "galerkin") -> GenStats:
    """Runs the solver over `seeds` and fills `store` with crops.  The target
    is the part of each step that a guess misses: residual_of "extrapolation"
    (u^n + (u^n - u^{n-1})) or "galerkin" (the engine's energy-minimising
    combination of the last two previous-step differences, which the solves
    then also start from -- tl_warm_start_history 2 and tl_warm_start_norm
    energy, the guess the samples' two coefficients restate; the solution is
    the same to the solver tolerance)."""
    lib = library()
    rng = rng or np.random.default_rng(0)
    st = GenStats()
    dev = store.device
    bufs = {k: torch.empty((n, n), dtype=torch.float64, device=dev)
            for k in ("rho", "prev2", "prev", "u0", "u", "kx", "ky")}
    coef = (ctypes.c_double * 3)()
    extra = ({"warm_start": "extrapolate", "warm_start_history": 2, "warm_start_norm": "energy"}
             if residual_of == "galerkin" else {})
    origins = torch.empty(2 * crops_per_step, dtype=torch.int32, device=dev)
    scales = (ctypes.c_double * 2)()
    crop = store.crop
    for k, seed in enumerate(seeds):
        deck = solver_deck(n, seed, steps, **extra).write(Path(tmpdir) / "gen.in")
        state = {"nprev": 0}

        def on_step(step):
            for fid, key in ((FIELD_DENSITY, "rho"), (XFIELD_U0, "u0"), (FIELD_U, "u"), (XFIELD_KX, "kx"),
                             (XFIELD_KY, "ky")):
                if key in ("rho", "kx", "ky") and step > 1:
                    continue
                if lib.tea_engine_copy_field(fid, ctypes.c_void_p(bufs[key].data_ptr())) != 0:
                    raise TeaError("tea_engine_copy_field failed")
            if state["nprev"] >= 1 and store.count < store.capacity:
                if residual_of == "galerkin":
                    lib.tea_engine_warm_start_coefficients(coef)
                    c0, c1 = coef[0], coef[1]
                else:
                    c0, c1 = 1.0, 0.0
                if state["nprev"] < 2:
                    c1 = 0.0
                m = min(crops_per_step, store.capacity - store.count)
                o = rng.integers(0, n - crop + 1, size=(m, 2)).astype(np.int32)
                origins[:2 * m].copy_(torch.from_numpy(o.ravel()))
                out = store.data[store.count:store.count + m]
                rc = lib.tea_ws_make_samples(*(ctypes.c_void_p(bufs[x].data_ptr())
                                               for x in ("rho", "prev2", "prev", "u0", "u", "kx", "ky")),
                                             n, n, c0, c1, ctypes.c_void_p(origins.data_ptr()), m, crop,
                                             ctypes.c_void_p(out.data_ptr()), scales)
                if rc == 0:
                    store.count += m
                    st.samples += m
            bufs["prev2"].copy_(bufs["prev"] if state["nprev"] >= 1 else bufs["u0"])
            bufs["prev"].copy_(bufs["u0"])
            torch.cuda.synchronize()
            state["nprev"] += 1

        t0 = time.perf_counter()
        res = run_deck(deck, on_step=on_step)
        st.solve_seconds += time.perf_counter() - t0
        st.steps += len(res.iterations)
        st.iterations.append(list(res.iterations))
        if log and (k % 10 == 0 or k == len(seeds) - 1):
            log(f"[warmstart] data: seed {seed} ({k + 1}/{len(seeds)}), {store.count} samples, "
                f"{st.solve_seconds:.1f} s solving")
    return st


def train(trainer: Trainer, store: SampleStore, epochs: float, lr: float, seed: int = 0, log=print,
          log_every: int = 200):
    """Adam over random batches of the store, cosine learning-rate decay."""
    b = trainer.shape[0]
    steps = max(1, int(epochs * store.count / b))
    g = torch.Generator(device="cpu")
    g.manual_seed(seed)
    history = []
    t0 = time.perf_counter()
    for s in range(steps):
        idx = torch.randint(0, store.count, (b,), generator=g).to(store.device)
        rate = lr * 0.5 * (1.0 + np.cos(np.pi * s / steps))
        want = s % log_every == 0 or s == steps - 1
        loss = trainer.step(store.batch(idx), rate, want_loss=want)
        if want:
            history.append((s, loss))
            if log:
                log(f"[warmstart] train step {s}/{steps} loss {loss:.4e} ({time.perf_counter() - t0:.1f} s)")
    trainer.sync()
    return history


# ----------------------------------------------------------------------------- evaluation

def evaluate(seeds, n: int, steps: int, variants: dict, tmpdir, inner_precision: str = "fp64", log=print):
    """Per seed and variant (name -> tl_warm_start value, None for the plain
    u = u^n guess): iterations, step times and true residuals over `steps`
    timesteps to the FP64 true residual 1e-10 of the u = u^n guess."""
    rows = []
    for k, seed in enumerate(seeds):
        row = {"seed": seed}
        for tag, ws in variants.items():
            extra = {} if ws is None else {"warm_start": str(ws)}
            deck = solver_deck(n, seed, steps, inner_precision=inner_precision, check_result=True, **extra)
            r = run_deck(deck.write(Path(tmpdir) / f"eval_{tag}.in"))
            row[tag] = {"iterations": r.iterations, "seconds": r.step_seconds, "exact": r.exact_error}
        rows.append(row)
        if log and (k % 10 == 0 or k == len(seeds) - 1):
            tot = {t: sum(sum(r[t]["iterations"][1:]) for r in rows) for t in variants}
            log(f"[warmstart] eval seed {seed} ({k + 1}/{len(seeds)}): iterations after step 1 {tot}")
    return rows


def summarise(rows, variants, base="plain") -> dict:
    """Totals over steps 2.. (the first step has no previous step to use)."""
    out = {}
    for t in variants:
        its = sum(sum(r[t]["iterations"][1:]) for r in rows)
        sec = sum(sum(r[t]["seconds"][1:]) for r in rows)
        worst = max(max(r[t]["exact"][1:]) for r in rows)
        out[t] = {"iterations": its, "seconds": round(sec, 4), "max_true_residual": worst}
    for t in variants:
        if t != base:
            out[t]["iterations_saved"] = round(1.0 - out[t]["iterations"] / out[base]["iterations"], 4)
            out[t]["time_saved"] = round(1.0 - out[t]["seconds"] / out[base]["seconds"], 4)
    return out


def _range(text):
    a, _, b = text.partition(":")
    return range(int(a), int(b))


def main(argv=None):
    """python -m tealeaf_cuda_amd.warmstart train|evaluate ... (see --help)."""
    import argparse
    import json
    import sys
    import tempfile

    ap = argparse.ArgumentParser(prog="python -m tealeaf_cuda_amd.warmstart")
    sub = ap.add_subparsers(dest="cmd", required=True)
    tr = sub.add_parser("train", help="generate solves, train on the GPU, write a weight file")
    tr.add_argument("--size", type=int, default=1024)
    tr.add_argument("--seeds", default="0:2000")
    tr.add_argument("--steps", type=int, default=20)
    tr.add_argument("--crops", type=int, default=4, help="crops per timestep")
    tr.add_argument("--crop", type=int, default=128)
    tr.add_argument("--batch", type=int, default=32)
    tr.add_argument("--iters", type=int, default=20000)
    tr.add_argument("--lr", type=float, default=1e-3)
    tr.add_argument("--loss", choices=sorted(LOSSES), default="energy")
    tr.add_argument("--target", choices=["galerkin", "extrapolation"], default="galerkin",
                    help="the guess whose remaining error the network learns")
    tr.add_argument("--init-seed", type=int, default=0)
    tr.add_argument("--out", required=True)
    ev = sub.add_parser("evaluate", help="iterations to 1e-10 with and without the warm start")
    ev.add_argument("--size", type=int, default=1024)
    ev.add_argument("--seeds", default="10000:10200")
    ev.add_argument("--steps", type=int, default=20)
    ev.add_argument("--weights", required=True)
    ev.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    ev.add_argument("--json", default=None)
    a = ap.parse_args(argv)
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    with tempfile.TemporaryDirectory() as tmp:
        if a.cmd == "train":
            seeds = _range(a.seeds)
            store = SampleStore(capacity=len(seeds) * (a.steps - 1) * a.crops, crop=a.crop)
            t0 = time.perf_counter()
            st = generate(seeds, a.size, a.steps, a.crops, store, tmp, rng=np.random.default_rng(a.init_seed), log=log,
                          residual_of=a.target)
            gen_s = time.perf_counter() - t0
            trainer = Trainer(a.batch, a.crop, a.crop, margin=6, loss=a.loss)
            init = _initial_params(a.init_seed)
            trainer.set_params(init)
            t0 = time.perf_counter()
            hist = train(trainer, store, a.iters * a.batch / store.count, a.lr, seed=a.init_seed, log=log,
                         log_every=1000)
            train_s = time.perf_counter() - t0
            save_weights(a.out, trainer.get_params())
            trainer.close()
            print(json.dumps({"weights": str(a.out), "size": a.size, "seeds": [seeds.start, seeds.stop],
                              "steps": a.steps, "samples": store.count, "solver_steps": st.steps,
                              "data_seconds": round(gen_s, 1), "train_iters": a.iters, "batch": a.batch,
                              "train_seconds": round(train_s, 1), "loss": a.loss, "target": a.target,
                              "loss_history": [(s, float(l)) for s, l in hist]}))
        else:
            variants = {"plain": None, "extrapolate": "extrapolate", "network": str(Path(a.weights).resolve())}
            rows = evaluate(_range(a.seeds), a.size, a.steps, variants, tmp, inner_precision=a.precision, log=log)
            summary = {"size": a.size, "seeds": a.seeds, "steps": a.steps, "precision": a.precision,
                       "weights": str(a.weights), "result": summarise(rows, variants)}
            if a.json:
                Path(a.json).write_text(json.dumps({"summary": summary, "rows": rows}))
            print(json.dumps(summary))


def _initial_params(seed: int) -> np.ndarray:
    """He-style initial weights; the output layer starts at zero (the first
    guess is then the Galerkin combination of the previous-step directions)."""
    rng = np.random.default_rng(seed)
    n = param_count()
    p = np.zeros(n, dtype=np.float64)
    k1, ch, taps = 27, 32, 9
    p[:k1 * ch] = rng.standard_normal(k1 * ch) * np.sqrt(2.0 / k1)
    o = k1 * ch + ch
    for _ in range(4):
        p[o:o + taps * ch * ch] = rng.standard_normal(taps * ch * ch) * np.sqrt(2.0 / (taps * ch))
        o += taps * ch * ch + ch
    return p.astype(np.float32)


if __name__ == "__main__":
    main()
