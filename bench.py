#!/usr/bin/env python3
"""Benchmark: ms per timestep to an FP64 relative residual of 1e-10 at 4096^2.

Workload: a seeded layout of random axis-aligned rectangles (density
log-uniform in [0.1, 100], energy log-uniform in [1e-4, 25], on a density-100 /
energy-1e-4 background) on a 4096 x 4096 mesh of [0,10]^2, PPCG with 10 inner
steps and the block-Jacobi preconditioner, dt 0.004, one chunk on one GPU.  A
step is one timestep: solver set-up, the CG pre-steps / PPCG iterations, and
the finalisation, all through the host driver over libtea_engine.so.

Five solves of the same deck are timed, each W warm-up steps (the first holds
the CG eigenvalue estimate) then exactly K timed steps.  Every solve but the
last stops when the FP64 true residual ||u0 - A u||_2 < 1e-10 ||r0||_2, r0
the residual of the u = u^n guess:

* headline: learned warm start + FP32 fields with FP64 dot accumulators and
  FP64 residual replacement (the network's weights were trained on the GPU
  from plain FP64 PPCG solves at this mesh size);
* fp32_extrapolate: the same without the network (the energy-minimising
  combination of the previous steps' differences only);
* fp32: FP32 refinement from the u = u^n guess;
* fp64_true: plain FP64 PPCG with the same true-residual stop;
* fp64_reference: plain FP64 PPCG stopping on the reference's own recurrence
  test at tl_eps 1e-10 (its true residual is reported beside it).

The timed region is bracketed by device synchronisation; HIP events recorded
on the engine's own stream give the GPU time of the region, and the engine's
per-launch profiler (start and stop timestamps of each dispatch) times the
inner-step passes, the dominant kernels, inside it for the roofline figure.
Two further steps, outside the timed region, time every launch for the
per-kernel breakdown.

The CPU baseline runs the plain FP64 deck through the same driver over the CPU
oracle (OpenMP, plain reductions) on a bounded sample of PPCG iterations and
extrapolates to the GPU plain run's iteration counts.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO))

METRIC = "ms/timestep to FP64 rel-res 1e-10 at 4096²: learned+FP32 vs plain FP64 PPCG"
HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per interior cell of each kernel launch: every distinct
# FP64 array the reference kernel reads or writes, once, 8 bytes each.
ALGO_BYTES = {
    # one whole inner step (the reference's update_r + calc_sd_new launches, jac_block)
    "k_ppcg_inner_fused": 112,
    "k_ppcg_inner_fused_f32": 56,     # the same arrays in FP32
    # n whole inner steps per launch (temporally blocked): n times the reference's bytes
    **{f"k_ppcg_inner{n}": 112 * n for n in range(1, 6)},
    **{f"k_ppcg_inner{n}_f32": 56 * n for n in range(1, 6)},
    "ppcg_update_r": 40,              # rtemp, sd, Kx, Ky read; rtemp written
    "k_ppcg_calc_sd": 72,             # jac_block: rtemp, sd, utemp, cp, bfp, Kx, Ky read; sd, utemp written
    "cg_calc_w": 32,                  # p, Kx, Ky read; w written
    "k_cg_calc_ur": 88,               # jac_block: u, p, r, w, cp, bfp, Kx, Ky read; u, r, z written
    "ppcg_calc_rrn": 24,
    "ppcg_calc_2norm": 16,
    "ppcg_store_r": 16,
    "ppcg_update_z": 16,
    "ppcg_calc_p": 24,
    "ppcg_init_sd_new": 40,           # r, z read; sd, rtemp, utemp written
}


LAYOUT_SEED = 4096


def weights_for(n):
    return REPO / "tealeaf_cuda_amd" / "weights" / f"warmstart_{n}.tlws"


WEIGHTS = {}


def variants(n):
    return {
        "headline": dict(inner_precision="fp32", warm_start=str(WEIGHTS.get(n, weights_for(n)))),
        "fp32_extrapolate": dict(inner_precision="fp32", warm_start="extrapolate"),
        "fp32": dict(inner_precision="fp32"),
        "fp64_true": dict(inner_precision="fp64"),
        "fp64_reference": dict(),
        # the plain FP64 solvers on the reference's rectangle layout (run with --only)
        "ref_layout_cg_diag": dict(solver="cg", prec="jac_diag", layout="reference"),
        "ref_layout_chebyshev": dict(solver="chebyshev", prec="none", layout="reference", ch_cg_presteps=20),
        "ref_layout_ppcg_block": dict(layout="reference"),
        # Chebyshev (block Jacobi) refined by restarts to the FP64 true residual, FP64 or FP32 inner solves
        "cheby_block_fp64": dict(solver="chebyshev", ch_cg_presteps=20, inner_precision="fp64", inner_eps=1e-4),
        "cheby_block_fp32": dict(solver="chebyshev", ch_cg_presteps=20, inner_precision="fp32", inner_eps=1e-4),
    }


DEFAULT_VARIANTS = ("headline", "fp32_extrapolate", "fp32", "fp64_true", "fp64_reference")


def make_deck(path, n, steps, solver="ppcg", prec="jac_block", eps=1e-10, max_iters=10000, inner=10,
              epslim=None, layout="random", **kw):
    from tealeaf_cuda_amd import Deck, random_layout

    if layout == "random":
        kw["states"] = random_layout(LAYOUT_SEED, cells=n)   # else the deck's default: the reference's rectangles
    d = Deck(x_cells=n, y_cells=n, end_step=steps, solver=solver, preconditioner=prec, eps=eps,
             max_iters=max_iters, halo_depth=1, ppcg_inner_steps=inner, summary_frequency=0,
             check_result=True, ch_cg_epslim=epslim, **kw)
    return d.write(path)


# Launches timed one by one inside the timed region: the inner passes (the
# dominant kernels) only, as every timed launch costs the run a few
# microseconds.  PROFILED_STEPS further steps time every launch for the
# per-kernel breakdown.
TIMED_PREFIX = "k_ppcg_inner"
PROFILED_STEPS = 2


def run_gpu(n, warmup, steps, tmp, variant):
    import torch
    import tealeaf_cuda_amd as tl

    deck = make_deck(Path(tmp) / f"tea_{variant}.in", n, warmup + steps + PROFILED_STEPS, **variants(n)[variant])
    eng = tl.Engine()  # same library instance the driver loads (RTLD_LOCAL handles are shared per path)
    ev = {}

    def on_step(step):
        stream = torch.cuda.ExternalStream(eng.lib.tea_engine_stream())
        if step > warmup and (step - warmup) % 20 == 0 and step < warmup + steps:
            print(f"[bench] {variant}: {step - warmup} of {steps} timed steps done", file=sys.stderr, flush=True)
        if step == warmup:
            eng.synchronize()
            eng.profile_reset()
            eng.profile_only(TIMED_PREFIX)
            eng.profile_enable(True)
            ev["start"] = torch.cuda.Event(enable_timing=True)
            ev["start"].record(stream)
            ev["t0"] = time.perf_counter()
        elif step == warmup + steps:
            ev["stop"] = torch.cuda.Event(enable_timing=True)
            ev["stop"].record(stream)
            ev["stop"].synchronize()
            ev["t1"] = time.perf_counter()
            eng.profile_enable(False)
            ev["profile"] = eng.profile()
            eng.profile_reset()
            eng.profile_only("")
            eng.profile_enable(True)
        elif step == warmup + steps + PROFILED_STEPS:
            eng.profile_enable(False)
            ev["breakdown"] = eng.profile()

    res = tl.run_deck(deck, on_step=on_step)
    gpu_ms = ev["start"].elapsed_time(ev["stop"])
    wall_ms = (ev["t1"] - ev["t0"]) * 1e3
    timed = slice(warmup, warmup + steps)
    print(f"[bench] {variant}: {gpu_ms / steps:.1f} ms/step, iterations {res.iterations[timed]}, "
          f"true residual {res.exact_error[timed]}", file=sys.stderr, flush=True)
    return res, gpu_ms, wall_ms, ev["profile"], timed, ev["breakdown"]


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summaries
    (lines 'TRAFFIC <kernel> <bytes>' in profiles/*.md), or None."""
    for p in sorted((REPO / "profiles").glob("*.md")):
        for line in p.read_text().splitlines():
            parts = line.split()
            if len(parts) == 3 and parts[0] == "TRAFFIC" and parts[1] == kernel:
                return int(parts[2])
    return None


def roofline(profile, n):
    """Dominant kernel (largest total time) with its algorithmic bandwidth."""
    best = None
    for name, (count, ms) in profile.items():
        if count and (best is None or ms > best[2]):
            best = (name, count, ms)
    if best is None:
        return None, None
    name, count, ms = best
    per_cell = ALGO_BYTES.get(name)
    avg_us = ms * 1e3 / count
    out = {"kernel": name, "launches": count, "avg_us": round(avg_us, 3)}
    if per_cell is None:
        return out, None
    bytes_per_launch = per_cell * n * n
    achieved = bytes_per_launch / (avg_us * 1e-6) / 1e9
    roof = {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": measured_traffic(name) if n == 4096 else None,
        "kernel": name,
        "bytes_per_launch": bytes_per_launch,
        "avg_launch_us": round(avg_us, 3),
    }
    return out, roof


def cpu_baseline(n, gpu_res, timed, tmp):
    """Oracle on the host cores: time two bounded first-step runs that differ
    only in the number of PPCG iterations, take the per-iteration cost, and
    scale it by the GPU run's outer iterations per timed step."""
    import tealeaf_cuda_amd as tl

    oracle = REPO / "oracle" / "liboracle.so"
    if not oracle.exists():
        return None
    cores = min(16, os.cpu_count() or 1)
    os.environ["OMP_NUM_THREADS"] = str(cores)
    os.environ["TEA_ORACLE_DOT"] = "plain"
    # the sample switches to PPCG right after the 25 CG pre-steps (epslim
    # lifted) so that the two runs differ only in PPCG iterations
    presteps = 25
    lo, hi = presteps + 2, presteps + 32
    times = {}
    for iters in (lo, hi):
        deck = make_deck(Path(tmp) / f"cpu_{iters}.in", n, 1, max_iters=iters, epslim=1e30)
        print(f"[bench] cpu baseline sample: {iters} iterations", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        tl.run_deck(deck, backend=oracle)
        times[iters] = time.perf_counter() - t0
    per_it = (times[hi] - times[lo]) / (hi - lo)
    its = gpu_res.iterations[timed]
    mean_its = sum(its) / len(its)
    ms_step = per_it * mean_its * 1e3
    return {
        "value": round(ms_step, 1),
        "unit": "ms/timestep",
        "cores": cores,
        "kind": "port",
        "sample": (f"oracle (C restatement of the kernels, OpenMP {cores} threads, plain reductions) through the "
                   f"same driver: first timestep at {n}^2 with tl_max_iters {lo} and {hi}; "
                   f"{per_it * 1e3:.1f} ms per PPCG iteration x {mean_its:.1f} iterations per timed GPU step"),
        "seconds": round(times[lo] + times[hi], 1),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=19)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--weights", default=None,
                    help="warm-start weight file (default: the one trained at --size)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--only", default=None,
                    help="comma-separated variants to run (profiling); the first one then gives value")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs one chunk on one GPU; multi-GPU runs are not part of this build "
              f"(asked for --gpus {args.gpus})", file=sys.stderr)
        sys.exit(2)
    n = args.size
    WEIGHTS[n] = Path(args.weights) if args.weights else weights_for(n)
    if not WEIGHTS[n].exists():
        print(f"bench.py: no warm-start weights for {n}^2 ({WEIGHTS[n]})", file=sys.stderr)
        sys.exit(2)
    chosen = [v for v in variants(n) if (v in DEFAULT_VARIANTS if args.only is None else v in args.only.split(","))]
    if not chosen:
        print(f"bench.py: --only names none of {list(variants(n))}", file=sys.stderr)
        sys.exit(2)
    main_variant = chosen[0]
    runs = {}
    with tempfile.TemporaryDirectory() as tmp:
        for variant in chosen:
            runs[variant] = run_gpu(n, args.warmup, args.steps, tmp, variant)
        res, gpu_ms, wall_ms, profile, timed, breakdown = runs[main_variant]
        ms_per_step = gpu_ms / args.steps
        dominant, roof = roofline(profile, n)
        plain_roof = roofline(runs["fp64_reference"][3], n)[1] if "fp64_reference" in runs else None
        cpu = None
        if not args.no_cpu_baseline and "fp64_reference" in runs:
            cpu = cpu_baseline(n, runs["fp64_reference"][0], timed, tmp)

    def summary(variant):
        r, g_ms, w_ms, _, t, _ = runs[variant]
        f = r.final
        return {"ms_per_step": round(g_ms / args.steps, 3), "wall_ms_per_step": round(w_ms / args.steps, 3),
                "iterations": r.iterations[t], "ppcg_inner": r.ppcg_inner[t],
                "residual_replacements": r.refinements[t], "exact_rel_residual": r.exact_error[t],
                "field_summary": {"step": f.step, "volume": f.volume, "mass": f.mass, "internal_energy": f.energy,
                                  "temperature": f.u}}

    comparison = {v: summary(v) for v in runs}
    # every variant's final field summary against plain FP64 PPCG with the same true-residual stop
    if "fp64_true" in comparison:
        ref = comparison["fp64_true"]["field_summary"]
        for v in runs:
            fs = comparison[v]["field_summary"]
            comparison[v]["field_summary_rel_err_vs_fp64_true"] = max(
                abs(fs[k] - ref[k]) / abs(ref[k]) for k in ("mass", "internal_energy", "temperature"))
    for v in [v for v in runs if v != main_variant]:
        comparison[f"speedup_vs_{v}"] = round(comparison[v]["ms_per_step"] / comparison[main_variant]["ms_per_step"],
                                              3)
    line = {
        "metric": METRIC,
        "value": round(ms_per_step, 3),
        "unit": "ms/timestep",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 fields, f64 accumulators and residual",
        "data": "synthetic",
        "config": {
            "workload": f"PPCG, jac_block, 10 inner steps, learned warm start ({WEIGHTS[n].name}), FP32 fields with FP64 residual "
                        f"replacement, {n}x{n} cells, random-rectangle layout seed {LAYOUT_SEED}, "
                        "stop at FP64 true relative residual 1e-10, dt 0.004, halo_depth 1",
            "cells": n * n,
            "solver": "ppcg",
            "preconditioner": "jac_block",
            "tl_eps": 1e-10,
            "layout_seed": LAYOUT_SEED,
        },
        "variant": main_variant,
        "wall_ms_per_step": round(wall_ms / args.steps, 3),
        "iterations": res.iterations[timed],
        "ppcg_inner": res.ppcg_inner[timed],
        "exact_rel_residual": res.exact_error[timed],
        "comparison": comparison,
        "roofline": roof,
        "roofline_fp64_reference": plain_roof,
        "dominant_kernel": dominant,
        "cpu_baseline": cpu,
        "kernels": {"steps": PROFILED_STEPS,
                    "note": f"every launch timed over {PROFILED_STEPS} steps after the timed region",
                    **{k: {"launches": c, "total_ms": round(ms, 3)} for k, (c, ms) in sorted(breakdown.items())}},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
