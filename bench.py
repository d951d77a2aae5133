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
* fp32_extrapolate: the same without the network (the residual-minimising
  combination of the previous steps' differences only);
* fp32: FP32 refinement from the u = u^n guess;
* fp64_true: plain FP64 PPCG with the same true-residual stop;
* fp64_reference: plain FP64 PPCG stopping on the reference's own recurrence
  test at tl_eps 1e-10 (its true residual is reported beside it).

The timed region is bracketed by device synchronisation; HIP events recorded
on the engine's own stream give the GPU time of the region, and the engine's
per-launch profiler (start and stop timestamps of a dispatch) times one in 8
of the inner-step passes, the dominant kernels, inside it for the roofline
figure.
Two further steps, outside the timed region, time every launch for the
per-kernel breakdown.

The CPU baseline runs the plain FP64 deck for three whole timesteps through
the reference's own Fortran driver (compiled from its sources into
oracle/_ref) over the CPU oracle (OpenMP, plain reductions) on the job's CPU
cores, and reports ms per timestep over steps 2 and 3 from the driver's own
wall clock.

`roofline` is the dominant kernel's compulsory HBM bytes per launch over its
average launch duration; `step_roofline` the compulsory bytes of every launch
of a timestep over the time per timestep.
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

# Bytes per interior cell of each launch counted as the reference's separate
# kernels would move them: every distinct FP64 array the reference kernel reads
# or writes, once, 8 bytes each (4 in FP32), and a pass of n inner steps
# charged n times.  Reported as `effective_gbs`; above the HBM peak for a
# temporally blocked pass, so it is no roofline fraction.
REF_STEP_BYTES = {
    "k_ppcg_inner_fused": 112,
    "k_ppcg_inner_fused_f32": 56,     # the same arrays in FP32
    **{f"k_ppcg_inner{n}": 112 * n for n in range(1, 6)},
    **{f"k_ppcg_inner{n}_f32": 56 * n for n in range(1, 6)},
}

# Compulsory HBM bytes per interior cell of one launch of each kernel as this
# engine runs it with the block-Jacobi preconditioner: every distinct array the
# launch must read or write, once (4 bytes per FP32 cell, 8 per FP64 cell).
# The roofline fraction of a launch is these bytes over its duration; a
# kernel absent here counts 0 bytes in the step-level figure (a lower bound).
COMPULSORY_BYTES = {
    # inner-step pass (any number of steps): sd, rtemp, utemp, Kx, Ky read; sd, rtemp, utemp written
    **{f"k_ppcg_inner{n}_f32": 32 for n in range(1, 6)},
    **{f"k_ppcg_inner{n}": 64 for n in range(1, 6)},
    # the last pass of an FP64 outer iteration also reads r and rstore for the end's r.z
    **{f"k_ppcg_inner{n}_end": 80 for n in range(1, 6)},
    # the first pass of an outer iteration with sd formed from r: r (as rtemp), Kx, Ky read;
    # sd, rtemp, utemp written
    **{f"k_ppcg_inner{n}_from_r_f32": 24 for n in range(1, 6)},
    **{f"k_ppcg_inner{n}_from_r": 48 for n in range(1, 6)},
    "k_ppcg_inner_fused_f32": 32, "k_ppcg_inner_fused": 64,
    # fused begin: p, z (utemp), Kx, Ky, r read; p', r_store written (w formed again by ur)
    "ppcg_begin_f32": 28, "ppcg_begin": 56,
    # fused ur + init_sd_new: u, p', r, Kx, Ky read; u, r written (z and sd left to the first pass,
    # which forms sd from r; after a residual replacement ur also writes z and sd, 8 or 16 more)
    "k_ppcg_ur_sd_f32": 28, "k_ppcg_ur_sd": 56,
    # fused end: utemp, r, r_store read (z owed to the next begin)
    "ppcg_end_f32": 12, "ppcg_end": 24,
    # the end with the next begin's store_r and r.z: utemp, r, r_store read; r_store written
    "ppcg_end_rr_f32": 16, "ppcg_end_rr": 32,
    # the begin after it: p, z (utemp), Kx, Ky read; p' written
    "ppcg_begin_pw_f32": 20, "ppcg_begin_pw": 40,
    # CG iteration kernels: w = Ap (p, Kx, Ky read; w written); u/r/z update (u, p, r, w, Kx, Ky read;
    # u, r, z written); p = z + beta p (z, p read; p written)
    "cg_calc_w_f32": 16, "cg_calc_w": 32,
    "k_cg_calc_ur_f32": 36, "k_cg_calc_ur": 72,
    "cg_calc_p_f32": 12, "cg_calc_p": 24, "ppcg_calc_p_f32": 12, "ppcg_calc_p": 24,
    # residual replacement in one pass: u, u0, Kx, Ky (FP64), u32, Kx32, Ky32 read;
    # u, r (FP64), r32, u0_32, z32, sd32, rtemp32, utemp32 written
    "mp_replace": 8 * 4 + 4 * 3 + 8 * 2 + 4 * 6,
    # the same with z32, sd32, rtemp32, utemp32 left to the first inner pass (formed from r32):
    # u, u0, Kx, Ky (FP64), u32 read; u, r (FP64), r32, u0_32 written
    "mp_replace_r": 8 * 4 + 4 + 8 * 2 + 4 * 2,
    "calc_residual": 40, "calc_2norm": 8, "calc_2norm_f32": 4,
}


LAYOUT_SEED = 4096


def weights_for(n):
    return REPO / "tealeaf_cuda_amd" / "weights" / f"warmstart_{n}.tlws"


WEIGHTS = {}


def variants(n):
    return {
        "headline": dict(inner_precision="fp32", warm_start=str(WEIGHTS.get(n, weights_for(n)))),
        "fp32_extrapolate": dict(inner_precision="fp32", warm_start="extrapolate"),
        # the Galerkin span over the differences of fewer previous steps (tl_warm_start_history, default 8)
        "headline_h2": dict(inner_precision="fp32", warm_start=str(WEIGHTS.get(n, weights_for(n))), warm_start_history=2),
        "fp32_extrapolate_h2": dict(inner_precision="fp32", warm_start="extrapolate", warm_start_history=2),
        "fp32_extrapolate_h4": dict(inner_precision="fp32", warm_start="extrapolate", warm_start_history=4),
        "fp32_extrapolate_h6": dict(inner_precision="fp32", warm_start="extrapolate", warm_start_history=6),
        # the coefficients that minimise the error's energy norm instead of the residual's 2-norm
        "fp32_extrapolate_energy": dict(inner_precision="fp32", warm_start="extrapolate", warm_start_norm="energy"),
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


# Launches timed one by one inside the timed region: one in TIMED_EVERY of
# the inner passes (the dominant kernels), as every timed launch costs the
# run a few microseconds.  PROFILED_STEPS further steps time every launch for the
# per-kernel breakdown.
TIMED_PREFIX = "k_ppcg_inner"
TIMED_EVERY = 8   # one in 8 of them: a timed launch leaves the GPU idle a few microseconds around it
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
            eng.profile_sample(TIMED_EVERY)
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
            eng.profile_sample(1)
            eng.profile_enable(True)
            ev["pstart"] = torch.cuda.Event(enable_timing=True)
            ev["pstart"].record(stream)
        elif step == warmup + steps + PROFILED_STEPS:
            ev["pstop"] = torch.cuda.Event(enable_timing=True)
            ev["pstop"].record(stream)
            ev["pstop"].synchronize()
            eng.profile_enable(False)
            ev["breakdown"] = eng.profile()

    res = tl.run_deck(deck, on_step=on_step)
    gpu_ms = ev["start"].elapsed_time(ev["stop"])
    wall_ms = (ev["t1"] - ev["t0"]) * 1e3
    timed = slice(warmup, warmup + steps)
    print(f"[bench] {variant}: {gpu_ms / steps:.1f} ms/step, iterations {res.iterations[timed]}, "
          f"true residual {res.exact_error[timed]}", file=sys.stderr, flush=True)
    # the profiled steps' own time (every launch timed: a little slower than untimed steps)
    profiled_ms = ev["pstart"].elapsed_time(ev["pstop"]) / PROFILED_STEPS
    return res, gpu_ms, wall_ms, ev["profile"], timed, (ev["breakdown"], profiled_ms)


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
    """Dominant kernel (largest total time): its compulsory bytes per launch
    over its average launch duration, against the HBM peak."""
    best = None
    for name, (count, ms) in profile.items():
        if count and (best is None or ms > best[2]):
            best = (name, count, ms)
    if best is None:
        return None, None
    name, count, ms = best
    avg_us = ms * 1e3 / count
    out = {"kernel": name, "launches_timed": count, "timed_every": TIMED_EVERY, "avg_us": round(avg_us, 3)}
    per_cell = COMPULSORY_BYTES.get(name)
    if per_cell is None:
        return out, None
    bytes_per_launch = per_cell * n * n
    achieved = bytes_per_launch / (avg_us * 1e-6) / 1e9
    traffic = measured_traffic(name) if n == 4096 else None
    roof = {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic,
        "kernel": name,
        "bytes_per_launch": bytes_per_launch,
        "compulsory_bytes_per_cell": per_cell,
        "avg_launch_us": round(avg_us, 3),
        # HBM bytes the counters measured per launch, over the same duration
        "traffic_gbs": round(traffic / (avg_us * 1e-6) / 1e9, 1) if traffic else None,
        # the reference's separate-launch bytes of all the pass's steps (not a roofline fraction)
        "effective_gbs": round(REF_STEP_BYTES[name] * n * n / (avg_us * 1e-6) / 1e9, 1)
        if name in REF_STEP_BYTES else None,
    }
    return out, roof


def step_roofline(profiled, steps, n):
    """Compulsory HBM bytes of every launch of the profiled timesteps over
    those steps' own GPU time: the step's achieved fraction of the HBM peak
    (a lower bound: kernels without an entry in COMPULSORY_BYTES count zero
    bytes, and timing every launch slows the steps a little)."""
    breakdown, ms_per_step = profiled
    total = 0.0
    known_ms = 0.0
    all_ms = 0.0
    for name, (count, ms) in breakdown.items():
        all_ms += ms
        if name in COMPULSORY_BYTES:
            total += COMPULSORY_BYTES[name] * n * n * count
            known_ms += ms
    per_step = total / steps
    gbs = per_step / (ms_per_step * 1e-3) / 1e9
    return {"bytes_per_step": int(per_step), "ms_per_step": round(ms_per_step, 3), "achieved": round(gbs, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4),
            "kernel_time_covered": round(known_ms / all_ms, 4) if all_ms else None}


def cpu_share():
    """CPU cores this job may use: the affinity mask, capped by a cgroup CPU
    quota and by OMP_NUM_THREADS when the environment sets one (a shared GPU
    host allots each job its share of the cores)."""
    cores = len(os.sched_getaffinity(0))
    info = {"os_cpu_count": os.cpu_count(), "affinity": cores}
    try:
        quota, period = Path("/sys/fs/cgroup/cpu.max").read_text().split()[:2]
        if quota != "max":
            info["cgroup_quota"] = int(quota) / int(period)
            cores = min(cores, max(1, int(int(quota) / int(period))))
    except (OSError, ValueError):
        pass
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        info["omp_num_threads_env"] = int(env)
        cores = min(cores, int(env))
    model = None
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    info["cpu_model"] = model
    return cores, info


def cpu_baseline(n, tmp, steps=3):
    """The plain FP64 deck (PPCG, block Jacobi, the reference's stop at tl_eps
    1e-10) through the reference's own Fortran driver compiled from its
    sources (oracle/_ref, linked to the oracle's C restatement of the kernels
    with OpenMP reductions) on the job's CPU cores, `steps` whole timesteps;
    ms per timestep from the driver's own per-step wall clock in tea.out,
    step 1 (CG pre-steps and eigenvalue estimate) excluded.  Without the
    reference build, the same deck through this repository's host driver over
    the same library, timed per step the same way."""
    import subprocess

    import tealeaf_cuda_amd as tl

    oracle = REPO / "oracle" / "liboracle.so"
    ref_exe = REPO / "oracle" / "_ref" / "tea_leaf_ref_oracle"
    if not oracle.exists():
        return None
    cores, info = cpu_share()
    env = dict(os.environ, OMP_NUM_THREADS=str(cores), OMP_PROC_BIND="close", OMP_PLACES="cores",
               TEA_ORACLE_DOT="plain")
    work = Path(tmp) / "cpu"
    work.mkdir(exist_ok=True)
    make_deck(work / "tea.in", n, steps)
    print(f"[bench] cpu baseline: {steps} timesteps at {n}^2 on {cores} threads", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    walls = []
    kind = "reference"
    if ref_exe.exists():
        proc = subprocess.run([str(ref_exe)], cwd=work, env=env, stdout=subprocess.DEVNULL,
                              stderr=subprocess.PIPE, text=True, timeout=900)
        text = (work / "tea.out").read_text() if (work / "tea.out").exists() else ""
        walls = [float(l.split()[-1].replace("D", "E")) for l in text.splitlines()
                 if l.strip().startswith("Wall clock") and len(l.split()) == 3]
        if proc.returncode != 0 or len(walls) < steps:
            print(f"[bench] reference driver failed ({proc.returncode}): {proc.stderr[-500:]}", file=sys.stderr)
            walls = []
    if not walls:
        kind = "port"
        saved = {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES", "TEA_ORACLE_DOT")}
        os.environ.update({k: env[k] for k in saved})
        stamps = []
        tl.run_deck(work / "tea.in", backend=oracle, on_step=lambda s: stamps.append(time.perf_counter()))
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
        walls = [t - t0 for t in stamps]
    elapsed = time.perf_counter() - t0
    per_step = [(walls[k] - walls[k - 1]) * 1e3 for k in range(1, steps)]
    ms_step = sum(per_step) / len(per_step)
    what = ("the reference Fortran driver (oracle/_ref/tea_leaf_ref_oracle)" if kind == "reference" else
            "this repository's host driver")
    return {
        "value": round(ms_step, 1),
        "unit": "ms/timestep",
        "cores": cores,
        "kind": kind,
        "sample": (f"plain FP64 PPCG (block Jacobi, 10 inner steps, reference stop at tl_eps 1e-10), {n}^2 "
                   f"random-rectangle layout seed {LAYOUT_SEED}: {what} over the oracle library (C restatement "
                   f"of the kernels, OpenMP {cores} threads, OMP_PROC_BIND=close, OMP_PLACES=cores), {steps} whole "
                   f"timesteps, ms per timestep over steps 2..{steps} from the driver's per-step wall clock"),
        "per_step_ms": [round(x, 1) for x in per_step],
        "seconds": round(elapsed, 1),
        "host": info,
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
        if not args.no_cpu_baseline:
            cpu = cpu_baseline(n, tmp)

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
        "step_roofline": step_roofline(breakdown, PROFILED_STEPS, n),
        "step_roofline_fp64_reference": (step_roofline(runs["fp64_reference"][5], PROFILED_STEPS, n)
                                         if "fp64_reference" in runs else None),
        "dominant_kernel": dominant,
        "cpu_baseline": cpu,
        "kernels": {"steps": PROFILED_STEPS,
                    "note": f"every launch timed over {PROFILED_STEPS} steps after the timed region",
                    "ms_per_step": round(breakdown[1], 3),
                    **{k: {"launches": c, "total_ms": round(ms, 3)} for k, (c, ms) in sorted(breakdown[0].items())}},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
