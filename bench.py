#!/usr/bin/env python3
"""Benchmark: HydroC99 split Godunov/MUSCL step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--nx 8192] [--no-cpu]

Workload: the 8192x8192 centred blast of the FP64 plain form (E = 1/dx^2 in the
centre cell, rho = 1, E = 1e-5 elsewhere, dx = 1/8192, reflective walls, CFL
0.8, MUSCL, slope_type 1, 10 Riemann iterations).  A step is one full time
step of the reference loop (dt on even steps, x+y sweeps), run on the GPU
with the state resident in HBM; W untimed steps, then K timed steps.

Prints ONE JSON line.  `value` is interior cell-updates per second (the
reference's MC/s unit x 1e6).  `roofline` is for the dominant sweep kernel:
algorithmic bytes (64 B per cell per sweep in FP64: read 4 doubles, write 4)
over its average duration, timed with HIP events recorded on the grid's own
stream around every launch in the timed region.  `cpu_baseline` is the
reference's OpenMP CPU path (HydroC99_2Domp, built by oracle/Makefile into
oracle/_ref/hydro_omp) on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-updates/s + % HBM roofline; batched solves/s & p99 latency vs plain FP64"
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
BYTES_PER_CELL_SWEEP = {0: 64, 1: 32}  # FP64 / FP32 storage: 4 vars read + 4 written


def gpu_bench(nx, ny, steps, warmup, precision, autotune=True):
    import hydro_amd
    p = hydro_amd.Params(nx=nx, ny=ny, dx=1.0 / nx, courant_factor=0.8, niter_riemann=10,
                         iorder=2, slope_type=1.0, tend=1e30, nstepmax=2**31 - 1,
                         precision=precision)
    g = hydro_amd.Grid(p)
    g.init(0)
    # launch shapes picked on this grid before timing (the state is untouched)
    tune = g.autotune(reps=2) if autotune else None
    if warmup:
        g.run(warmup, log=False)
    g.set_profile(True)
    g.sync()
    t0 = time.perf_counter()
    done = g.run(steps, log=False)
    t1 = time.perf_counter()
    prof = g.profile()
    g.set_profile(False)
    tstate = g.state()
    g.close()
    assert done == steps, (done, steps)
    elapsed = t1 - t0
    cells = nx * ny
    kern = max(("x_sweep", "y_sweep"), key=lambda k: prof[k][0])
    ms, n = prof[kern]
    per_launch_bytes = BYTES_PER_CELL_SWEEP[precision] * cells
    avg_s = ms / n / 1e3
    achieved = per_launch_bytes / avg_s / 1e9
    return {
        "tune": tune,
        "value": cells * steps / elapsed,
        "ms_per_step": elapsed / steps * 1e3,
        "elapsed_s": elapsed,
        "t_end": tstate[0],
        "kernels": {k: {"avg_ms": (v[0] / v[1] if v[1] else None), "launches": v[1]}
                    for k, v in prof.items()},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None, "kernel": kern,
                     "algorithmic_bytes_per_launch": per_launch_bytes},
    }


def cpu_threads():
    n = os.cpu_count() or 1
    # one GPU's share of the box's host cores
    return max(1, min(n, int(os.environ.get("HYDRO_CPU_THREADS", "16"))))


def cpu_baseline(nx, ny, nsteps=8):
    """Time the reference OpenMP CPU path on `nsteps` steps of the same grid.

    MC/s per the reference (HydroC99_2Domp/Src/main.c:426-427, averaged over
    steps > 5); reported in cell-updates/s."""
    ref = os.path.join(ROOT, "oracle", "_ref", "hydro_omp")
    threads = cpu_threads()
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="spread",
               OMP_PLACES="cores")
    deck = ("&RUN\ntend=1e30\nnstepmax=%d\n/\n&MESH\nnx=%d\nny=%d\nnxystep=1024\ndx=%r\n"
            "boundary_left=1\nboundary_right=1\nboundary_down=1\nboundary_up=1\ntestcase=0\n/\n"
            "&HYDRO\ncourant_factor=0.8\nniter_riemann=10\niorder=2\nslope_type=1.\n"
            "scheme=muscl\nsmallr=1e-10\nsmallc=1e-10\n/\n" % (nsteps, nx, ny, 1.0 / nx))
    sample = f"{nx}x{ny} centred blast, {nsteps} steps, MC/s averaged over steps > 5"
    if os.path.exists(ref):
        with tempfile.TemporaryDirectory() as d:
            path = os.path.join(d, "bench.nml")
            with open(path, "w") as f:
                f.write(deck)
            out = subprocess.run([ref, "-i", path], cwd=d, env=env, capture_output=True,
                                 text=True, timeout=600).stdout
        per_step = [float(m) for m in re.findall(r"--> step=\s*\d+,.*?([0-9.]+) MC/s", out)]
        avg = re.search(r"Average MC/s: ([0-9.]+)", out)
        mcs = float(avg.group(1)) if avg else (sum(per_step[5:]) / max(1, len(per_step[5:])))
        return {"value": mcs * 1e6, "unit": "cell-updates/s", "cores": threads,
                "kind": "reference", "sample": sample + " (HydroC99_2Domp, gcc -O3 -fopenmp)"}
    import numpy as np
    import oracle
    p = oracle.Params(nx=nx, ny=ny, dx=1.0 / nx, courant_factor=0.8, tend=1e30, testcase=0)
    port = oracle.Port()
    u = port.init(p)
    port.run(p, u, 2)
    t0 = time.perf_counter()
    port.run(p, u, 2, state=None)
    dt = time.perf_counter() - t0
    del np
    return {"value": nx * ny * 2 / dt, "unit": "cell-updates/s", "cores": threads,
            "kind": "port", "sample": f"{nx}x{ny} centred blast, 2 steps (oracle restatement)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--nx", type=int, default=8192)
    ap.add_argument("--ny", type=int, default=0)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=8)
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {a.gpus}); the "
              "multi-GPU row-band decomposition is a design, not built", file=sys.stderr)
        sys.exit(2)
    ny = a.ny or a.nx
    prec = 0 if a.precision == "fp64" else 1
    r = gpu_bench(a.nx, ny, a.steps, a.warmup, prec, autotune=not a.no_autotune)
    line = {
        "metric": METRIC,
        "value": r["value"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == 0 else "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"{a.nx}x{ny} centred blast, split Godunov/MUSCL, "
                               f"{'FP64' if prec == 0 else 'mixed FP32/FP64'}, "
                               "CFL 0.8, reflective walls",
                   "nx": a.nx, "ny": ny, "testcase": 0, "steps_timed": a.steps,
                   "first_timed_step": a.warmup, "launch_shape": r["tune"]},
        "roofline": r["roofline"],
        "kernels": r["kernels"],
    }
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_baseline(a.nx, ny, a.cpu_steps)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
