#!/usr/bin/env python3
"""Benchmark: HydroC99 split Godunov/MUSCL step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--nx 8192]
                    [--precision fast|fp64|mixed] [--no-cpu]

Workload: the 8192x8192 centred blast (E = 1/dx^2 in the centre cell, rho = 1,
E = 1e-5 elsewhere, dx = 1/8192, reflective walls, CFL 0.8, MUSCL,
slope_type 1, 10 Riemann iterations).  A step is one full time step of the
reference loop (dt on even steps, x+y sweeps), run on the GPU with the state
resident in HBM; W untimed steps, then K timed steps.  The headline form is
the FP64 fast form (FP64 storage and arithmetic, within a per-field relative
L1 of 1e-12 of the reference); the bit-for-bit FP64 form and the mixed form
are measured beside it on the same grid.

Prints ONE JSON line.  `value` is interior cell-updates per second (the
reference's MC/s unit x 1e6) over the timed steps, run as the product runs
them: replayed HIP graphs, no per-launch instrumentation.  Per-kernel
durations come from a second, instrumented pass over the same grid (HIP
events recorded on the grid's own stream around every launch).  `roofline` is
for the dominant sweep kernel: algorithmic bytes (64 B per cell per sweep in
FP64: read 4 doubles, write 4) over its average duration; its `traffic` is
the HBM bytes the kernel really moves (two rocprofv3 --pmc passes,
calibrated on a copy kernel), and its `copy_ceiling` the rate of a plain copy
in the same access shape (the sweep's bytes with no arithmetic: what this
pattern reaches on the machine, beside the published peak).  `valu` is that
kernel's VALU work from a third counter pass.  `c2` is the 23000x23000 corner blast (50 steps).
`batched` and `serving` time the many-grid workloads (512 batched instances;
256 Poisson arrivals at a nominal and a saturating rate, mixed precision)
against the plain per-grid FP64 loop, timed cold (each grid created, run and
freed in turn) and warm (grids created before the clock starts).
`cpu_baseline` is the reference's OpenMP CPU path (HydroC99_2Domp, built by
oracle/Makefile into oracle/_ref/hydro_omp) on a bounded sample of the same
workload.
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
BYTES_PER_CELL_SWEEP = {0: 64, 1: 32, 2: 64}  # FP64 / FP32 / FP64: 4 vars read + 4 written


PRECISIONS = {"fp64": 0, "mixed": 1, "fast": 2}
FORM_NAMES = {0: "FP64 (bit for bit)", 1: "mixed FP32/FP64", 2: "FP64 fast form (1e-12)"}
PREC_ARG = {v: k for k, v in PRECISIONS.items()}


_T0 = time.perf_counter()


def progress(what):
    """One progress line on stderr (the JSON result line goes to stdout)."""
    print(f"bench.py [{time.perf_counter() - _T0:7.1f} s] {what}", file=sys.stderr, flush=True)


def gpu_bench(nx, ny, steps, warmup, precision, autotune=True, testcase=0, prof_steps=20):
    """Times `steps` steps of one grid on the product path (graph replay), then
    an instrumented pass of prof_steps steps for per-kernel durations."""
    import hydro_amd
    p = hydro_amd.Params(nx=nx, ny=ny, dx=1.0 / nx, courant_factor=0.8, niter_riemann=10,
                         iorder=2, slope_type=1.0, tend=1e30, nstepmax=2**31 - 1,
                         precision=precision)
    g = hydro_amd.Grid(p)
    g.init(testcase)
    # launch shapes picked on this grid before timing (the state is untouched)
    tune = g.autotune(reps=5) if autotune else None
    if warmup:
        g.run(warmup, log=False)
    g.sync()
    t0 = time.perf_counter()
    done = g.run(steps, log=False)
    t1 = time.perf_counter()
    assert done == steps, (done, steps)
    tstate = g.state()
    g.set_profile(True)
    g.run(prof_steps, log=False)
    prof = g.profile()
    g.set_profile(False)
    g.close()
    elapsed = t1 - t0
    cells = nx * ny
    progress(f"{nx}x{ny} {FORM_NAMES[precision]}: {cells * steps / elapsed / 1e9:.2f} G "
             "cell-updates/s")
    kern = max(("x_sweep", "y_sweep"), key=lambda k: prof[k][0])
    ms, n = prof[kern]
    bpc = BYTES_PER_CELL_SWEEP[precision]
    per_launch_bytes = bpc * cells
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
        "kernels_note": f"instrumented pass of {prof_steps} steps after the timed region",
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None, "kernel": kern,
                     "algorithmic_bytes_per_launch": per_launch_bytes,
                     "bytes_per_cell_sweep": bpc},
        "step_roofline_frac": round(cells * steps * 2 * bpc / elapsed / 1e9 / HBM_PEAK_GBS, 4),
    }


def form_summary(r, precision):
    return {"form": FORM_NAMES[precision], "value": r["value"], "ms_per_step": r["ms_per_step"],
            "roofline_frac": r["roofline"]["frac"], "step_roofline_frac": r["step_roofline_frac"],
            "kernels": r["kernels"], "launch_shape": r["tune"]}


def batched_bench(n=512, steps=100):
    """Batched workload: n seeded blast instances (sides log-uniform in [64, 1024]),
    `steps` steps each.  The plain per-grid FP64 loop runs one Grid per
    instance in turn, timed cold (each grid created, initialised, run and
    freed inside the clock) and warm (every grid created and initialised
    before the clock starts; only the runs are timed); the batches (FP64,
    FP64 fast form, mixed) are timed from their first step, with admission,
    initialisation and the autotuner (3 timed repetitions per shape) before
    the clock, and that setup reported beside."""
    import hydro_amd
    from hydro_amd import workloads as W
    out = {"instances": n, "steps": steps,
           "cell_updates": W.cell_updates(W.batch_instances(n, steps=steps))}
    runs = W.batch_instances(n, steps=steps)
    t0 = time.perf_counter()
    for r in runs:
        with hydro_amd.Grid(r.params) as g:
            g.init_blast(r.bx, r.by, r.e)
            g.run(steps, log=False)
    t_cold = time.perf_counter() - t0
    grids = [hydro_amd.Grid(r.params) for r in runs]
    for g, r in zip(grids, runs):
        g.init_blast(r.bx, r.by, r.e)
    for g in grids:
        g.sync()
    t0 = time.perf_counter()
    for g in grids:
        g.run(steps, log=False)
    t_warm = time.perf_counter() - t0
    for g in grids:
        g.close()
    del grids
    out["plain_fp64_per_grid"] = {"seconds": t_warm, "solves_per_s": n / t_warm,
                                  "cold_seconds": t_cold, "cold_solves_per_s": n / t_cold}
    progress(f"batched: plain per-grid FP64 loop {t_warm:.2f} s warm, {t_cold:.2f} s cold")
    for name, prec in (("batched_fp64", hydro_amd.FP64), ("batched_fp64_fast", hydro_amd.FP64_FAST),
                       ("batched_mixed", hydro_amd.MIXED)):
        runs = W.batch_instances(n, steps=steps, precision=prec)
        ts = time.perf_counter()
        with hydro_amd.Batch(prec) as b:
            ids = [b.add(r.params) for r in runs]
            for i, r in zip(ids, runs):
                b.init_blast(i, r.bx, r.by, r.e)
            tune = b.autotune(reps=3)
            b.sync()
            t0 = time.perf_counter()
            b.run(steps)
            t = time.perf_counter() - t0
            q = b.query()
            assert all(v[1] == steps for v in q.values())
        out[name] = {"seconds": t, "solves_per_s": n / t, "cell_updates_per_s":
                     out["cell_updates"] / t, "tune": tune, "setup_seconds": t0 - ts,
                     "speedup_vs_plain_fp64": t_warm / t,
                     "speedup_vs_plain_fp64_cold": t_cold / t}
        progress(f"batched: {name} {t:.3f} s")
    return out


def _serve(runs, chunk):
    """Feed `runs` to a mixed-precision Server at their arrival offsets (real
    time) and pump it until every run is done; returns (server, request ids,
    per-request info, span in seconds from the first arrival)."""
    import hydro_amd
    srv = hydro_amd.Server(precision=hydro_amd.MIXED, budget_bytes=200 * 10**9,
                           chunk_steps=chunk, keep_results=True)
    # pinned room for every result (FP32 state incl. ghosts, 4 variables)
    srv.reserve_results(sum(16 * (r.params.nx + 4) * (r.params.ny + 4) + 256 for r in runs))
    t_start = hydro_amd.now()
    ids, k, n = [], 0, len(runs)
    while True:
        now = hydro_amd.now() - t_start
        while k < n and runs[k].arrival <= now:
            ids.append(srv.submit(runs[k].params, runs[k].bx, runs[k].by, runs[k].e))
            k += 1
        left = srv.pump()  # admit, advance everything running by one chunk, retire
        if left == 0:
            if k == n:
                break
            time.sleep(max(0.0, runs[k].arrival - (hydro_amd.now() - t_start)))
    info = [srv.info(i) for i in ids]
    assert all(x[0] == srv.DONE for x in info), [x[0] for x in info]
    return srv, ids, info, max(x[3] for x in info) - t_start


def serving_bench(n=256, rates=(20.0, 1000.0), chunk=20):
    """Serving workload: n runs (sides log-uniform in [128, 4096], t_end 0.1 or 500
    steps) arriving as a Poisson process; mixed precision served by the
    continuous-batching loop against the plain per-grid FP64 loop (FIFO, one
    run at a time, each run's service time measured) on the same requests.
    The plain loop's service time is measured warm (grid created and
    initialised before its clock starts: run + download) and cold (creation
    and initialisation included).  Latency = finished - arrived (results on
    the host in both).  The first rate is the nominal one; the last is far
    above either's capacity, so its solves/s is each side's throughput under a
    full queue.  The mixed results' error: relative L1 and L-infinity of all
    four fields against the FP64 run of the same request."""
    import numpy as np
    import hydro_amd
    from hydro_amd import workloads as W
    res = {"requests": n, "chunk_steps": chunk, "budget_bytes": 200 * 10**9, "rates": []}
    served = []
    for i, rate in enumerate(rates):
        runs = W.serving_requests(n, rate=rate, precision=hydro_amd.MIXED)
        srv, ids, info, span = _serve(runs, chunk)
        lat = np.array([x[3] - x[1] for x in info])
        served.append({"rate_per_s": rate, "runs": runs,
                       "served_mixed": {"solves_per_s": n / span, "seconds": span,
                                        "p50_s": float(np.percentile(lat, 50)),
                                        "p99_s": float(np.percentile(lat, 99)),
                                        "max_s": float(lat.max()),
                                        "fp64_reruns": srv.reruns}})
        progress(f"serving: {n} requests at {rate:g}/s in {span:.2f} s")
        if i == 0:
            keep_srv, keep_ids = srv, ids
            res["cell_updates"] = float(sum(r.params.nx * r.params.ny * x[5]
                                            for r, x in zip(runs, info)))
        else:
            srv.close()
    # the plain per-grid FP64 loop: service time of each run alone, and the
    # mixed results' error against it
    warm, cold, errs = [], [], []
    for rid, r in zip(keep_ids, served[0]["runs"]):
        p64 = hydro_amd.Params(**{**r.params.__dict__, "precision": hydro_amd.FP64})
        tc = time.perf_counter()
        with hydro_amd.Grid(p64) as g:
            g.init_blast(r.bx, r.by, r.e)
            g.sync()
            t0 = time.perf_counter()
            g.run(10**6, log=False)
            u64 = g.download()
            t1 = time.perf_counter()
        warm.append(t1 - t0)
        cold.append(t1 - tc)
        um = keep_srv.result(rid)
        keep_srv.release(rid)
        e = []
        for v in range(4):
            a, b = um[v, 2:-2, 2:-2], u64[v, 2:-2, 2:-2]
            d = np.abs(a - b)
            e.append((float(d.sum() / np.abs(b).sum()), float(d.max() / np.abs(b).max())))
        errs.append(e)
        del um, u64
        if len(warm) % 64 == 0:
            progress(f"serving: plain FP64 loop, {len(warm)} of {n} runs")
    keep_srv.close()
    for s in served:
        runs = s.pop("runs")
        for key, service in (("plain_fp64_per_grid", warm), ("plain_fp64_per_grid_cold", cold)):
            done, lat_plain = 0.0, []
            for r, sv in zip(runs, service):  # FIFO queue of one server
                done = max(done, r.arrival) + sv
                lat_plain.append(done - r.arrival)
            s[key] = {"solves_per_s": n / done, "seconds": done,
                      "p50_s": float(np.percentile(lat_plain, 50)),
                      "p99_s": float(np.percentile(lat_plain, 99))}
        s["speedup_solves_per_s"] = s["served_mixed"]["solves_per_s"] / s[
            "plain_fp64_per_grid"]["solves_per_s"]
        s["speedup_solves_per_s_vs_cold"] = s["served_mixed"]["solves_per_s"] / s[
            "plain_fp64_per_grid_cold"]["solves_per_s"]
        res["rates"].append(s)
    errs = np.array(errs)  # [run, field, (L1, Linf)]
    names = ("density", "momentum_x", "momentum_y", "energy")
    res["plain_fp64_service_s_total"] = {"warm": float(sum(warm)), "cold": float(sum(cold))}
    res["mixed_error_vs_fp64"] = {
        **{f"{nm}_l1_max": float(errs[:, v, 0].max()) for v, nm in enumerate(names)},
        **{f"{nm}_linf_max": float(errs[:, v, 1].max()) for v, nm in enumerate(names)},
        "note": "relative L1 (sum|a-b|/sum|b|) and L-infinity (max|a-b|/max|b|) of the interior "
                "per field against the FP64 run of the same request, worst over the requests"}
    return res


def cpu_threads():
    """Threads for the CPU baseline: one GPU's share of the box's host cores
    (16 on these boxes, which allot that share to a one-GPU job: OpenMP pools
    stay within it), unless HYDRO_CPU_THREADS says otherwise."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(n, int(os.environ.get("HYDRO_CPU_THREADS", "16"))))


def host_facts():
    """The host's CPU as lscpu reports it (model, sockets, physical cores,
    threads per core), the CPUs this process may run on, and the cgroup CPU
    quota, if any."""
    out = {}
    try:
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        kv = dict(l.split(":", 1) for l in txt.splitlines() if ":" in l)
        kv = {k.strip(): v.strip() for k, v in kv.items()}
        sockets = int(kv.get("Socket(s)", "0") or 0)
        per = int(kv.get("Core(s) per socket", "0") or 0)
        out = {"model": kv.get("Model name"), "sockets": sockets,
               "physical_cores": sockets * per if sockets and per else None,
               "threads_per_core": int(kv.get("Thread(s) per core", "0") or 0) or None,
               "logical_cpus": int(kv.get("CPU(s)", "0") or 0) or None}
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    try:
        out["affinity_cpus"] = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        pass
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            out["cgroup_cpu_max"] = f.read().strip()
    except OSError:
        pass
    return out


def cpu_baseline(nx, ny, nsteps=8, testcase=0):
    """Time the reference OpenMP CPU path on `nsteps` steps of the same grid.

    MC/s per the reference (HydroC99_2Domp/Src/main.c:426-427, averaged over
    steps > 5; with fewer steps, over every step after the first, which pays
    the first touch of the arrays); reported in cell-updates/s."""
    ref = os.path.join(ROOT, "oracle", "_ref", "hydro_omp")
    threads = cpu_threads()
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="spread",
               OMP_PLACES="cores")
    deck = ("&RUN\ntend=1e30\nnstepmax=%d\n/\n&MESH\nnx=%d\nny=%d\nnxystep=1024\ndx=%r\n"
            "boundary_left=1\nboundary_right=1\nboundary_down=1\nboundary_up=1\ntestcase=%d\n/\n"
            "&HYDRO\ncourant_factor=0.8\nniter_riemann=10\niorder=2\nslope_type=1.\n"
            "scheme=muscl\nsmallr=1e-10\nsmallc=1e-10\n/\n" % (nsteps, nx, ny, 1.0 / nx, testcase))
    blast = {0: "centred blast", 1: "corner blast"}.get(testcase, f"testcase {testcase}")
    over = "steps > 5" if nsteps > 5 else "steps 2.." + str(nsteps)
    sample = f"{nx}x{ny} {blast}, {nsteps} steps, MC/s averaged over {over}"
    if os.path.exists(ref):
        with tempfile.TemporaryDirectory() as d:
            path = os.path.join(d, "bench.nml")
            with open(path, "w") as f:
                f.write(deck)
            out = subprocess.run([ref, "-i", path], cwd=d, env=env, capture_output=True,
                                 text=True, timeout=600).stdout
        per_step = [float(m) for m in re.findall(r"--> step=\s*\d+,.*?([0-9.]+) MC/s", out)]
        avg = re.search(r"Average MC/s: ([0-9.]+)", out)
        mcs = float(avg.group(1)) if avg else 0.0
        if not mcs > 0:
            tail = per_step[1:] or per_step
            mcs = sum(tail) / max(1, len(tail))
        return {"value": mcs * 1e6, "unit": "cell-updates/s", "cores": threads,
                "kind": "reference", "sample": sample + " (HydroC99_2Domp, gcc -O3 -fopenmp)",
                "host": host_facts(),
                "cores_note": "OpenMP threads = this job's share of the host (one GPU's share, "
                              "the OMP_NUM_THREADS the box sets); the host's whole CPU is in "
                              "`host`"}
    import oracle
    p = oracle.Params(nx=nx, ny=ny, dx=1.0 / nx, courant_factor=0.8, tend=1e30,
                      testcase=testcase)
    port = oracle.Port()
    u = port.init(p)
    t0 = time.perf_counter()
    port.run(p, u, 2)
    dt = time.perf_counter() - t0
    return {"value": nx * ny * 2 / dt, "unit": "cell-updates/s", "cores": threads,
            "kind": "port", "sample": f"{nx}x{ny} {blast}, 2 steps (oracle restatement)"}


def pmc_workload(nx, precision, steps, tune=None):
    """A short fixed workload for counter runs (rocprofv3 --pmc): one plain
    copy kernel of 1 GiB (the byte-counter calibration), then `steps` steps of
    the bench grid with the given launch shape (default shape if None)."""
    import hydro_amd
    gbs = hydro_amd.copy_probe(1 << 30, 1)
    p = hydro_amd.Params(nx=nx, ny=nx, dx=1.0 / nx, courant_factor=0.8, tend=1e30,
                         precision=precision)
    with hydro_amd.Grid(p) as g:
        g.init(0)
        if tune:
            g.set_tune(*tune)
        g.run(steps, log=False)
    print(json.dumps({"pmc_workload": True, "copy_bytes": 1 << 30, "copy_GBps": gbs,
                      "nx": nx, "steps": steps}), flush=True)


def pmc_batch_workload(precision, n, steps):
    """A short fixed workload for counter and trace runs of the batched
    kernels: the 1 GiB copy kernel, then `steps` steps of the n-instance
    batch (sides log-uniform in [64, 1024]) with the batch's default task
    shape."""
    import hydro_amd
    from hydro_amd import workloads as W
    gbs = hydro_amd.copy_probe(1 << 30, 1)
    runs = W.batch_instances(n, steps=steps, precision=precision)
    with hydro_amd.Batch(precision) as b:
        ids = [b.add(r.params) for r in runs]
        for i, r in zip(ids, runs):
            b.init_blast(i, r.bx, r.by, r.e)
        b.run(steps)
    print(json.dumps({"pmc_batch_workload": True, "copy_bytes": 1 << 30, "copy_GBps": gbs,
                      "instances": n, "steps": steps,
                      "cells": sum(r.params.nx * r.params.ny for r in runs)}), flush=True)


COPY_PROBE_BYTES = 1 << 30  # each of read and written by pmc_workload's copy kernel


def _pmc(counters, args):
    """{counter: {kernel: mean value per dispatch}} from one rocprofv3 --pmc
    run of this script with `args` (kernels: copy, x_sweep, y_sweep), or None
    if the profiler is absent or fails."""
    import csv
    import shutil
    exe = shutil.which("rocprofv3")
    if not exe:
        return None
    with tempfile.TemporaryDirectory(dir="/tmp") as d:
        cmd = [exe, "--pmc", *counters, "--output-format", "csv", "-d", d, "-o", "run", "--",
               sys.executable, os.path.abspath(__file__), "--pmc-workload"] + args
        try:
            subprocess.run(cmd, cwd=d, env=dict(os.environ, TMPDIR="/tmp"), timeout=150,
                           capture_output=True, check=True)
        except (subprocess.SubprocessError, OSError):
            return None
        path = os.path.join(d, "run_counter_collection.csv")
        if not os.path.exists(path):
            return None
        acc = {}
        with open(path) as f:
            for r in csv.DictReader(f):
                name = r["Kernel_Name"]
                key = ("copy" if "k_copy_probe" in name else "x_sweep" if "k_sweep_x" in name
                       else "y_sweep" if "k_sweep_y" in name else None)
                if key and r.get("Counter_Name") in counters:
                    acc.setdefault(r["Counter_Name"], {}).setdefault(key, []).append(
                        float(r["Counter_Value"]))
        return {c: {k: sum(v) / len(v) for k, v in m.items()} for c, m in acc.items()}


def _pmc_kib(counter, args):
    r = _pmc([counter], args)
    return r.get(counter) if r else None


def measured_traffic(nx, precision, tune):
    """HBM bytes per launch of each sweep kernel: rocprofv3 FETCH_SIZE and
    WRITE_SIZE passes over pmc_workload, each scaled by what the same counter
    reads for the copy kernel (COPY_PROBE_BYTES read and written)."""
    args = ["--steps", "2", "--nx", str(nx), "--precision", PREC_ARG[precision]]
    if tune:
        args += ["--tune", ",".join(str(tune[k]) for k in ("yseg", "ywaves", "xwaves", "xiter"))]
    fetch, write = _pmc_kib("FETCH_SIZE", args), _pmc_kib("WRITE_SIZE", args)
    if not fetch or not write or "copy" not in fetch or "copy" not in write:
        return None
    out = {"fetch_scale": COPY_PROBE_BYTES / (fetch["copy"] * 1024),
           "write_scale": COPY_PROBE_BYTES / (write["copy"] * 1024)}
    for k in ("x_sweep", "y_sweep"):
        if k in fetch and k in write:
            out[k] = (fetch[k] * 1024 * out["fetch_scale"] + write[k] * 1024 * out["write_scale"])
    return out


VALU_COUNTERS = ["SQ_INSTS_VALU", "SQ_INSTS_VALU_FMA_F64", "SQ_INSTS_VALU_MUL_F64",
                 "SQ_INSTS_VALU_ADD_F64", "SQ_INSTS_VALU_TRANS_F64", "GRBM_GUI_ACTIVE"]
FP64_VALU_PEAK = 78.6e12  # MI355X FP64 vector FLOP/s
SIMDS = 1024              # 256 CUs x 4 SIMDs
XCDS = 8                  # GRBM_GUI_ACTIVE is summed over the XCDs


def measured_valu(nx, precision, tune, kern, avg_s):
    """VALU work of the sweep kernel `kern` from one rocprofv3 --pmc pass over
    pmc_workload: VALU instructions per cell, FP64 FLOP/s (FMA = 2) over the
    kernel's measured duration `avg_s`, and the share of SIMD cycles issuing
    VALU, counting each wave64 instruction at 4 cycles (a 16-lane SIMD; the
    FP32-heavy mixed form reads above 1, so some instructions take fewer)."""
    args = ["--steps", "2", "--nx", str(nx), "--precision", PREC_ARG[precision]]
    if tune:
        args += ["--tune", ",".join(str(tune[k]) for k in ("yseg", "ywaves", "xwaves", "xiter"))]
    r = _pmc(VALU_COUNTERS, args)
    if not r or any(kern not in r.get(c, {}) for c in VALU_COUNTERS):
        return None
    v = {c: r[c][kern] for c in VALU_COUNTERS}
    cells = nx * nx
    flops = 64 * (2 * v["SQ_INSTS_VALU_FMA_F64"] + v["SQ_INSTS_VALU_MUL_F64"]
                  + v["SQ_INSTS_VALU_ADD_F64"] + v["SQ_INSTS_VALU_TRANS_F64"])
    cycles = v["GRBM_GUI_ACTIVE"] / XCDS
    return {"kernel": kern,
            "valu_instr_per_cell": round(64 * v["SQ_INSTS_VALU"] / cells, 1),
            "fp64_instr_per_cell": {c[len("SQ_INSTS_VALU_"):].lower(): round(64 * v[c] / cells, 1)
                                    for c in VALU_COUNTERS[1:5]},
            "fp64_flops_per_cell": round(flops / cells, 1),
            "fp64_tflops": round(flops / avg_s / 1e12, 2),
            "fp64_peak_tflops": FP64_VALU_PEAK / 1e12,
            "fp64_frac": round(flops / avg_s / FP64_VALU_PEAK, 4),
            "valu_issue_busy": round(v["SQ_INSTS_VALU"] * 4 / (SIMDS * cycles), 3),
            "valu_issue_note": "SQ_INSTS_VALU x 4 cycles / (1024 SIMDs x GPU cycles): every "
                               "wave64 VALU instruction counted at 4 cycles; above 1 where some "
                               "instructions issue in fewer"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--nx", type=int, default=8192)
    ap.add_argument("--ny", type=int, default=0)
    ap.add_argument("--precision", choices=sorted(PRECISIONS), default="fast")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=16)
    ap.add_argument("--testcase", type=int, default=0)
    ap.add_argument("--no-forms", action="store_true",
                    help="skip the other arithmetic forms on the same grid")
    ap.add_argument("--no-c2", action="store_true", help="skip the 23000x23000 corner blast")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--batch-n", type=int, default=512)
    ap.add_argument("--serve-n", type=int, default=256)
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--pmc-workload", action="store_true",
                    help="short fixed run for rocprofv3 counter passes")
    ap.add_argument("--pmc-batch-workload", action="store_true",
                    help="short fixed batched run (--batch-n instances, --steps steps) for "
                         "rocprofv3 counter and trace passes")
    ap.add_argument("--tune", default="", help="yseg,ywaves,xwaves,xiter (pmc workload)")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the rocprofv3 counter passes behind roofline.traffic and valu")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: runs on one GPU only (asked for --gpus {a.gpus}); the "
              "multi-GPU row-band decomposition is a design, not built", file=sys.stderr)
        sys.exit(2)
    ny = a.ny or a.nx
    prec = PRECISIONS[a.precision]
    if a.pmc_workload:
        pmc_workload(a.nx, prec, a.steps,
                     [int(v) for v in a.tune.split(",")] if a.tune else None)
        return
    if a.pmc_batch_workload:
        pmc_batch_workload(prec, a.batch_n, a.steps)
        return
    r = gpu_bench(a.nx, ny, a.steps, a.warmup, prec, autotune=not a.no_autotune,
                  testcase=a.testcase)
    blast = "centred" if a.testcase == 0 else "corner"
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
        "dtype": "f32+f64" if prec == 1 else "f64",
        "data": "synthetic",
        "config": {"workload": f"{a.nx}x{ny} {blast} blast, split Godunov/MUSCL, "
                               f"{FORM_NAMES[prec]}, CFL 0.8, reflective walls",
                   "form": FORM_NAMES[prec], "nx": a.nx, "ny": ny, "testcase": a.testcase,
                   "steps_timed": a.steps, "first_timed_step": a.warmup,
                   "launch_shape": r["tune"], "timed_path": "replayed HIP graphs"},
        "roofline": r["roofline"],
        "step_roofline_frac": r["step_roofline_frac"],
        "kernels": r["kernels"],
        "kernels_note": r["kernels_note"],
    }
    # the bandwidth this access shape reaches with no arithmetic, beside the
    # published peak
    import hydro_amd
    ceil = hydro_amd.stream_probe((a.nx + 4) * (ny + 4), 10)
    line["roofline"]["copy_ceiling"] = {
        "GBps": round(ceil, 1), "frac": round(line["roofline"]["achieved"] / ceil, 4),
        "kernel": "k_stream_probe: a copy in the sweep's access shape (4 arrays of "
                  f"{(a.nx + 4) * (ny + 4)} doubles in, 4 out, one cell per thread, "
                  "non-temporal loads and stores), 64 bytes per cell"}
    progress(f"streaming-copy ceiling {ceil:.0f} GB/s")
    if not a.no_forms:
        forms = {a.precision: form_summary(r, prec)}
        for name, p2 in sorted(PRECISIONS.items()):
            if p2 != prec:
                forms[name] = form_summary(
                    gpu_bench(a.nx, ny, a.steps, a.warmup, p2, autotune=not a.no_autotune,
                              testcase=a.testcase), p2)
        line["forms"] = forms
    if not a.no_c2:
        c2 = {}
        for name in (a.precision, "fp64") if a.precision != "fp64" else ("fp64",):
            p2 = PRECISIONS[name]
            rc = gpu_bench(23000, 23000, 50, 4, p2, autotune=not a.no_autotune, testcase=1,
                           prof_steps=6)
            c2[name] = {**form_summary(rc, p2), "roofline": rc["roofline"]}
        line["c2"] = {"workload": "23000x23000 corner blast, 50 steps (33.9 GB of FP64 state "
                                  "in HBM)", **c2}
    if not a.no_traffic:
        tr = measured_traffic(a.nx, prec, r["tune"])
        kern = line["roofline"]["kernel"]
        if tr and kern in tr:
            line["roofline"]["traffic"] = round(tr[kern])
            line["roofline"]["traffic_note"] = (
                "HBM bytes per launch from rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled by "
                f"{tr['fetch_scale']:.3f} / {tr['write_scale']:.3f} (copy-kernel calibration)")
        va = measured_valu(a.nx, prec, r["tune"], kern, r["kernels"][kern]["avg_ms"] / 1e3)
        if va:
            line["valu"] = va
        progress("rocprofv3 traffic and VALU passes done")
    if not a.no_batched:
        line["batched"] = batched_bench(a.batch_n)
        progress("batched workload done")
    if not a.no_serving:
        line["serving"] = serving_bench(a.serve_n, (a.serve_rate, 1000.0))
        progress("serving workload done")
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_baseline(a.nx, ny, a.cpu_steps, a.testcase)
        progress("CPU baseline done")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
