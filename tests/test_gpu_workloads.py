"""The batched and serving workloads at their full configuration, on the GPU,
against the oracle.

* Batched: the 512 seeded blast instances (sides log-uniform in [64, 1024],
  gamma uniform in [1.3, 1.67], one blast per instance), 100 steps of each
  instance's own time loop (HydroC99_2DMpi/Src/main.c:244-298) in one FP64
  batch: every instance bit for bit its single-grid run, and the four
  smallest instances and the largest bit for bit the oracle's run.
* Serving: the 32 seeded requests of the serving mix (sides log-uniform in
  [128, 4096], t_end 0.1 or 500 steps) through the mixed-precision server
  with a 200 GB budget, all arriving at once: every result bit for bit the
  same mixed run alone; the relative L1 and L-infinity of all four fields
  against the oracle's FP64 run (sides up to 512; for larger sides against
  the FP64 grid, which is the oracle bit for bit on every workload tested),
  density L1 <= 1e-5; beside them the error of the reference's own all-FP32
  build (-DPREC_SP) on the same runs, where that build is present.
"""
import numpy as np
import pytest

import hydro_amd
import oracle
from helpers import bits_differ, to_oracle
from hydro_amd import workloads as W
from oracle import bindings
from test_gpu_fast import field_errors

pytestmark = pytest.mark.gpu


def test_batched_512_instances_100_steps():
    runs = W.batch_instances(512, steps=100)
    with hydro_amd.Batch(hydro_amd.FP64) as b:
        ids = [b.add(r.params) for r in runs]
        for i, r in zip(ids, runs):
            b.init_blast(i, r.bx, r.by, r.e)
        b.run(100)
        q = b.query()
        assert all(q[i][1] == 100 and q[i][3] == 0 for i in ids)
        got = {i: b.download(i) for i in ids}
    for k, (i, r) in enumerate(zip(ids, runs)):
        with hydro_amd.Grid(r.params) as g:
            g.init_blast(r.bx, r.by, r.e)
            g.run(100, log=False)
            u = g.download()
            t, n, _ = g.state()
        assert (q[i][0], q[i][1]) == (t, n), k
        assert bits_differ(got[i], u) == 0, f"instance {k} ({r.params.nx}^2)"
    order = sorted(range(len(runs)), key=lambda k: runs[k].params.nx)
    port = oracle.Port()
    for k in order[:4] + order[-1:]:
        r = runs[k]
        po = to_oracle(r.params)
        ur = port.init_blast(po, r.bx, r.by, r.e)
        port.run(po, ur, 100)
        assert bits_differ(got[ids[k]], ur) == 0, f"instance {k} ({r.params.nx}^2) vs oracle"


def ref_sp_run(r):
    """The reference's all-FP32 build on request r (its blast state set
    through the harness), or None when that build is absent."""
    if not __import__("os").path.exists(bindings.REF_SP_SO):
        return None
    po = to_oracle(r.params)
    ref = bindings.Ref(sp=True)
    ref.open(po, gamma=r.params.gamma)
    try:
        ref.set_uold(oracle.Port().init_blast(po, r.bx, r.by, r.e))
        ref.steps(r.params.nstepmax)
        return ref.get_uold()
    finally:
        ref.close()


def test_serving_mix_32_requests_against_oracle():
    runs = W.serving_requests(32, rate=1000.0, precision=hydro_amd.MIXED)
    with hydro_amd.Server(precision=hydro_amd.MIXED, budget_bytes=200 * 10**9, chunk_steps=20,
                          keep_results=True) as s:
        ids = [s.submit(r.params, r.bx, r.by, r.e) for r in runs]
        while s.pump() > 0:
            pass
        results = []
        for rid in ids:
            assert s.info(rid)[0] == s.DONE
            assert s.result_precision(rid) == hydro_amd.MIXED
            results.append(s.result(rid))
            s.release(rid)
    port = oracle.Port()
    worst = np.zeros((4, 2))
    lines = []
    for r, um in zip(runs, results):
        with hydro_amd.Grid(r.params) as g:  # the same mixed run alone
            g.init_blast(r.bx, r.by, r.e)
            g.run(10**6, log=False)
            assert bits_differ(um, g.download()) == 0, f"{r.params.nx}^2: mix != alone"
        p64 = hydro_amd.Params(**{**r.params.__dict__, "precision": hydro_amd.FP64})
        if r.params.nx <= 512:
            po = to_oracle(p64)
            u64 = port.init_blast(po, r.bx, r.by, r.e)
            port.run(po, u64, po.nstepmax)
            against = "oracle"
        else:
            with hydro_amd.Grid(p64) as g:
                g.init_blast(r.bx, r.by, r.e)
                g.run(10**6, log=False)
                u64 = g.download()
            against = "FP64 grid"
        e = np.array(field_errors(um, u64))
        worst = np.maximum(worst, e)
        line = f"{r.params.nx}^2 vs {against}: mixed L1/Linf " + " ".join(
            f"{a:.2e}/{b:.2e}" for a, b in e)
        if against == "oracle":
            usp = ref_sp_run(r)
            if usp is not None:
                esp = np.array(field_errors(usp, u64))
                line += "; reference FP32 build " + " ".join(f"{a:.2e}/{b:.2e}" for a, b in esp)
        lines.append(line)
        assert e[0, 0] <= 1e-5, line
    print("\n".join(lines))
    print("worst over the mix (rho, rho u, rho v, E; L1 / Linf):", worst.tolist())
