"""Serving on the GPU: continuous batching with every result identical to
the same run alone, budget admission, failure isolation, and the
mixed-precision error bound against the FP64 path."""
import numpy as np
import pytest

import hydro_amd
from helpers import bits_differ
from hydro_amd import workloads as W

pytestmark = pytest.mark.gpu


def small_runs(n, seed, precision, lo=16, hi=96, nstepmax=60):
    return W.serving_requests(n, rate=1000.0, seed=seed, lo=lo, hi=hi, precision=precision,
                              tend=0.1, nstepmax=nstepmax)


def alone(run):
    with hydro_amd.Grid(run.params) as g:
        g.init_blast(run.bx, run.by, run.e)
        g.run(10**6, log=False)
        return g.download(), g.state()


@pytest.mark.parametrize("precision", [hydro_amd.FP64, hydro_amd.MIXED])
def test_results_same_alone_and_in_the_mix(precision):
    runs = small_runs(10, 3, precision)
    with hydro_amd.Server(precision=precision, budget_bytes=10**9, chunk_steps=7,
                          keep_results=True) as s:
        ids = [s.submit(r.params, r.bx, r.by, r.e) for r in runs[:4]]
        s.pump()
        s.pump()
        ids += [s.submit(r.params, r.bx, r.by, r.e) for r in runs[4:]]  # arrive mid-run
        while s.pump() > 0:
            pass
        assert s.counts()["done"] == len(runs)
        for rid, run in zip(ids, runs):
            st, t_sub, t_adm, t_done, t_sim, nstep = s.info(rid)
            assert st == s.DONE and t_sub <= t_adm <= t_done
            u_ref, (t_ref, n_ref, _) = alone(run)
            assert (t_sim, nstep) == (t_ref, n_ref)
            assert bits_differ(s.result(rid), u_ref) == 0, f"request {rid}"


def test_results_through_pinned_arena():
    """Results copied by DMA into reserved pinned memory, with the arena too
    small for all of them (the rest take the staged copy), released and
    reused: every result equals the run alone."""
    runs = small_runs(8, 6, hydro_amd.MIXED)
    with hydro_amd.Server(precision=hydro_amd.MIXED, budget_bytes=10**9, chunk_steps=9,
                          keep_results=True) as s:
        room = sum(16 * (r.params.nx + 4) * (r.params.ny + 4) + 256 for r in runs[:3])
        s.reserve_results(room)
        ids = [s.submit(r.params, r.bx, r.by, r.e) for r in runs]
        while s.pump() > 0:
            pass
        for rid, run in zip(ids, runs):
            assert bits_differ(s.result(rid), alone(run)[0]) == 0, f"request {rid}"
        for rid in ids:
            s.release(rid)
        again = [s.submit(r.params, r.bx, r.by, r.e) for r in runs[:3]]
        while s.pump() > 0:
            pass
        for rid, run in zip(again, runs[:3]):
            assert bits_differ(s.result(rid), alone(run)[0]) == 0, f"request {rid} (reused)"


def test_budget_and_oversized_request():
    runs = small_runs(6, 4, hydro_amd.MIXED, lo=64, hi=64)
    with hydro_amd.Batch(hydro_amd.MIXED) as b:
        one = b.bytes_needed(64, 64)
    with hydro_amd.Server(precision=hydro_amd.MIXED, budget_bytes=int(2.5 * one),
                          chunk_steps=10) as s:
        big = hydro_amd.Params(nx=128, ny=128, dx=1 / 128.0, courant_factor=0.8, tend=0.1,
                               nstepmax=10, precision=hydro_amd.MIXED)
        ids = [s.submit(r.params, r.bx, r.by, r.e) for r in runs]
        rbig = s.submit(big, 5, 5, 100.0)
        peak = 0
        while s.pump() > 0:
            peak = max(peak, s.counts()["running"])
        assert peak <= 2  # never more than the budget holds
        assert all(s.info(i)[0] == s.DONE for i in ids)
        assert s.info(rbig)[0] == s.FAILED


def test_failed_run_is_isolated():
    runs = small_runs(4, 5, hydro_amd.FP64)
    with hydro_amd.Server(precision=hydro_amd.FP64, budget_bytes=10**9, chunk_steps=5,
                          keep_results=True) as s:
        good = [s.submit(r.params, r.bx, r.by, r.e) for r in runs]
        # an infinite blast energy gives an infinite wave speed, so dt = 0
        bad = s.submit(runs[0].params, 1, 1, float("inf"))
        while s.pump() > 0:
            pass
        assert s.info(bad)[0] == s.FAILED
        for rid, run in zip(good, runs):
            assert s.info(rid)[0] == s.DONE
            assert bits_differ(s.result(rid), alone(run)[0]) == 0


def test_mixed_precision_error_vs_fp64():
    """Density and energy L1 (relative) of the mixed form against the FP64 path
    on serving-like runs; the FP64 path is bitwise the reference's."""
    runs = W.serving_requests(3, rate=1.0, seed=9, lo=128, hi=512, precision=hydro_amd.MIXED,
                              nstepmax=500)
    for r in runs:
        um, _ = alone(r)
        p64 = hydro_amd.Params(**{**r.params.__dict__, "precision": hydro_amd.FP64})
        with hydro_amd.Grid(p64) as g:
            g.init_blast(r.bx, r.by, r.e)
            g.run(10**6, log=False)
            u64 = g.download()
        for v, tol in ((0, 1e-5), (3, 1e-4)):
            a, b = um[v, 2:-2, 2:-2], u64[v, 2:-2, 2:-2]
            l1 = np.abs(a - b).sum() / np.abs(b).sum()
            assert l1 <= tol, (r.params.nx, v, l1)
