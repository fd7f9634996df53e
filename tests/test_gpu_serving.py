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


def test_requests_that_can_never_run_fail_at_once():
    """Requests a batch would always refuse -- another precision, invalid
    parameters, a state larger than the whole budget -- placed ahead of good
    ones fail at submission, and the good ones behind them are served without
    waiting for the queue to drain."""
    runs = small_runs(4, 8, hydro_amd.MIXED, lo=32, hi=64)
    with hydro_amd.Batch(hydro_amd.MIXED) as b:
        budget = 4 * b.bytes_needed(64, 64)
    base = runs[0].params
    bad = [hydro_amd.Params(**{**base.__dict__, "precision": hydro_amd.FP64}),
           hydro_amd.Params(**{**base.__dict__, "gamma": 0.9}),
           hydro_amd.Params(**{**base.__dict__, "scheme": 7}),
           hydro_amd.Params(**{**base.__dict__, "nx": 1024, "ny": 1024, "dx": 1 / 1024.0})]
    with hydro_amd.Server(precision=hydro_amd.MIXED, budget_bytes=budget, chunk_steps=10,
                          keep_results=True) as s:
        bids = [s.submit(p, 1, 1, 100.0) for p in bad]
        for i in bids:
            assert s.info(i)[0] == s.FAILED  # before any pump
        ids = [s.submit(r.params, r.bx, r.by, r.e) for r in runs]
        s.pump()
        assert s.counts()["running"] > 0  # the good ones were admitted at the first pump
        while s.pump() > 0:
            pass
        for rid, run in zip(ids, runs):
            assert s.info(rid)[0] == s.DONE
            assert bits_differ(s.result(rid), alone(run)[0]) == 0
        assert s.counts()["failed"] == len(bad)


def hypersonic_state(seed):
    """A state of the random-configuration draws with cells whose kinetic
    energy dwarfs their internal energy (velocities to +-8, pressures down to
    1e-3): the FP32 total energy of the mixed form loses their pressure."""
    from test_gpu_random_configs import draw
    p, u, steps = draw(seed)
    return p, u, steps


def test_mixed_breakdown_is_run_again_in_fp64():
    """Mixed-precision requests that break down -- a blast whose fluxes overflow
    FP32 (E = 1e35) and a hypersonic random state on which the mixed form goes
    non-finite -- end DONE, run again in FP64, with results bit for bit those
    of the FP64 run alone; an ordinary request beside them stays mixed."""
    blast = hydro_amd.Params(nx=48, ny=40, dx=1 / 48.0, courant_factor=0.8, tend=1e30,
                             nstepmax=30, precision=hydro_amd.MIXED)
    # a random-configuration state on which the mixed form breaks down alone
    hyper = None
    for seed in range(20000, 20200):
        p, u0, steps = hypersonic_state(seed)
        if p.nx * p.ny > 300 * 300:
            continue
        p.nstepmax, p.tend = steps + 20, 1e30
        p64 = hydro_amd.Params(**{**p.__dict__, "precision": hydro_amd.FP64})
        with hydro_amd.Grid(p64) as g:
            g.upload(u0)
            g.run(p.nstepmax, log=False)
            u64 = g.download()
        if not np.isfinite(u64).all():
            continue
        p.precision = hydro_amd.MIXED
        with hydro_amd.Grid(p) as g:
            g.upload(u0)
            g.run(p.nstepmax, log=False)
            um = g.download()
        if not np.isfinite(um).all():
            hyper = (p, u0, u64)
            break
    assert hyper is not None, "no random state broke the mixed form"
    ok = small_runs(1, 11, hydro_amd.MIXED)[0]
    with hydro_amd.Server(precision=hydro_amd.MIXED, budget_bytes=10**9, chunk_steps=8,
                          keep_results=True) as s:
        rb = s.submit(blast, 20, 10, 1e35)
        rh = s.submit_state(hyper[0], hyper[1])
        ro = s.submit(ok.params, ok.bx, ok.by, ok.e)
        while s.pump() > 0:
            pass
        assert s.info(rb)[0] == s.DONE and s.info(rh)[0] == s.DONE and s.info(ro)[0] == s.DONE
        assert s.result_precision(rb) == hydro_amd.FP64
        assert s.result_precision(rh) == hydro_amd.FP64
        assert s.result_precision(ro) == hydro_amd.MIXED
        assert s.reruns == 2
        b64 = hydro_amd.Params(**{**blast.__dict__, "precision": hydro_amd.FP64})
        with hydro_amd.Grid(b64) as g:
            g.init_blast(20, 10, 1e35)
            g.run(10**6, log=False)
            ub = g.download()
        assert np.isfinite(ub).all()
        assert bits_differ(s.result(rb), ub) == 0
        assert bits_differ(s.result(rh), hyper[2]) == 0
        assert bits_differ(s.result(ro), alone(ok)[0]) == 0
