"""The FP64 fast form (HYDRO_AMD_FP64_FAST) against the oracle, to the
tolerance of the parity contract rather than bit for bit: per-field relative
L1 <= 1e-12 after N steps, the dt sequence to 1e-12 (relative), identical
step counts, and mass conserved to round-off under reflective walls.

The form fuses multiply-adds and takes quotients and square roots from the
hardware estimates refined to about an ulp (hydro_numerics.hpp, prim_fast /
trace_fast / riemann_fast), so every operation differs from the reference's
by a few ulp at most.  The single-grid and batched kernels of the form share
their source and unit, so a batch instance equals its single-grid run bit for
bit.  (The reference loop: HydroC99_2DMpi/Src/main.c:244-298.)
"""
import numpy as np
import pytest

import hydro_amd
import oracle
from helpers import CORNER_250x125, bits_differ, to_oracle
from test_gpu_random_configs import COUNT, FIRST_SEED, draw

pytestmark = pytest.mark.gpu

TOL_L1 = 1e-12   # per-field relative L1 of the interior (the parity contract)
TOL_DT = 1e-12   # relative difference of every dt


def field_errors(a, b):
    """Relative L1 and L-infinity of each of the four interior fields of a
    against b (sum |a - b| / sum |b|, max |a - b| / max |b|)."""
    out = []
    for v in range(4):
        x, y = a[v, 2:-2, 2:-2], b[v, 2:-2, 2:-2]
        d = np.abs(x - y)
        den1, deninf = np.abs(y).sum(), np.abs(y).max()
        out.append((float(d.sum() / den1) if den1 > 0 else float(d.sum()),
                    float(d.max() / deninf) if deninf > 0 else float(d.max())))
    return out


def assert_within(u, ur, dts, rdts, what):
    errs = field_errors(u, ur)
    assert len(dts) == len(rdts), (what, len(dts), len(rdts))
    dtd = float(np.max(np.abs(dts - rdts) / np.abs(rdts))) if len(rdts) else 0.0
    assert dtd <= TOL_DT, (what, dtd)
    for v, (l1, _) in enumerate(errs):
        assert l1 <= TOL_L1, (what, v, errs)
    return errs


def test_c0_fast_100_steps_within_contract():
    """The 250x125 corner blast over 100 steps, run to a t_end that the
    reference reaches in exactly 100 steps (between its t after step 99 and
    after step 100), so the step count is a result of both runs."""
    p0 = hydro_amd.Params(**CORNER_250x125)
    po = to_oracle(p0, 1)
    po.nstepmax = 1000
    port = oracle.Port()
    ur = port.init(po)
    rdts, _ = port.run(po, ur.copy(), 100)
    t99, t100 = float(np.sum(rdts[:99])), float(np.sum(rdts))
    tend = 0.5 * (t99 + t100)
    po.tend = tend
    ur = port.init(po)
    rdts, rstate = port.run(po, ur, 1000)
    assert len(rdts) == 100
    p = hydro_amd.Params(**{**CORNER_250x125, "nstepmax": 1000, "tend": tend,
                            "precision": hydro_amd.FP64_FAST})
    with hydro_amd.Grid(p) as g:
        g.init(1)
        dts = g.run(1000)
        u = g.download()
        t, nstep, _ = g.state()
    assert nstep == 100 and len(dts) == 100, (nstep, len(dts))
    errs = assert_within(u, ur, dts, rdts, "C0")
    print("C0 fast: relative L1 / Linf per field", errs)
    assert abs(t - rstate[0]) <= TOL_DT * rstate[0]
    mass = np.sum(u[0, 2:-2, 2:-2])
    assert abs(mass - 250 * 125) / (250 * 125) < 1e-13, mass


def test_fast_grid_and_batch_identical():
    """A batch instance of the fast form equals its single-grid run bit for
    bit (one source, one unit), under a shared batch."""
    p = hydro_amd.Params(**{**CORNER_250x125, "precision": hydro_amd.FP64_FAST, "nstepmax": 30})
    q = hydro_amd.Params(**{**CORNER_250x125, "nx": 97, "ny": 70, "dx": 1.0 / 97,
                            "precision": hydro_amd.FP64_FAST, "nstepmax": 30})
    outs = []
    for par, tc in ((p, 1), (q, 0)):
        with hydro_amd.Grid(par) as g:
            g.init(tc)
            g.run(30)
            outs.append(g.download())
    with hydro_amd.Batch(hydro_amd.FP64_FAST) as b:
        i = b.add(p)
        j = b.add(q)
        b.init(i, 1)
        b.init(j, 0)
        b.run(30)
        assert bits_differ(b.download(i), outs[0]) == 0
        assert bits_differ(b.download(j), outs[1]) == 0


def test_fast_random_configurations():
    """The seeded random configurations of the bitwise test (sizes 3-900,
    every scheme option and boundary type, blasts, random and blocky states)
    through the fast form, on every run whose reference trajectory stays
    finite.

    Blast states (the benchmark's kind of flow): within the contract, each.

    Random and blocky states put cold gas at Mach 10-1000 (velocities to +-8
    over pressures down to 1e-3), where the pressure E/rho - |v|^2/2 cancels
    by up to five orders of magnitude; there the reference's own result
    moves beyond 1e-12 under a change of rounding alone -- its build with FMA
    contraction (the oracle with FMA contraction, as the reference compiled
    with -march=native would be) differs from it on about two thirds of these
    runs, by up to ~5e-6.  That build is the yardstick here.  Runs whose
    reference trajectory is numerically unstable (its largest energy grows
    more than 100-fold in a few steps: the collela and plmde predictors at
    high CFL numbers, 11 of the 150 seeds, growing by up to 1e25) amplify any
    difference without bound and are only reported, as are runs with
    niter_riemann <= 2, whose unconverged Riemann solves the fast form's
    regrouped Newton step does not follow to the yardstick's level (up to
    ~5e-4, against ~5e-6 for the FMA build).  On the others (niter_riemann >=
    3) the fast form must stay as close to the reference as the FMA build:
    its worst relative L1 within 10x the build's worst, and no more runs
    beyond 1e-12 than the build has plus a tenth of the runs."""
    fma = oracle.Port(fma=True)
    port = oracle.Port()
    blast_rows, rough, failures = [], [], []
    for seed in range(FIRST_SEED, FIRST_SEED + COUNT):
        p, u0, steps = draw(seed)
        po = to_oracle(p)
        ur = u0.copy()
        rdts, _ = port.run(po, ur, steps)
        if not (np.isfinite(ur).all() and np.isfinite(rdts).all()):
            continue
        p.precision = hydro_amd.FP64_FAST
        with hydro_amd.Grid(p) as g:
            g.upload(u0)
            dts = g.run(steps)
            ug = g.download()
        l1 = max(e[0] for e in field_errors(ug, ur))
        dtd = float(np.max(np.abs(dts - rdts) / np.abs(rdts)))
        if len(dts) != len(rdts):
            failures.append((seed, "step count", len(dts), len(rdts)))
            continue
        if np.count_nonzero(u0[1]) == 0:  # a blast on the quiescent background
            blast_rows.append((seed, l1, dtd))
            if l1 > TOL_L1 or dtd > TOL_DT:
                failures.append((seed, "blast", l1, dtd))
            continue
        uf = u0.copy()
        fma.run(po, uf, steps)
        l1f = max(e[0] for e in field_errors(uf, ur))
        growth = float(np.abs(ur[3, 2:-2, 2:-2]).max() / np.abs(u0[3, 2:-2, 2:-2]).max())
        kind = ("unstable" if growth > 100 else
                "truncated" if p.niter_riemann <= 2 else "rough")
        rough.append((kind, seed, l1, dtd, l1f, growth))
    print(f"fast form: {len(blast_rows)} blast runs, worst L1 "
          f"{max(r[1] for r in blast_rows):.3g}, worst dt {max(r[2] for r in blast_rows):.3g}")
    for kind in ("rough", "truncated", "unstable"):
        rows = [r for r in rough if r[0] == kind]
        if rows:
            f = np.array([r[2] for r in rows])
            q = np.array([r[4] for r in rows])
            print(f"{kind}: {len(rows)} runs; fast form worst L1 {f.max():.3g} (over 1e-12: "
                  f"{int((f > TOL_L1).sum())}); FMA-built reference worst {q.max():.3g} "
                  f"(over 1e-12: {int((q > TOL_L1).sum())})")
    for r in rough:
        print("  %-9s seed %d: fast L1 %.2e dt %.2e | FMA reference L1 %.2e | growth %.3g" % r)
    assert len(blast_rows) + len(rough) >= COUNT * 9 // 10
    assert not failures, failures
    rows = [r for r in rough if r[0] == "rough"]
    f = np.array([r[2] for r in rows])
    q = np.array([r[4] for r in rows])
    assert f.max() <= 10 * q.max(), (f.max(), q.max())
    assert (f > TOL_L1).sum() <= (q > TOL_L1).sum() + len(rows) // 10
    for r in rough:
        assert np.isfinite(r[2]), r
