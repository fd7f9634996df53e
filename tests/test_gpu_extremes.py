"""Parity far from the benchmark's magnitudes (GPU vs oracle, bit for bit).

The sweeps take shortcuts that are exact only inside a range: the square
root drops the library sequence's range fix-up when every lane's argument is
in [2^-767, DBL_MAX]; divisions drop v_div_scale when a wave's cells and
faces pass the range tests (densities and pressures in [2^-41, 2^40),
velocities, momenta and energies zero or in that range, the Newton pressure
within [2^-151, 2^150)); and the x sweep's limiter is specialised for
slope_type == 1.  These states push arguments out of those ranges (tiny
floors, near-vacuum cells, huge energies, zero and negative inputs, values
straddling the division tests' bounds) so the general paths run, alone and
beside the fast ones, and compare every value with the oracle (a NaN matches
a NaN).
"""
import numpy as np
import pytest

import hydro_amd
import oracle
from helpers import bits_differ, to_oracle

pytestmark = pytest.mark.gpu


def same(a, b):
    """Bit for bit (signed zeros included), a NaN matching a NaN."""
    return bits_differ(np.asarray(a), np.asarray(b)) == 0


def params(nx, ny, **kw):
    base = dict(nx=nx, ny=ny, dx=1.0 / nx, courant_factor=0.8, tend=1e30)
    base.update(kw)
    return hydro_amd.Params(**base)


def run_both(p, u0, steps, form="grid"):
    """The state after `steps` steps on the GPU and in the oracle.  form
    "grid": a single grid (these small grids take the sweep variants without
    the range-gated divisions); "batch": a one-instance batch, whose
    persistent kernels always carry them, so the gates' fast and IEEE
    branches both run here."""
    if form == "grid":
        with hydro_amd.Grid(p) as g:
            g.upload(u0)
            dts = g.run(steps)
            u = g.download()
    else:
        with hydro_amd.Batch(hydro_amd.FP64) as b:
            i = b.add(p)
            b.upload(i, u0)
            b.run(steps)
            u = b.download(i)
            t, n, dt, _ = b.query()[i]
        dts = None
    port = oracle.Port()
    ur = u0.copy()
    rdts, state = port.run(to_oracle(p), ur, steps)
    if dts is None:  # the batch reports the run's end state, not each dt
        assert (t, n) == (state[0], state[1]) or (np.isnan(t) and np.isnan(state[0]))
        dts = rdts
    return u, dts, ur, rdts


FORMS = ["grid", "batch"]


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("slope_type", [1.0, 1.5])
def test_tiny_scale_state(slope_type, form):
    """A smooth blast scaled to rho ~ p ~ 1e-120 (smallr = smallc = 1e-200):
    every Riemann square root (gamma * p * rho ~ 1e-240) is below 2^-767, so
    the general sequence runs everywhere; the state stays finite."""
    p = params(90, 70, smallr=1e-200, smallc=1e-200, slope_type=slope_type)
    u = np.zeros(p.shape())
    rng = np.random.default_rng(11)
    s = 1e-120
    u[0, 2:-2, 2:-2] = s * (1 + 0.1 * rng.random((70, 90)))
    u[3, 2:-2, 2:-2] = s * (2.5 + 0.1 * rng.random((70, 90)))
    u[3, 30:34, 40:44] = s * 25.0
    g, dts, r, rdts = run_both(p, u, 8, form)
    assert len(dts) == 8 and np.isfinite(dts).all() and np.isfinite(g).all()
    assert np.array_equal(dts, rdts), (dts, rdts)
    assert same(g, r), f"{np.count_nonzero(g != r)} values differ"


def test_near_vacuum_blow_up(form="grid"):
    # (a batch stops an instance whose dt stops being finite, which the
    # oracle's run does not, so this breakdown is compared on a single grid)
    """Densities spread over 250 decades next to a blast: the run breaks down
    into NaNs within a few steps, identically (NaN matching NaN)."""
    p = params(90, 70, smallr=1e-300, smallc=1e-300)
    u = np.zeros(p.shape())
    rng = np.random.default_rng(11)
    u[0, 2:-2, 2:-2] = 10.0 ** rng.uniform(-250, 0, (70, 90))
    u[1, 2:-2, 2:-2] = u[0, 2:-2, 2:-2] * rng.uniform(-1, 1, (70, 90))
    u[2, 2:-2, 2:-2] = u[0, 2:-2, 2:-2] * rng.uniform(-1, 1, (70, 90))
    u[3, 2:-2, 2:-2] = 10.0 ** rng.uniform(-290, -200, (70, 90)) + 0.5 * (
        u[1, 2:-2, 2:-2] ** 2 + u[2, 2:-2, 2:-2] ** 2) / u[0, 2:-2, 2:-2]
    u[3, 30:34, 40:44] = 1e5
    g, dts, r, rdts = run_both(p, u, 8, form)
    assert len(dts) >= 2
    assert same(dts, rdts), (dts, rdts)
    assert same(g, r), f"{np.count_nonzero(~((g == r) | (np.isnan(g) & np.isnan(r))))} differ"


def test_huge_and_zero_values(form="grid"):
    """Energies near 1e300, exact zeros in density and energy, smallc = 0."""
    p = params(64, 48, smallc=0.0, smallr=1e-10)
    u = np.zeros(p.shape())
    rng = np.random.default_rng(12)
    u[0, 2:-2, 2:-2] = rng.choice([0.0, 1e-300, 1.0, 1e150], size=(48, 64))
    u[3, 2:-2, 2:-2] = rng.choice([0.0, 1e-310, 1e-5, 1e300], size=(48, 64))
    u[1, 2:-2, 2:-2] = rng.choice([0.0, -0.0, 1e-200, -3.0], size=(48, 64))
    g, dts, r, rdts = run_both(p, u, 3, form)
    assert len(dts) >= 1
    assert same(dts, rdts), (dts, rdts)
    assert same(g, r)


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("seed", [21, 22])
def test_division_gate_boundaries(seed, form):
    """Densities, pressures and momenta scattered just inside and just outside
    the division tests' bounds (2^-41, 2^40), in some columns and rows only:
    waves and faces of one sweep take the fast and the IEEE divisions side by
    side, and the result is the oracle's bit for bit."""
    p = params(150, 90)
    rng = np.random.default_rng(seed)
    u = np.zeros(p.shape())
    shape = (90, 150)
    u[0, 2:-2, 2:-2] = 1.0 + 0.1 * rng.random(shape)
    u[3, 2:-2, 2:-2] = 2.5 + 0.1 * rng.random(shape)
    # columns 40..59 near the upper bound, rows 30..44 near the lower one
    hi = 2.0 ** rng.uniform(39.5, 40.5, (90, 20))
    u[0, 2:-2, 42:62] = hi
    u[3, 2:-2, 42:62] = hi * rng.uniform(0.5, 2.0, (90, 20))
    lo = 2.0 ** rng.uniform(-41.5, -40.5, (15, 150))
    u[0, 32:47, 2:-2] = np.maximum(lo, 1e-9)
    u[1, 32:47, 2:-2] = lo * rng.uniform(-1, 1, (15, 150))
    u[3, 32:47, 2:-2] = lo * rng.uniform(0.5, 2.0, (15, 150))
    u[3, 70:74, 100:104] = 1e6
    g, dts, r, rdts = run_both(p, u, 6, form)
    assert np.array_equal(dts, rdts), (dts, rdts)
    assert same(g, r), f"{np.count_nonzero(~((g == r) | (np.isnan(g) & np.isnan(r))))} differ"


@pytest.mark.parametrize("form,nx,ny", [("grid", 200, 150), ("batch", 200, 150),
                                        ("grid", 1024, 4096)])
def test_signed_zero_states(form, nx, ny):
    """In-range states with +0 and -0 momenta and energies: the range-gated
    divisions take them (zero is in range), and the quotients that drop
    v_div_fixup must give the zeros' signs exactly where the reference's do
    (the zeros meet the first step's primitives, traces and Riemann inputs;
    the flow has filled them in after one step).
    The 1024 x 4096 grid is large enough for the single-grid kernels with the
    gated divisions; the batch form always has them."""
    p = params(nx, ny)
    rng = np.random.default_rng(31)
    u = np.zeros(p.shape())
    shp = (ny, nx)
    u[0, 2:-2, 2:-2] = rng.uniform(0.5, 2.0, shp)
    for v in (1, 2):
        u[v, 2:-2, 2:-2] = np.where(rng.random(shp) < 0.6,
                                    rng.choice([0.0, -0.0], size=shp),
                                    rng.uniform(-1.0, 1.0, shp))
    u[3, 2:-2, 2:-2] = np.where(rng.random(shp) < 0.1, rng.choice([0.0, -0.0], size=shp),
                                rng.uniform(0.5, 5.0, shp))
    for steps in (1, 4):
        g, dts, r, rdts = run_both(p, u, steps, form)
        assert np.array_equal(dts, rdts), (steps, dts, rdts)
        assert same(g, r), f"after {steps} steps: {bits_differ(g, r)} values differ in their bits"
