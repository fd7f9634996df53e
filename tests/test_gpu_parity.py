"""GPU parity: the HIP path (through the C ABI) against the oracle, bitwise.

The oracle (oracle/hydro_oracle.c) is itself bit-identical to the reference
HydroC99_2DMpi sources (tests/test_oracle.py), so equality here is equality
with the reference.  The bar for the FP64 path is bitwise; the tests also
state the 1e-12 relative-L1 bound the contract allows, so a failure report
shows which of the two broke.
"""
import numpy as np
import pytest

import hydro_amd
import oracle
from helpers import CORNER_250x125, bits_differ, rel_l1, to_oracle

pytestmark = pytest.mark.gpu

TOL_L1 = 1e-12  # per-field relative L1 allowed for FP64 by the contract


@pytest.fixture(scope="module")
def port():
    return oracle.Port()


def assert_same(gpu, ref, what):
    l1 = rel_l1(gpu, ref)
    diff = np.argwhere(gpu != ref)
    msg = f"{what}: {len(diff)} differing values, rel L1 {l1:.3e}" + (
        f", first at {tuple(diff[0])}: {gpu[tuple(diff[0])]!r} vs {ref[tuple(diff[0])]!r}"
        if len(diff) else "")
    assert l1 <= TOL_L1, msg
    assert len(diff) == 0, msg
    nb = bits_differ(gpu, ref)
    assert nb == 0, f"{what}: {nb} values differ in their bits (signed zeros)"


def surface_loop(g, steps):
    """HydroC99_2DMpi/Src/main.c:244-298 driven from the host through the grid API."""
    dts = []
    dt = 0.0
    for n in range(steps):
        if n % 2 == 0:
            dt = g.compute_deltat()
            if n == 0:
                dt = dt / 2.0
        g.godunov(1 if n % 2 == 0 else 2, dt)
        dts.append(dt)
    return np.array(dts)


def test_c0_host_loop_100_steps(port):
    """The 250x125 corner blast, 100 steps, host-driven loop."""
    p = hydro_amd.Params(**CORNER_250x125)
    with hydro_amd.Grid(p) as g:
        g.init(1)
        dts = surface_loop(g, 100)
        u = g.download()
    po = to_oracle(p, testcase=1)
    ur = port.init(po)
    rdt, st = port.run(po, ur, 100)
    assert np.array_equal(dts, rdt), f"dt differs: max rel {np.max(np.abs(dts/rdt-1)):.3e}"
    assert_same(u, ur, "uold after 100 steps")


def test_c0_device_loop_100_steps(port):
    """The same run as a device-side loop (hipGraph replay); t and nstep too."""
    p = hydro_amd.Params(**CORNER_250x125)
    with hydro_amd.Grid(p) as g:
        g.init(1)
        dts = g.run(100)
        t, n, _ = g.state()
        u = g.download()
    po = to_oracle(p, testcase=1)
    ur = port.init(po)
    rdt, (rt, rn, _) = port.run(po, ur, 100)
    assert n == rn == 100
    assert np.array_equal(dts, rdt)
    assert t == rt
    assert_same(u, ur, "uold after 100 device steps")


def test_device_loop_chunks_and_tend(port):
    """Runs split into odd-sized chunks and stopped by tend give the same steps."""
    p = hydro_amd.Params(**{**CORNER_250x125, "nx": 61, "ny": 67, "dx": 1 / 61.0})
    po = to_oracle(p, testcase=0)
    ur = port.init(po)
    rdt, (rt, _, _) = port.run(po, ur, 40)
    tend = float(np.cumsum(rdt)[24])  # stops after step 25 (t >= tend)
    p2 = hydro_amd.Params(**{**CORNER_250x125, "nx": 61, "ny": 67, "dx": 1 / 61.0, "tend": tend})
    po2 = to_oracle(p2, testcase=0)
    ur2 = port.init(po2)
    rdt2, (rt2, rn2, _) = port.run(po2, ur2, 1000)
    with hydro_amd.Grid(p2) as g:
        g.init(0)
        got = []
        for k in (3, 1, 4, 2, 7, 100):
            got.extend(g.run(k))
        t, n, _ = g.state()
        u = g.download()
    assert n == rn2 and len(got) == rn2
    assert np.array_equal(np.array(got), rdt2)
    assert t == rt2
    assert_same(u, ur2, "uold after tend stop")


@pytest.mark.parametrize("nx,ny", [(1, 1), (3, 2), (5, 130), (59, 7), (60, 60), (61, 64),
                                   (121, 33), (64, 257)])
def test_shapes(port, nx, ny):
    """Ragged shapes: widths around the 60-cell x segment, tiny and tall grids."""
    p = hydro_amd.Params(nx=nx, ny=ny, dx=1.0 / max(nx, ny), courant_factor=0.8, tend=1e30)
    po = to_oracle(p, testcase=0)
    ur = port.init(po)
    rdt, _ = port.run(po, ur, 12)
    with hydro_amd.Grid(p) as g:
        g.init(0)
        dts = g.run(12)
        u = g.download()
    assert np.array_equal(dts, rdt)
    assert_same(u, ur, f"{nx}x{ny}")


VARIANTS = {
    "first_order": dict(iorder=1),
    "slope_half": dict(slope_type=0.5),
    "slope_two": dict(slope_type=2.0),
    "plmde": dict(scheme=2),
    "collela": dict(scheme=3),
    "newton0": dict(niter_riemann=0),
    "newton1": dict(niter_riemann=1),
    "gamma_5_3": dict(gamma=5.0 / 3.0),
    "absorbing": dict(boundary_left=2, boundary_right=2, boundary_down=2, boundary_up=2),
    "periodic": dict(boundary_left=3, boundary_right=3, boundary_down=3, boundary_up=3),
    "mixed_walls": dict(boundary_left=1, boundary_right=2, boundary_down=3, boundary_up=0),
}


@pytest.mark.parametrize("name", sorted(VARIANTS))
def test_scheme_variants(port, name):
    p = hydro_amd.Params(**{**CORNER_250x125, "nx": 97, "ny": 70, "dx": 1 / 97.0, **VARIANTS[name]})
    po = to_oracle(p, testcase=0)
    ur = port.init(po)
    rdt, _ = port.run(po, ur, 30)
    with hydro_amd.Grid(p) as g:
        g.init(0)
        dts = g.run(30)
        u = g.download()
    assert np.array_equal(dts, rdt), name
    assert_same(u, ur, name)


def test_surface_calls_on_uploaded_state(port):
    """make_boundary / compute_deltat / godunov on an arbitrary uploaded state."""
    p = hydro_amd.Params(**{**CORNER_250x125, "nx": 77, "ny": 45, "dx": 1 / 77.0,
                            "boundary_left": 1, "boundary_right": 2,
                            "boundary_down": 3, "boundary_up": 1})
    po = to_oracle(p, testcase=1)
    rng = np.random.default_rng(7)
    u = port.init(po)
    port.run(po, u, 6)
    u[:, 2:-2, 2:-2] *= 1 + 0.05 * rng.standard_normal(u[:, 2:-2, 2:-2].shape)
    u[0] = np.abs(u[0])
    u[3] = np.abs(u[3]) + 1e-3
    with hydro_amd.Grid(p) as g:
        g.upload(u)
        assert g.compute_deltat() == port.compute_deltat(po, u)
        for idim in (1, 2):
            g.make_boundary(idim)
            port.make_boundary(idim, po, u)
            assert_same(g.download(), u, f"make_boundary({idim})")
        for idim, dt in ((1, 3e-5), (2, 2e-5)):
            g.godunov(idim, dt)
            port.godunov(idim, dt, po, u)
            assert_same(g.download(), u, f"godunov({idim})")
        assert g.compute_deltat() == port.compute_deltat(po, u)


@pytest.mark.parametrize("yseg,ywaves,xwaves,xiter", [(1, 1, 1, 1), (7, 2, 3, 5), (64, 4, 4, 8),
                                                     (1000, 3, 2, 64)])
def test_tiling_independence(port, yseg, ywaves, xwaves, xiter):
    """Results do not depend on the sweep launch shapes."""
    p = hydro_amd.Params(**{**CORNER_250x125, "nx": 130, "ny": 150, "dx": 1 / 130.0})
    po = to_oracle(p, testcase=0)
    ur = port.init(po)
    rdt, _ = port.run(po, ur, 10)
    with hydro_amd.Grid(p) as g:
        g.set_tune(yseg, ywaves, xwaves, xiter)
        g.init(0)
        dts = g.run(10)
        u = g.download()
    assert np.array_equal(dts, rdt)
    assert_same(u, ur, f"tune {yseg} {ywaves} {xwaves} {xiter}")


def test_autotune_keeps_state_and_results(port):
    """The autotuner only writes scratch: the state and the run are unchanged."""
    p = hydro_amd.Params(**{**CORNER_250x125, "nx": 300, "ny": 200, "dx": 1 / 300.0})
    po = to_oracle(p, testcase=1)
    ur = port.init(po)
    rdt, _ = port.run(po, ur, 8)
    with hydro_amd.Grid(p) as g:
        g.init(1)
        before = g.download()
        best = g.autotune(reps=1)
        assert set(best) == {"yseg", "ywaves", "xwaves", "xiter"}
        assert np.array_equal(g.download(), before)
        dts = g.run(8)
        u = g.download()
    assert np.array_equal(dts, rdt)
    assert_same(u, ur, f"after autotune {best}")


def test_autotune_mid_run_keeps_the_run(port):
    """Tuned between steps (after an odd step, with dt still to come from the
    CFL maxima its last sweep folded): the autotuner's whole-step trials run
    from a saved copy of the state and the CFL slots, so the run continues
    bit for bit -- against the oracle in the FP64 form, against an untuned
    run in the fast form."""
    p = hydro_amd.Params(**{**CORNER_250x125, "nx": 300, "ny": 200, "dx": 1 / 300.0})
    po = to_oracle(p, testcase=1)
    ur = port.init(po)
    rdt, _ = port.run(po, ur, 9)
    with hydro_amd.Grid(p) as g:
        g.init(1)
        d1 = g.run(3)
        state = g.state()
        g.autotune(reps=1)
        assert g.state() == state
        d2 = g.run(6)
        u = g.download()
    assert np.array_equal(np.concatenate([d1, d2]), rdt)
    assert_same(u, ur, "FP64, tuned after step 3")
    pf = hydro_amd.Params(**{**p.__dict__, "precision": hydro_amd.FP64_FAST})
    outs = []
    for tune in (False, True):
        with hydro_amd.Grid(pf) as g:
            g.init(1)
            d1 = g.run(3)
            if tune:
                g.autotune(reps=1)
            d2 = g.run(6)
            outs.append((np.concatenate([d1, d2]), g.download()))
    assert np.array_equal(outs[0][0], outs[1][0])
    assert bits_differ(outs[0][1], outs[1][1]) == 0


def test_mid_size_against_oracle(port):
    """1024x768 centred blast, 20 steps, bitwise."""
    p = hydro_amd.Params(**{**CORNER_250x125, "nx": 1024, "ny": 768, "dx": 1 / 1024.0})
    po = to_oracle(p, testcase=0)
    ur = port.init(po)
    rdt, _ = port.run(po, ur, 20)
    with hydro_amd.Grid(p) as g:
        g.init(0)
        dts = g.run(20)
        u = g.download()
    assert np.array_equal(dts, rdt)
    assert_same(u, ur, "1024x768")
