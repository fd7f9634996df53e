"""Parity on general flows (GPU vs oracle, bit for bit).

The benchmark's blasts are mostly quiescent background, where every face
takes the same branches.  These states vary everywhere (densities 0.1-10,
velocities up to +-5, pressures 0.01-100, with jumps), so shocks,
rarefactions and several Newton iterations meet the range-gated divisions
in every kernel form: the single-grid x sweep with the gated divisions (a
512 x 480 grid has enough x waves), the y sweep without them, and the batch
kernels, which always carry them.
"""
import numpy as np
import pytest

import hydro_amd
import oracle
from helpers import bits_differ, to_oracle

pytestmark = pytest.mark.gpu


def random_state(p, seed):
    rng = np.random.default_rng(seed)
    ny, nx = p.ny, p.nx
    rho = np.exp(rng.uniform(np.log(0.1), np.log(10.0), (ny, nx)))
    u = rng.uniform(-5.0, 5.0, (ny, nx))
    v = rng.uniform(-5.0, 5.0, (ny, nx))
    pr = np.exp(rng.uniform(np.log(0.01), np.log(100.0), (ny, nx)))
    # smooth the fields over blocks so faces see both jumps and near-equal states
    for f in (rho, u, v, pr):
        blk = rng.integers(1, 9)
        f[:] = np.repeat(np.repeat(f[::blk, ::blk], blk, 0), blk, 1)[:ny, :nx]
    U = np.zeros(p.shape())
    U[0, 2:-2, 2:-2] = rho
    U[1, 2:-2, 2:-2] = rho * u
    U[2, 2:-2, 2:-2] = rho * v
    U[3, 2:-2, 2:-2] = pr / (p.gamma - 1.0) + 0.5 * rho * (u * u + v * v)
    return U


@pytest.mark.parametrize("form", ["grid", "batch"])
@pytest.mark.parametrize("seed,slope_type,bounds", [(41, 1.0, (1, 1, 1, 1)),
                                                    (42, 1.5, (2, 2, 1, 1)),
                                                    (43, 2.0, (3, 3, 3, 3))])
def test_random_flows_bitwise(form, seed, slope_type, bounds):
    bl, br, bd, bu = bounds
    p = hydro_amd.Params(nx=512, ny=480, dx=1.0 / 512, courant_factor=0.8, tend=1e30,
                         slope_type=slope_type, boundary_left=bl, boundary_right=br,
                         boundary_down=bd, boundary_up=bu)
    u0 = random_state(p, seed)
    steps = 20
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
            t, n, _, _ = b.query()[i]
    ur = u0.copy()
    rdts, state = oracle.Port().run(to_oracle(p), ur, steps)
    if form == "grid":
        assert np.array_equal(dts, rdts), (dts, rdts)
    else:
        assert (t, n) == (state[0], state[1])
    assert np.isfinite(ur).all()
    nb = bits_differ(u, ur)
    assert nb == 0, f"{nb} values differ in their bits"


@pytest.mark.parametrize("form", ["grid", "batch"])
@pytest.mark.parametrize("gamma,smallr,smallc", [(1.01, 2.0 ** -100, 2.0 ** -50),
                                                 (3.0, 1e-10, 1e-10),
                                                 (1.67, 0.5, 2.0)])
def test_random_flows_at_parameter_bounds(form, gamma, smallr, smallc):
    """The gated divisions are enabled for gamma in [1.01, 3], smallr in
    [2^-100, 2^30] and smallc in [2^-50, 2^30]; these runs sit on those
    bounds (and on floors large enough to bind in many cells: smallr = 0.5
    against densities from 0.1, smallc = 2 against sound speeds of a few)."""
    p = hydro_amd.Params(nx=512, ny=480, dx=1.0 / 512, courant_factor=0.8, tend=1e30,
                         gamma=gamma, smallr=smallr, smallc=smallc)
    u0 = random_state(p, 50)
    steps = 12
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
            t, n, _, _ = b.query()[i]
    ur = u0.copy()
    rdts, state = oracle.Port().run(to_oracle(p), ur, steps)
    if form == "grid":
        assert np.array_equal(dts, rdts), (dts, rdts)
    else:
        assert (t, n) == (state[0], state[1])
    nb = bits_differ(u, ur)
    assert nb == 0, f"{nb} values differ in their bits"
