"""Testcase 2, the Sod column (HydroC99_2DMpi/Src/hydro_funcs.c:121-128:
E = (1/dx)/dx in padded column 2 of every padded row from 2 on, the two
upper ghost rows included), on the GPU bit for bit: the initial state and the
per-step hashes of the whole uold (ghosts included) of the reference's own
trajectory (tests/golden/traj_sod64x16.npz, written by the compiled
reference), the oracle, the single-grid device loop step by step and a batch
instance."""
import hashlib
import json
import os

import numpy as np
import pytest

import hydro_amd
import oracle
from helpers import bits_differ, to_oracle

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "traj_sod64x16.npz")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def fixture():
    g = np.load(GOLD)
    par = json.loads(str(g["params"]))
    p = hydro_amd.Params(nx=par["nx"], ny=par["ny"], dx=par["dx"],
                         courant_factor=par["courant_factor"], tend=par["tend"])
    assert par["testcase"] == 2
    return p, g


def test_sod_grid_step_by_step_matches_reference_hashes(fixture):
    p, g = fixture
    with hydro_amd.Grid(p) as grid:
        grid.init(2)
        u0 = grid.download()
        assert bits_differ(u0, g["uold0"]) == 0, "initial state"
        for s in range(len(g["dt"])):
            dt = grid.run(1)
            assert len(dt) == 1 and dt[0] == g["dt"][s], (s, dt, g["dt"][s])
            u = grid.download()
            assert sha(u) == g["hashes"][s], f"uold differs after step {s + 1}"
        t, n, _ = grid.state()
    assert n == len(g["dt"]) and t == g["t"][-1]
    assert bits_differ(u, g["uold"]) == 0


def test_sod_device_loop_and_batch_match_oracle(fixture):
    p, g = fixture
    steps = len(g["dt"])
    port = oracle.Port()
    po = to_oracle(p, 2)
    ur = port.init(po)
    rdts, (rt, rn, _) = port.run(po, ur, steps)
    assert np.array_equal(rdts, g["dt"]) and sha(ur) == g["hashes"][-1]
    with hydro_amd.Grid(p) as grid:
        grid.init(2)
        dts = grid.run(steps)
        u = grid.download()
    assert np.array_equal(dts, rdts) and bits_differ(u, ur) == 0
    with hydro_amd.Batch(hydro_amd.FP64) as b:
        i = b.add(p)
        j = b.add(hydro_amd.Params(**{**p.__dict__, "nx": 40, "ny": 30, "dx": 1 / 40.0}))
        b.init(i, 2)
        b.init(j, 2)
        b.run(steps)
        ub = b.download(i)
        t, n, _, _ = b.query()[i]
    assert (t, n) == (rt, rn) == (g["t"][-1], steps)
    assert bits_differ(ub, ur) == 0
