"""The benchmark grids at their full size, on the GPU, against the oracle.

The 8192x8192 centred blast (25 steps: past the warm-up and into the timed
window of the benchmark, both sweep orders, the halved first dt and every
dt reuse) and the 23000x23000 corner blast (2 steps, 33.9 GB of state) are
compared bit for bit with the oracle's OpenMP restatement in the FP64 form,
and to the contract's tolerance (per-field relative L1 <= 1e-12, dt to 1e-12)
in the FP64 fast form; the 23000x23000 states are checked for conservation
of mass and total energy (reflective walls: conserved to round-off).  The
100-step 8192x8192 comparison runs when HYDRO_LONG=1 (the oracle needs about
a minute for it).

At the benchmark's full lengths (200 steps of 8192x8192, 50 of 23000x23000)
the oracle would need many minutes, so there the fast form is held to the
contract against the FP64 form on the GPU, which is the oracle bit for bit
on these grids (above) and on every trajectory test.
"""
import math
import os

import numpy as np
import pytest

import hydro_amd
import oracle
from helpers import bits_differ
from test_gpu_fast import TOL_DT, TOL_L1, field_errors

pytestmark = pytest.mark.gpu


def run_gpu(n, testcase, steps, precision):
    p = hydro_amd.Params(nx=n, ny=n, dx=1.0 / n, courant_factor=0.8, tend=1e30,
                         precision=precision)
    with hydro_amd.Grid(p) as g:
        g.init(testcase)
        dts = g.run(steps)
        u = g.download()
    return u, dts


def run_oracle(n, testcase, steps):
    po = oracle.Params(nx=n, ny=n, dx=1.0 / n, courant_factor=0.8, tend=1e30, testcase=testcase)
    port = oracle.Port()
    ur = port.init(po)
    rdts, _ = port.run(po, ur, steps)
    return ur, rdts


def check_fast(u, dts, ur, rdts, what):
    errs = field_errors(u, ur)
    dtd = float(np.max(np.abs(dts - rdts) / rdts))
    print(f"{what} fast form: relative L1 / Linf per field {errs}, dt {dtd:.3g}")
    assert len(dts) == len(rdts)
    assert dtd <= TOL_DT, dtd
    assert all(e[0] <= TOL_L1 for e in errs), errs


@pytest.fixture(scope="module")
def c1_oracle_25():
    return run_oracle(8192, 0, 25)


def test_8192_centred_blast_25_steps_bitwise(c1_oracle_25):
    ur, rdts = c1_oracle_25
    u, dts = run_gpu(8192, 0, 25, hydro_amd.FP64)
    assert np.array_equal(dts, rdts), (dts, rdts)
    bad = bits_differ(u, ur)
    assert bad == 0, f"{bad} values differ (bit for bit)"


def test_8192_centred_blast_25_steps_fast_form(c1_oracle_25):
    ur, rdts = c1_oracle_25
    u, dts = run_gpu(8192, 0, 25, hydro_amd.FP64_FAST)
    check_fast(u, dts, ur, rdts, "8192^2, 25 steps")


@pytest.mark.skipif(os.environ.get("HYDRO_LONG") != "1", reason="set HYDRO_LONG=1 (about a minute)")
def test_8192_centred_blast_100_steps_bitwise_and_fast():
    ur, rdts = run_oracle(8192, 0, 100)
    u, dts = run_gpu(8192, 0, 100, hydro_amd.FP64)
    assert np.array_equal(dts, rdts)
    assert bits_differ(u, ur) == 0
    del u
    u, dts = run_gpu(8192, 0, 100, hydro_amd.FP64_FAST)
    check_fast(u, dts, ur, rdts, "8192^2, 100 steps")


def conserved(u, n):
    """Interior sums of rho and E against the initial state of the corner blast
    (rho = 1 everywhere, E = 1e-5 plus (1/dx)/dx in one cell): row sums, then
    an exactly rounded sum of the rows (a running float64 total over 23000
    rows would itself drift by ~1e-12 of the total)."""
    dx = 1.0 / n
    mass0 = float(n) * n
    energy0 = math.fsum([1e-5 * (float(n) * n - 1), (1.0 / dx) / dx])
    mass = math.fsum(np.sum(u[0, 2:-2, 2:-2], axis=1))
    energy = math.fsum(np.sum(u[3, 2:-2, 2:-2], axis=1))
    assert abs(mass - mass0) / mass0 < 1e-13, (mass, mass0)
    assert abs(energy - energy0) / energy0 < 1e-13, (energy, energy0)


@pytest.mark.skipif(os.environ.get("HYDRO_SKIP_23000") == "1", reason="asked to skip")
def test_23000_corner_blast_bitwise_fast_and_conserved():
    n = 23000
    ur, rdts = run_oracle(n, 1, 2)
    u, dts = run_gpu(n, 1, 2, hydro_amd.FP64)
    assert np.array_equal(dts, rdts), (dts, rdts)
    for v in range(4):
        bad = bits_differ(u[v], ur[v])
        assert bad == 0, f"variable {v}: {bad} values differ (bit for bit)"
    conserved(u, n)
    del u
    u, dts = run_gpu(n, 1, 2, hydro_amd.FP64_FAST)
    check_fast(u, dts, ur, rdts, "23000^2, 2 steps")
    conserved(u, n)


def test_8192_fast_form_200_steps_against_fp64_form():
    """The 8192x8192 centred blast over the benchmark's 200 steps."""
    u64, d64 = run_gpu(8192, 0, 200, hydro_amd.FP64)
    u, dts = run_gpu(8192, 0, 200, hydro_amd.FP64_FAST)
    check_fast(u, dts, u64, d64, "8192^2, 200 steps, against the FP64 form")


@pytest.mark.skipif(os.environ.get("HYDRO_SKIP_23000") == "1", reason="asked to skip")
def test_23000_fast_form_50_steps_against_fp64_form():
    """The 23000x23000 corner blast over the benchmark's 50 steps (two
    16.9 GB downloads); both forms conserve mass and total energy."""
    n = 23000
    u64, d64 = run_gpu(n, 1, 50, hydro_amd.FP64)
    conserved(u64, n)
    u, dts = run_gpu(n, 1, 50, hydro_amd.FP64_FAST)
    conserved(u, n)
    check_fast(u, dts, u64, d64, "23000^2, 50 steps, against the FP64 form")
