"""Helpers shared by the tests: parameter conversion and comparisons."""
import dataclasses

import numpy as np

import hydro_amd
import oracle

# 250x125 corner blast (E = 1/dx^2 in interior cell (0, 0)), CFL 0.8, 100 steps
CORNER_250x125 = dict(nx=250, ny=125, dx=1.0 / 250, courant_factor=0.8, niter_riemann=10,
                      iorder=2, slope_type=1.0, smallr=1e-10, smallc=1e-10, tend=1e30,
                      nstepmax=100)


def to_oracle(p: hydro_amd.Params, testcase: int = 0) -> oracle.Params:
    d = {f.name: getattr(p, f.name) for f in dataclasses.fields(oracle.Params)
         if hasattr(p, f.name)}
    d["testcase"] = testcase
    return oracle.Params(**d)


def bits_differ(a: np.ndarray, b: np.ndarray) -> int:
    """Number of values whose bit patterns differ, the sign of a zero included;
    a NaN matches a NaN whatever its payload."""
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    assert a.shape == b.shape and a.dtype == b.dtype, (a.shape, b.shape, a.dtype, b.dtype)
    ut = {8: np.uint64, 4: np.uint32}[a.dtype.itemsize]
    diff = a.view(ut) != b.view(ut)
    diff &= ~(np.isnan(a) & np.isnan(b))
    return int(np.count_nonzero(diff))


def rel_l1(a: np.ndarray, b: np.ndarray) -> float:
    """Relative L1 difference of the interiors of two [4, nyt, nxt] states, per field max."""
    a = a[:, 2:-2, 2:-2]
    b = b[:, 2:-2, 2:-2]
    worst = 0.0
    for v in range(a.shape[0]):
        den = np.abs(b[v]).sum()
        num = np.abs(a[v] - b[v]).sum()
        worst = max(worst, num / den if den > 0 else num)
    return worst
