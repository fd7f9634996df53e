"""Seeded synthetic workloads for the batched and serving forms.

Both draw independent blast runs with the physics of the plain form (ideal
gas, MUSCL, CFL 0.8, reflective walls, rho = 1 and E = 1e-5 background):

* batch_instances -- N square grids, side log-uniform in [lo, hi], gamma
  uniform in [1.3, 1.67], one blast at a uniformly random cell with energy
  uniform in [0.1, 10] / dx^2, dx = 1 / side.
* serving_requests -- the same draws with Poisson arrival times at a given
  rate, each run to t_end (default 0.1) or nstepmax (default 500) steps.

numpy's PCG64 generator with a fixed seed makes the lists reproducible; the
first entries of each are pinned in tests/test_workloads.py.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from . import FP64, Params

SEED = 20261202


@dataclasses.dataclass
class Run:
    params: Params
    bx: int
    by: int
    e: float
    arrival: float = 0.0


def _draw(rng, lo, hi, precision, tend, nstepmax):
    side = int(np.exp(rng.uniform(np.log(lo), np.log(hi))))
    side = max(lo, min(hi, side))
    gamma = float(rng.uniform(1.3, 1.67))
    bx = int(rng.integers(0, side))
    by = int(rng.integers(0, side))
    dx = 1.0 / side
    e = float(rng.uniform(0.1, 10.0)) / dx / dx
    p = Params(nx=side, ny=side, dx=dx, gamma=gamma, courant_factor=0.8, niter_riemann=10,
               iorder=2, slope_type=1.0, tend=tend, nstepmax=nstepmax, precision=precision)
    return Run(p, bx, by, e)


def batch_instances(n: int = 512, seed: int = SEED, lo: int = 64, hi: int = 1024,
                    precision: int = FP64, steps: int = 100):
    rng = np.random.default_rng(seed)
    return [_draw(rng, lo, hi, precision, 1e30, steps) for _ in range(n)]


def serving_requests(n: int, rate: float, seed: int = SEED + 1, lo: int = 128, hi: int = 4096,
                     precision: int = FP64, tend: float = 0.1, nstepmax: int = 500):
    """n runs arriving as a Poisson process of `rate` per second (arrival
    offsets in seconds from the first)."""
    rng = np.random.default_rng(seed)
    runs = []
    t = 0.0
    for _ in range(n):
        r = _draw(rng, lo, hi, precision, tend, nstepmax)
        r.arrival = t
        runs.append(r)
        t += float(rng.exponential(1.0 / rate))
    return runs


def cell_updates(runs, steps=None) -> float:
    """Interior cell-updates if every run takes `steps` (default nstepmax) steps."""
    return float(sum(r.params.nx * r.params.ny * (steps or r.params.nstepmax) for r in runs))
