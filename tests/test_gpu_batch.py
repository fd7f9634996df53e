"""Batched grids on the GPU: every instance bit-identical to running it alone.

Single-grid runs (hydro_amd.Grid) are pinned to the oracle by
test_gpu_parity.py; a few instances here are also checked against the oracle
directly.
"""
import numpy as np
import pytest

import hydro_amd
import oracle
from helpers import bits_differ, to_oracle

pytestmark = pytest.mark.gpu


def make_instances(seed, count, lo=8, hi=180, precision=hydro_amd.FP64):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        nx = int(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        ny = int(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        g = float(rng.uniform(1.3, 1.67))
        p = hydro_amd.Params(nx=nx, ny=ny, dx=1.0 / max(nx, ny), gamma=g, courant_factor=0.8,
                             tend=1e30, precision=precision)
        bx, by = int(rng.integers(0, nx)), int(rng.integers(0, ny))
        e = float(rng.uniform(0.1, 10)) / p.dx ** 2
        out.append((p, bx, by, e))
    return out


def alone(p, bx, by, e, steps):
    with hydro_amd.Grid(p) as g:
        g.init_blast(bx, by, e)
        dts = g.run(steps)
        return g.download(), g.state(), dts


def check_same(u, ref, what):
    d = np.argwhere(u != ref)
    assert len(d) == 0, f"{what}: {len(d)} values differ, first {tuple(d[0])}"
    nb = bits_differ(u, ref)
    assert nb == 0, f"{what}: {nb} values differ in their bits (signed zeros)"


def test_batch_matches_single_grids():
    insts = make_instances(1, 12)
    with hydro_amd.Batch() as b:
        ids = []
        for p, bx, by, e in insts:
            i = b.add(p)
            b.init_blast(i, bx, by, e)
            ids.append(i)
        b.run(31)
        q = b.query()
        for i, (p, bx, by, e) in zip(ids, insts):
            u_ref, (t, n, dt), _ = alone(p, bx, by, e, 31)
            assert q[i][:3] == (t, n, dt), (i, q[i], (t, n, dt))
            check_same(b.download(i), u_ref, f"instance {i} {p.nx}x{p.ny}")


def test_batch_against_oracle():
    port = oracle.Port()
    insts = make_instances(2, 3, lo=20, hi=70)
    with hydro_amd.Batch() as b:
        ids = [b.add(p) for p, *_ in insts]
        for i, (p, bx, by, e) in zip(ids, insts):
            b.init_blast(i, bx, by, e)
        b.run(25)
        for i, (p, bx, by, e) in zip(ids, insts):
            po = to_oracle(p)
            ur = port.init_blast(po, bx, by, e)
            port.run(po, ur, 25)
            check_same(b.download(i), ur, f"oracle instance {i}")


def test_staggered_admission_and_removal():
    """Instances join with the others mid-run (mixed step parities) and leave."""
    insts = make_instances(3, 6)
    with hydro_amd.Batch() as b:
        first = [b.add(p) for p, *_ in insts[:3]]
        for i, (p, bx, by, e) in zip(first, insts[:3]):
            b.init_blast(i, bx, by, e)
        b.run(7)
        late = [b.add(p) for p, *_ in insts[3:]]
        for i, (p, bx, by, e) in zip(late, insts[3:]):
            b.init_blast(i, bx, by, e)
        b.run(9)
        b.remove(first[1])
        b.run(4)
        q = b.query()
        expect = {first[0]: 20, first[2]: 20, late[0]: 13, late[1]: 13, late[2]: 13}
        allinst = dict(zip(first + late, insts))
        for i, steps in expect.items():
            p, bx, by, e = allinst[i]
            u_ref, (t, n, dt), _ = alone(p, bx, by, e, steps)
            assert q[i][1] == n == steps
            assert q[i][0] == t and q[i][2] == dt
            check_same(b.download(i), u_ref, f"instance {i}")
        assert first[1] not in q


def test_instances_stop_at_their_own_limits():
    insts = make_instances(4, 4, lo=16, hi=64)
    limits = [(1e30, 5), (1e30, 12), (None, 1000), (1e30, 1000)]
    with hydro_amd.Batch() as b:
        ids = []
        for (p, bx, by, e), (tend, nmax) in zip(insts, limits):
            if tend is None:  # stop on time: t after 9 steps of this run
                _, (t9, _, _), _ = alone(p, bx, by, e, 9)
                tend = t9
            p.tend, p.nstepmax = tend, nmax
            i = b.add(p)
            b.init_blast(i, bx, by, e)
            ids.append(i)
        b.run(3)
        b.run(18)
        q = b.query()
        for i, (p, bx, by, e) in zip(ids, insts):
            u_ref, (t, n, dt), _ = alone(p, bx, by, e, 21)
            assert q[i][:3] == (t, n, dt)
            check_same(b.download(i), u_ref, f"instance {i}")
        assert [q[i][1] for i in ids] == [5, 12, 9, 21]
        assert [q[i][3] for i in ids] == [0, 0, 0, 1]


def test_batch_mixed_slope_types():
    """Instances with slope_type 1 and others share a phase (their tasks run in
    two kernel variants); each stays identical to its single-grid run."""
    insts = make_instances(7, 6)
    for k, (p, *_rest) in enumerate(insts):
        p.slope_type = (1.0, 2.0, 0.5)[k % 3]
    with hydro_amd.Batch() as b:
        ids = [b.add(p) for p, *_ in insts]
        for i, (p, bx, by, e) in zip(ids, insts):
            b.init_blast(i, bx, by, e)
        b.run(13)
        for i, (p, bx, by, e) in zip(ids, insts):
            u_ref, _, _ = alone(p, bx, by, e, 13)
            check_same(b.download(i), u_ref, f"slope_type {p.slope_type} instance {i}")


def test_budget_admission():
    with hydro_amd.Batch(budget_bytes=0) as probe:
        need = probe.bytes_needed(100, 100)
    p = hydro_amd.Params(nx=100, ny=100, dx=0.01, courant_factor=0.8, tend=1e30)
    with hydro_amd.Batch(budget_bytes=int(2.5 * need)) as b:
        a = b.add(p)
        c = b.add(p)
        assert b.try_add(p) is None
        assert "budget" in hydro_amd.last_error()
        b.remove(a)
        assert b.try_add(p) is not None
        assert b.bytes_used == 2 * need
        del c


@pytest.mark.parametrize("tune", [(1, 1, 1), (2, 3, 7), (4, 8, 64)])
def test_batch_tune_independence(tune):
    insts = make_instances(5, 5)
    with hydro_amd.Batch() as b:
        b.set_tune(*tune)
        ids = [b.add(p) for p, *_ in insts]
        for i, (p, bx, by, e) in zip(ids, insts):
            b.init_blast(i, bx, by, e)
        b.run(11)
        for i, (p, bx, by, e) in zip(ids, insts):
            u_ref, _, _ = alone(p, bx, by, e, 11)
            check_same(b.download(i), u_ref, f"tune {tune} instance {i}")


def test_batch_autotune_keeps_state():
    insts = make_instances(6, 6)
    with hydro_amd.Batch() as b:
        ids = [b.add(p) for p, *_ in insts]
        for i, (p, bx, by, e) in zip(ids, insts):
            b.init_blast(i, bx, by, e)
        b.run(3)
        before = {i: b.download(i) for i in ids}
        q0 = b.query()
        best = b.autotune(reps=1)
        assert set(best) == {"waves", "xiter", "yseg"}
        assert b.query() == q0
        for i in ids:
            check_same(b.download(i), before[i], "autotune changed state")
        b.run(6)
        for i, (p, bx, by, e) in zip(ids, insts):
            u_ref, _, _ = alone(p, bx, by, e, 9)
            check_same(b.download(i), u_ref, f"after autotune instance {i}")
