"""The oracle pinned: golden vectors, the live reference, and invariants (CPU).

Three independent anchors for oracle/hydro_oracle.c (the oracle every GPU
test compares against):
 1. the golden vectors written by tests/golden/make_golden.py from the
    reference's own compiled sources -- every stage, and whole trajectories
    (dt, t and a hash of the full uold after every step);
 2. the compiled reference itself (oracle/_ref) on fresh random states, where
    it is present, and its unmodified executable's printed step lines;
 3. properties fixed by the mathematics, which do not depend on the
    reference at all: conservation with reflective walls, transposition
    symmetry of the split step (pins the IU/IV swap and the index
    conventions), the limiter on linear data and extrema, the trace with zero
    slopes, fluxes of a fluid at rest, mirror symmetry of the Riemann solver,
    the dx scaling and transposition invariance of dt.
"""
import hashlib
import json
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import oracle
from oracle import Params

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def port():
    return oracle.Port()


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def same(a, b):
    """Bitwise equality (NaN payloads and signed zeros included)."""
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.shape == b.shape and a.tobytes() == b.tobytes()


# ---- 1. golden vectors ---------------------------------------------------------

def test_manifest_intact():
    man = json.load(open(os.path.join(GOLD, "MANIFEST.json")))
    for rel, h in man.items():
        data = open(os.path.join(GOLD, rel), "rb").read()
        assert hashlib.sha256(data).hexdigest() == h, rel


def test_stages_match_reference_vectors(port):
    g = np.load(os.path.join(GOLD, "stages.npz"))
    dtdx = float(g["dtdx"])
    nface = int(g["nface"])
    for k, u in enumerate(g["u"]):
        q, e = port.constoprim(u, 1e-10)
        assert same(q, g["constoprim_q"][k]) and same(e, g["constoprim_e"][k]), k
        q2, c = port.eos(e, q, 1e-10, 1.4)
        assert same(q2, g["eos_q"][k]) and same(c, g["eos_c"][k]), k
        for st in (1.0, 0.5, 2.0):
            dq = port.slope(q2, st)
            # the reference never writes the two end cells; compare where defined
            assert same(dq[:, 1:-1], g[f"slope_{st}"][k][:, 1:-1]), (k, st)
        dq = port.slope(q2, 1.0)
        for sc in (1, 2, 3):
            qxm, qxp = port.trace(q2, dq, c, dtdx, sc)
            assert same(qxm[:, 1:-1], g[f"trace{sc}_qxm"][k][:, 1:-1]), (k, sc)
            assert same(qxp[:, 1:-1], g[f"trace{sc}_qxp"][k][:, 1:-1]), (k, sc)
        qxm, qxp = port.trace(q2, dq, c, dtdx, 1)
        ql, qr = port.qleftright(nface, qxm, qxp)
        assert same(ql[:, :nface], g["ql"][k][:, :nface]) and same(qr[:, :nface],
                                                                   g["qr"][k][:, :nface])
        for gam, tag in ((1.4, "14"), (5.0 / 3.0, "53")):
            for it in (10, 1, 0):
                qg, sg = port.riemann(ql, qr, 1e-10, 1e-10, gam, it, nface=nface)
                assert same(qg[:, :nface], g[f"riemann{tag}_{it}_qg"][k][:, :nface]), (k, tag, it)
                assert same(sg[:nface], g[f"riemann{tag}_{it}_sgnm"][k][:nface]), (k, tag, it)
        qg, _ = port.riemann(ql, qr, 1e-10, 1e-10, 1.4, 10, nface=nface)
        fl = port.cmpflx(qg, 1.4, nface=nface)
        assert same(fl[:, :nface], g["cmpflx"][k][:, :nface]), k


TRAJ = sorted(f[5:-4] for f in os.listdir(GOLD) if f.startswith("traj_"))


@pytest.mark.parametrize("name", TRAJ)
def test_trajectory_matches_reference(port, name):
    g = np.load(os.path.join(GOLD, f"traj_{name}.npz"))
    p = Params(**{k: v for k, v in json.loads(str(g["params"])).items()})
    u = g["uold0"].copy()
    state = None
    for s in range(len(g["dt"])):
        dts, state = port.run(p, u, 1, state)
        assert len(dts) == 1 and dts[0] == g["dt"][s], (name, s)
        assert state[0] == g["t"][s], (name, s)
        assert sha(u) == g["hashes"][s], f"{name}: uold differs after step {s + 1}"
    assert same(u, g["uold"])


# ---- 2. the live reference -----------------------------------------------------

needs_ref = pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built")


@needs_ref
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_states_against_live_reference(port, seed):
    rng = np.random.default_rng(seed)
    nx, ny = int(rng.integers(5, 70)), int(rng.integers(5, 70))
    types = [int(x) for x in rng.integers(0, 4, 4)]
    p = Params(nx=nx, ny=ny, dx=float(rng.uniform(0.01, 1)), courant_factor=0.8, tend=1e30,
               boundary_left=types[0], boundary_right=types[1], boundary_down=types[2],
               boundary_up=types[3], slope_type=float(rng.choice([0.5, 1.0, 1.5])),
               niter_riemann=int(rng.integers(0, 12)))
    u = port.init(p)
    u[0] = rng.uniform(0.1, 3, u[0].shape)
    u[1] = rng.uniform(-1, 1, u[1].shape)
    u[2] = rng.uniform(-1, 1, u[2].shape)
    u[3] = rng.uniform(1, 5, u[3].shape)
    gam = float(rng.uniform(1.3, 1.67))
    p.gamma = gam
    ref = oracle.Ref()
    ref.open(p, gamma=gam)
    try:
        ref.set_uold(u)
        assert ref.compute_deltat() == port.compute_deltat(p, u)
        for idim, dt in ((1, 1e-3), (2, 7e-4), (2, 1e-3)):
            ref.make_boundary(3 - idim)
            port.make_boundary(3 - idim, p, u)
            assert same(ref.get_uold(), u)
            ref.godunov(idim, dt)
            port.godunov(idim, dt, p, u)
            assert same(ref.get_uold(), u), (seed, idim)
    finally:
        ref.close()


@needs_ref
def test_time_loop_against_reference_executable(port):
    """The unmodified HydroC99_2DMpi executable's printed step lines (t, dt to
    6 digits, main.c:349) match the oracle's loop (main.c:244-298 restated)."""
    p = Params(nx=60, ny=40, dx=1 / 60.0, courant_factor=0.8, tend=1e30, nstepmax=30,
               testcase=1)
    with tempfile.TemporaryDirectory() as d:
        deck = oracle.write_deck(p, os.path.join(d, "d.nml"))
        out = subprocess.run([oracle.bindings.REF_D, "-i", deck], cwd=d, capture_output=True,
                             text=True, env=dict(os.environ, OMP_NUM_THREADS="2"),
                             timeout=120).stdout
    lines = re.findall(r"--> step=\s*(\d+),\s*(\S+),\s*(\S+)", out)
    u = port.init(p)
    dts, _ = port.run(p, u, 30)
    t = np.cumsum(dts)
    assert len(lines) == 30
    for (n, ts, dts_s), tt, dd in zip(lines, t, dts):
        assert ts == "%.5e" % tt and dts_s == "%.5e" % dd, (n, ts, dts_s, tt, dd)


# ---- 3. properties fixed by the mathematics ------------------------------------

def transpose_state(u):
    """Swap the x and y axes: density and energy transpose, momenta swap slots."""
    return np.ascontiguousarray(np.stack([u[0].T, u[2].T, u[1].T, u[3].T]))


@pytest.mark.parametrize("btype", [1, 3])
def test_split_step_transposition_symmetry(port, btype):
    """An x-then-y step on S equals, transposed, a y-then-x step on S^T, exactly.
    Pins the y-sweep's IU/IV swap and both sweeps' index conventions."""
    rng = np.random.default_rng(5)
    p = Params(nx=41, ny=23, dx=0.03, courant_factor=0.8, tend=1e30, boundary_left=btype,
               boundary_right=btype, boundary_down=btype, boundary_up=btype)
    pt = Params(nx=23, ny=41, dx=0.03, courant_factor=0.8, tend=1e30, boundary_left=btype,
                boundary_right=btype, boundary_down=btype, boundary_up=btype)
    u = port.init(p)
    u[0, 2:-2, 2:-2] = rng.uniform(0.5, 2, (23, 41))
    u[1, 2:-2, 2:-2] = rng.uniform(-1, 1, (23, 41))
    u[2, 2:-2, 2:-2] = rng.uniform(-1, 1, (23, 41))
    u[3, 2:-2, 2:-2] = rng.uniform(2, 4, (23, 41))
    ut = transpose_state(u)
    assert port.compute_deltat(p, u) == port.compute_deltat(pt, ut)
    port.godunov(1, 2e-3, p, u)
    port.godunov(2, 2e-3, pt, ut)
    a = transpose_state(u)[:, 2:-2, 2:-2]
    b = ut[:, 2:-2, 2:-2]
    assert same(a, b)


def test_conservation_reflective_walls(port):
    """Mass and total energy are conserved to round-off with reflective walls."""
    p = Params(nx=250, ny=125, dx=1 / 250.0, courant_factor=0.8, tend=1e30, testcase=1)
    u = port.init(p)
    m0, e0 = u[0, 2:-2, 2:-2].sum(), u[3, 2:-2, 2:-2].sum()
    port.run(p, u, 100)
    m1, e1 = u[0, 2:-2, 2:-2].sum(), u[3, 2:-2, 2:-2].sum()
    assert abs(m1 - m0) / m0 < 1e-13
    assert abs(e1 - e0) / e0 < 1e-13
    assert (u[0, 2:-2, 2:-2] > 0).all()


def test_slope_limiter_cases(port):
    n = 16
    x = np.arange(n, dtype=float)
    q = np.stack([3 + 0.5 * x, -2 * x, 0 * x, 7 + x])
    dq = port.slope(q, 1.0)
    # values (a flat variable gives -0.0: the sign is -1 unless dc > 0)
    assert np.array_equal(dq[:, 1:-1], np.stack([np.full(n - 2, 0.5), np.full(n - 2, -2.0),
                                                np.zeros(n - 2), np.ones(n - 2)]))
    peak = np.stack([np.minimum(x, 15 - x)] * 4)  # extremum -> zero slope there
    assert (port.slope(peak, 1.0)[:, 7:9] == 0).all()
    assert np.array_equal(port.slope(-q, 1.0)[:, 1:-1], -dq[:, 1:-1])


def test_trace_with_zero_slopes_is_identity(port):
    rng = np.random.default_rng(3)
    n = 20
    q = np.stack([rng.uniform(0.5, 2, n), rng.uniform(-1, 1, n), rng.uniform(-1, 1, n),
                  rng.uniform(0.5, 2, n)])
    c = np.sqrt(1.4 * q[3] / q[0])
    qxm, qxp = port.trace(q, np.zeros_like(q), c, 0.3, 1)
    assert same(qxm[:, 1:-1], q[:, 1:-1]) and same(qxp[:, 1:-1], q[:, 1:-1])


def test_fluxes_of_fluid_at_rest(port):
    n = 8
    qg = np.stack([np.full(n, 1.3), np.zeros(n), np.zeros(n), np.full(n, 0.7)])
    f = port.cmpflx(qg, 1.4)
    assert same(f, np.stack([np.zeros(n), np.full(n, 0.7), np.zeros(n), np.zeros(n)]))
    # equal states at rest on both sides: the Godunov state is that state
    qg2, _ = port.riemann(qg, qg, 1e-10, 1e-10, 1.4, 10)
    assert same(qg2[[0, 3]], qg[[0, 3]])


def test_riemann_mirror_symmetry(port):
    """Swapping the states and negating normal velocities mirrors the solution."""
    rng = np.random.default_rng(11)
    n = 200
    ql = np.stack([rng.uniform(0.1, 2, n), rng.uniform(-1, 1, n), rng.uniform(-1, 1, n),
                   rng.uniform(0.1, 2, n)])
    qr = np.stack([rng.uniform(0.1, 2, n), rng.uniform(-1, 1, n), rng.uniform(-1, 1, n),
                   rng.uniform(0.1, 2, n)])
    mirror = np.array([1, -1, 1, 1])[:, None]
    g1, s1 = port.riemann(ql, qr, 1e-10, 1e-10, 1.4, 10)
    g2, s2 = port.riemann(qr * mirror, ql * mirror, 1e-10, 1e-10, 1.4, 10)
    # ustar = 0 exactly breaks the tie differently on the two sides; skip those
    ok = s1 == -s2
    assert ok.mean() > 0.99
    np.testing.assert_allclose(g2[0][ok], g1[0][ok], rtol=1e-12)
    np.testing.assert_allclose(g2[1][ok], -g1[1][ok], rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(g2[3][ok], g1[3][ok], rtol=1e-12)


def test_dt_scaling_and_transposition(port):
    rng = np.random.default_rng(2)
    p = Params(nx=30, ny=20, dx=0.01, courant_factor=0.8, tend=1e30)
    u = port.init(p)
    u[:, 2:-2, 2:-2] = np.stack([rng.uniform(0.5, 2, (20, 30)), rng.uniform(-1, 1, (20, 30)),
                                 rng.uniform(-1, 1, (20, 30)), rng.uniform(2, 3, (20, 30))])
    dt = port.compute_deltat(p, u)
    p2 = Params(nx=30, ny=20, dx=0.02, courant_factor=0.8, tend=1e30)
    assert port.compute_deltat(p2, u) == 2 * dt
    pt = Params(nx=20, ny=30, dx=0.01, courant_factor=0.8, tend=1e30)
    assert port.compute_deltat(pt, transpose_state(u)) == dt
    # brute force on a single cell: dt = cfl dx / (c + |u|)
    p1 = Params(nx=1, ny=1, dx=0.5, courant_factor=0.8, tend=1e30)
    u1 = port.init(p1)
    u1[:, 2, 2] = [2.0, 2.0 * 0.75, 0.0, 2.5 / 0.4 + 0.5 * 2.0 * 0.75 ** 2]
    c = np.sqrt(1.4 * 2.5 / 2.0)
    np.testing.assert_allclose(port.compute_deltat(p1, u1), 0.8 * 0.5 / (c + 0.75), rtol=1e-14)


def test_blast_energy_placement(port):
    p = Params(nx=250, ny=125, dx=1 / 250.0)
    for tc, (x, y) in ((0, (127, 64)), (1, (2, 2))):
        u = port.init(p, testcase=tc)
        assert u[3, y, x] == (1 / p.dx) / p.dx
        assert np.count_nonzero(u[3] > 1) == 1
