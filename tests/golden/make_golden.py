#!/usr/bin/env python3
"""Writes the golden vectors under tests/golden/ from the reference itself.

Run in a container that holds the reference sources (they are compiled in
place into oracle/_ref by `make -C oracle`); the vectors are data only and are
committed, the reference never travels.

    python tests/golden/make_golden.py            (everything)
    python tests/golden/make_golden.py images     (images.npz only)

Files written:
  stages.npz      per-stage inputs/outputs of the reference's own constoprim,
                  equation_of_state, slope, trace, qleftright, riemann and
                  cmpflx (HydroC99_2DMpi/Src/*.c) on seeded random and
                  edge-case lines
  traj_*.npz      time-loop runs of the reference (dt and t per step, sha256 of
                  the full uold after every step, final uold)
  vtk_ref/        the reference vtkfile() output for a small state
  images.npz      PPM density images written by the reference OpenMP
                  executable (HydroC99_2Domp, built in place as
                  oracle/_ref/hydro_omp) for three decks, with the decks and
                  the step -> image markers of its step lines
  decks/          copies of the reference's input decks (data files)
  MANIFEST.json   file list with sha256
"""
import hashlib
import json
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from oracle import Params  # noqa: E402

REF_INPUT = "/root/reference/Input"
REF_DECK6000 = "/root/reference/HydroC/input6000.nml"


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


# ---------------------------------------------------------------------------
# stage vectors

def random_lines(rng, n, count):
    """Conserved-variable lines [count][4][n]: smooth, noisy, jumps, edge cases."""
    out = []
    for k in range(count):
        kind = k % 6
        x = np.linspace(0, 1, n)
        if kind == 0:  # smooth waves
            rho = 1 + 0.5 * np.sin(2 * np.pi * x * rng.uniform(1, 4))
            u = rng.uniform(-2, 2) * np.cos(2 * np.pi * x)
            v = rng.uniform(-1, 1) * np.ones(n)
            p = 1 + 0.3 * np.cos(3 * np.pi * x)
        elif kind == 1:  # noise
            rho = rng.uniform(0.05, 10, n)
            u = rng.uniform(-5, 5, n)
            v = rng.uniform(-5, 5, n)
            p = rng.uniform(1e-3, 100, n)
        elif kind == 2:  # Sod-like jump
            m = rng.integers(10, n - 10)
            rho = np.where(np.arange(n) < m, 1.0, 0.125)
            u = np.zeros(n)
            v = np.zeros(n)
            p = np.where(np.arange(n) < m, 1.0, 0.1)
        elif kind == 3:  # strong blast in a cold background
            rho = np.ones(n)
            u = np.zeros(n)
            v = np.zeros(n)
            p = np.full(n, 1e-5 * 0.4)
            p[rng.integers(5, n - 5)] = rng.uniform(1e3, 1e8)
        elif kind == 4:  # vacuum / floors: rho below smallr, negative internal energy
            rho = rng.uniform(0, 2, n)
            rho[::7] = 0.0
            rho[3::11] = 1e-12
            u = rng.uniform(-3, 3, n)
            v = rng.uniform(-3, 3, n)
            p = rng.uniform(-0.5, 2, n)
        else:  # uniform state (all slopes zero) with signed zeros
            rho = np.full(n, rng.uniform(0.5, 2))
            u = np.zeros(n)
            u[::2] = -0.0
            v = np.zeros(n)
            p = np.full(n, rng.uniform(0.1, 2))
        gm1 = 0.4
        E = p / gm1 + 0.5 * rho * (u * u + v * v)
        out.append(np.stack([rho, rho * u, rho * v, E]))
    return np.array(out)


def stage_vectors(ref):
    rng = np.random.default_rng(20261202)
    n = 64
    lines = random_lines(rng, n, 36)
    d = {"u": lines}
    q_all, e_all, c_all, qe_all = [], [], [], []
    dq_all = {st: [] for st in (1.0, 0.5, 2.0)}
    tr_all = {sc: ([], []) for sc in (1, 2, 3)}
    lr_all = ([], [])
    rm_all = {(g, it): ([], []) for g in (1.4, 5.0 / 3.0) for it in (10, 1, 0)}
    fl_all = []
    dtdx = 0.37
    nface = n - 3
    for u in lines:
        q, e = ref.constoprim(u, 1e-10)
        q_all.append(q); e_all.append(e)
        q2, c = ref.eos(e, q, 1e-10, 1.4)
        qe_all.append(q2); c_all.append(c)
        for st in dq_all:
            dq_all[st].append(ref.slope(q2, st))
        dq = dq_all[1.0][-1]
        for sc in tr_all:
            qxm, qxp = ref.trace(q2, dq, c, dtdx, sc)
            tr_all[sc][0].append(qxm); tr_all[sc][1].append(qxp)
        qxm, qxp = tr_all[1][0][-1], tr_all[1][1][-1]
        ql, qr = ref.qleftright(nface, qxm, qxp)
        lr_all[0].append(ql); lr_all[1].append(qr)
        for (g, it) in rm_all:
            qg, sg = ref.riemann(ql, qr, 1e-10, 1e-10, g, it, nface=nface)
            rm_all[(g, it)][0].append(qg); rm_all[(g, it)][1].append(sg)
        qg = rm_all[(1.4, 10)][0][-1]
        fl_all.append(ref.cmpflx(qg, 1.4, nface=nface))
    d["constoprim_q"] = np.array(q_all)
    d["constoprim_e"] = np.array(e_all)
    d["eos_q"] = np.array(qe_all)
    d["eos_c"] = np.array(c_all)
    for st, v in dq_all.items():
        d[f"slope_{st}"] = np.array(v)
    for sc, (m, p) in tr_all.items():
        d[f"trace{sc}_qxm"] = np.array(m)
        d[f"trace{sc}_qxp"] = np.array(p)
    d["ql"] = np.array(lr_all[0])
    d["qr"] = np.array(lr_all[1])
    for (g, it), (qg, sg) in rm_all.items():
        tag = f"{'14' if g == 1.4 else '53'}_{it}"
        d[f"riemann{tag}_qg"] = np.array(qg)
        d[f"riemann{tag}_sgnm"] = np.array(sg)
    d["cmpflx"] = np.array(fl_all)
    d["dtdx"] = np.array(dtdx)
    d["nface"] = np.array(nface)
    np.savez_compressed(os.path.join(HERE, "stages.npz"), **d)


# ---------------------------------------------------------------------------
# trajectories

def trajectory(ref, name, p: Params, steps, gamma=None, uold=None):
    ref.open(p, gamma=gamma)
    try:
        if uold is not None:
            ref.set_uold(uold)
        u0 = ref.get_uold()
        dts, ts, hashes = [], [], []
        for _ in range(steps):
            dt, t = ref.steps(1)
            if len(dt) == 0:
                break
            dts.append(dt[0]); ts.append(t[0])
            hashes.append(sha(ref.get_uold()))
        final = ref.get_uold()
    finally:
        ref.close()
    np.savez_compressed(os.path.join(HERE, f"traj_{name}.npz"), dt=np.array(dts),
                        t=np.array(ts), hashes=np.array(hashes), uold0=u0, uold=final,
                        params=json.dumps({**p.__dict__, "gamma": gamma or p.gamma}))


def blast_state(p: Params, bx, by, e):
    port = oracle.Port()
    return port.init_blast(p, bx, by, e)


# ---------------------------------------------------------------------------
# images of the reference OpenMP executable

IMAGE_DECKS = {
    # 250x125 corner blast, an image every 10 steps
    "corner_n10": dict(nx=250, ny=125, dx=0.004, testcase=1, nstepmax=20, extra="nimage=10"),
    # 1100x60 centred blast: shrink 2 (550x30 pixels), an image every 2 steps
    "shrink2_n2": dict(nx=1100, ny=60, dx=1.0 / 1100, testcase=0, nstepmax=4, extra="nimage=2"),
    # 250x125 corner blast, an image each time t passes a multiple of 1e-4
    "corner_dt": dict(nx=250, ny=125, dx=0.004, testcase=1, nstepmax=20, extra="dtimage=1e-4"),
}


def image_deck(d):
    return ("&RUN\ntend=1e30\nnstepmax=%d\n%s\n/\n&MESH\nnx=%d\nny=%d\ndx=%r\n"
            "boundary_left=1\nboundary_right=1\nboundary_down=1\nboundary_up=1\ntestcase=%d\n/\n"
            "&HYDRO\ncourant_factor=0.8\nniter_riemann=10\niorder=2\nslope_type=1.\n"
            "scheme=muscl\nsmallr=1e-10\nsmallc=1e-10\n/\n"
            % (d["nstepmax"], d["extra"], d["nx"], d["ny"], d["dx"], d["testcase"]))


def images():
    import re
    import subprocess
    import tempfile
    exe = os.path.join(ROOT, "oracle", "_ref", "hydro_omp")
    out = {}
    for name, d in IMAGE_DECKS.items():
        deck = image_deck(d)
        with tempfile.TemporaryDirectory() as tmp:
            with open(os.path.join(tmp, "deck.nml"), "w") as f:
                f.write(deck)
            log = subprocess.run([exe, "-i", "deck.nml"], cwd=tmp, capture_output=True, text=True,
                                 env=dict(os.environ, OMP_NUM_THREADS="4"), check=True).stdout
            files = sorted(f for f in os.listdir(tmp) if f.startswith("IMAGE_"))
            for f in files:
                out[f"{name}/{f}"] = np.frombuffer(open(os.path.join(tmp, f), "rb").read(), np.uint8)
        marks = [(int(a), int(b)) for a, b in re.findall(r"--> step=\s*(\d+),.*<(\d+)>", log)]
        out[f"{name}/deck"] = np.frombuffer(deck.encode(), np.uint8)
        out[f"{name}/marks"] = np.array(marks, dtype=np.int64).reshape(-1, 2)
    np.savez_compressed(os.path.join(HERE, "images.npz"), **out)


def manifest():
    man = {}
    for dp, _, fs in os.walk(HERE):
        for f in fs:
            if f.endswith((".py", ".pyc")) or f == "MANIFEST.json":
                continue
            path = os.path.join(dp, f)
            man[os.path.relpath(path, HERE)] = hashlib.sha256(open(path, "rb").read()).hexdigest()
    with open(os.path.join(HERE, "MANIFEST.json"), "w") as fh:
        json.dump(dict(sorted(man.items())), fh, indent=1)
    print("golden vectors written:", len(man), "files")


def main():
    if sys.argv[1:] == ["images"]:
        oracle.build(ref=True)
        images()
        manifest()
        return
    oracle.build(ref=True)
    ref = oracle.Ref()
    stage_vectors(ref)
    # 250x125 corner blast, dx = 1/250, CFL 0.8, 100 steps
    corner = Params(nx=250, ny=125, dx=1.0 / 250, courant_factor=0.8, tend=1e30, nstepmax=100,
                testcase=1, nxystep=125)
    trajectory(ref, "corner250x125", corner, 100)
    # the reference's own deck Input/input_250x125_corner.nml (dx 0.05), 100 steps
    deck = Params(nx=250, ny=125, dx=0.05, courant_factor=0.8, tend=200.0, nstepmax=100,
                  testcase=1, nxystep=125)
    trajectory(ref, "deck250", deck, 100)
    # centred blast, odd sizes, 60 steps
    trajectory(ref, "center97x70", Params(nx=97, ny=70, dx=1 / 97.0, courant_factor=0.8,
                                          tend=1e30, testcase=0), 60)
    # Sod column (testcase 2), 40 steps
    trajectory(ref, "sod64x16", Params(nx=64, ny=16, dx=1 / 64.0, courant_factor=0.8,
                                       tend=1e30, testcase=2), 40)
    # boundary types and scheme variants on a small grid, 40 steps each
    variants = {
        "absorb": dict(boundary_left=2, boundary_right=2, boundary_down=2, boundary_up=2),
        "periodic": dict(boundary_left=3, boundary_right=3, boundary_down=3, boundary_up=3),
        "mixed": dict(boundary_left=1, boundary_right=2, boundary_down=3, boundary_up=0),
        "plmde_o1": dict(scheme=2, iorder=1),
        "collela_st2": dict(scheme=3, slope_type=2.0),
        "newton1": dict(niter_riemann=1),
    }
    for name, kw in variants.items():
        trajectory(ref, name, Params(nx=24, ny=17, dx=1 / 24.0, courant_factor=0.8, tend=1e30,
                                     testcase=0, **kw), 40)
    # instances like the batched and serving workloads' (random gamma, blast
    # cell and energy); the reference runs with gamma set on its struct
    rng = np.random.default_rng(20261202)
    for k in range(3):
        n = int(rng.integers(20, 48))
        g = float(rng.uniform(1.3, 1.67))
        bx, by = int(rng.integers(0, n)), int(rng.integers(0, n))
        e = float(rng.uniform(0.1, 10)) * n * n
        p = Params(nx=n, ny=n, dx=1.0 / n, courant_factor=0.8, tend=1e30, testcase=1, gamma=g)
        trajectory(ref, f"blast{k}", p, 60, gamma=g, uold=blast_state(p, bx, by, e))
    # VTK output of the reference for a small state
    vdir = os.path.join(HERE, "vtk_ref")
    shutil.rmtree(vdir, ignore_errors=True)
    os.makedirs(vdir)
    cwd = os.getcwd()
    try:
        os.chdir(vdir)
        ref.open(Params(nx=10, ny=6, dx=0.1, courant_factor=0.8, tend=1e30, testcase=0))
        ref.steps(5)
        ref.lib.ref_vtkfile(1)
        ref.steps(2)
        ref.lib.ref_vtkfile(2)
        np.save("uold_step7.npy", ref.get_uold())
        ref.close()
    finally:
        os.chdir(cwd)
    # the reference's input decks (data files)
    ddir = os.path.join(HERE, "decks")
    os.makedirs(ddir, exist_ok=True)
    for f in sorted(os.listdir(REF_INPUT)):
        shutil.copy(os.path.join(REF_INPUT, f), os.path.join(ddir, f))
    shutil.copy(REF_DECK6000, os.path.join(ddir, "input6000.nml"))
    images()
    manifest()


if __name__ == "__main__":
    main()
