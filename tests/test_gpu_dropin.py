"""Drop-in at the call surface, end to end on the GPU.

Three executables run the same deck:
  oracle/_ref/hydro_d       the reference HydroC99_2DMpi, CPU, unmodified
  oracle/_ref/hydro_dropin  the reference's own main.c linked against
                            libhydro_amd.so (the GPU) instead of its numerics
  hydro_amd/bin/hydro       this package's C99 driver (host loop and -native)
Their step lines (t, dt) and every VTK file they write must be identical.
"""
import filecmp
import os
import re
import subprocess

import pytest

import oracle
from oracle import bindings

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HYDRO = os.path.join(ROOT, "hydro_amd", "bin", "hydro")
DROPIN = os.path.join(ROOT, "oracle", "_ref", "hydro_dropin")

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not (os.path.exists(bindings.REF_D) and os.path.exists(DROPIN)),
                                 reason="oracle/_ref executables not built")]


def run(exe, deck, cwd, *extra):
    out = subprocess.run([exe, "-i", deck, *extra], cwd=cwd, capture_output=True, text=True,
                         timeout=300, env=dict(os.environ, OMP_NUM_THREADS="4"))
    assert out.returncode == 0, out.stderr[-2000:]
    return re.findall(r"--> step=\s*(\d+),\s*(\S+),\s*(\S+)", out.stdout)


def vtk_files(d):
    out = []
    for dp, _, fs in os.walk(d):
        for f in fs:
            if f.endswith((".vtr", ".pvtr", ".pvd")):
                out.append(os.path.relpath(os.path.join(dp, f), d))
    return sorted(out)


@pytest.mark.parametrize("testcase", [0, 1, 2])
def test_reference_driver_on_gpu_library(tmp_path, testcase):
    p = oracle.Params(nx=120, ny=70, dx=1 / 120.0, courant_factor=0.8, tend=1e30, nstepmax=40,
                      testcase=testcase)
    deck = oracle.write_deck(p, str(tmp_path / "deck.nml"))
    with open(deck, "a") as f:
        f.write("&RUN\nnoutput=20\n/\n")
    dirs = {k: tmp_path / k for k in ("ref", "dropin", "ours", "native")}
    for d in dirs.values():
        d.mkdir()
    ref = run(bindings.REF_D, deck, dirs["ref"])
    assert len(ref) == 40
    assert run(DROPIN, deck, dirs["dropin"]) == ref
    assert run(HYDRO, deck, dirs["ours"]) == ref
    assert run(HYDRO, deck, dirs["native"], "-native", "-chunk", "7") == ref
    files = vtk_files(dirs["ref"])
    assert len(files) == 2 * 3 + 1  # dumps 1..3: .vtr + .pvtr each, and Hydro.pvd
    for k in ("dropin", "ours", "native"):
        assert vtk_files(dirs[k]) == files, k
        for f in files:
            assert filecmp.cmp(dirs["ref"] / f, dirs[k] / f, shallow=False), (k, f)


def run_env(exe, deck, cwd, env, *extra):
    out = subprocess.run([exe, "-i", deck, *extra], cwd=cwd, capture_output=True, text=True,
                         timeout=300, env=dict(os.environ, OMP_NUM_THREADS="4", **env))
    assert out.returncode == 0, out.stderr[-2000:]
    return re.findall(r"--> step=\s*(\d+),\s*(\S+),\s*(\S+)", out.stdout), out


@pytest.mark.parametrize("mode", [(), ("-native", "-chunk", "3")])
def test_time_guard_checkpoint_restart(tmp_path, mode):
    """The time guard stops a run (forced from the host: HYDROC_START_TIME
    leaves 15 s of a 1000 s allotment, below the 20 s guard), the checkpoint
    is written, and the restarted run ends bit-identical to an uninterrupted
    one, with the same step lines from the restart on."""
    import time
    import numpy as np
    p = oracle.Params(nx=96, ny=64, dx=1 / 96.0, courant_factor=0.8, tend=1e30, nstepmax=25,
                      testcase=1)
    deck = oracle.write_deck(p, str(tmp_path / "deck.nml"))
    full, _ = run_env(HYDRO, deck, tmp_path, {}, *mode, "-dump", "full.bin")
    assert len(full) == 25
    env = {"HYDROC_MAXTIME": "1000", "HYDROC_START_TIME": str(int(time.time()) - 985)}
    first, out = run_env(HYDRO, deck, tmp_path, env, *mode, "-checkpoint", "ck.bin")
    assert "stops by time limit" in out.stderr
    k = len(first)
    assert 1 <= k < 25 and first == full[:k]
    # a second checkpoint keeps the first as .prev
    second, _ = run_env(HYDRO, deck, tmp_path, {}, *mode, "-restart", "ck.bin",
                        "-checkpoint", "ck.bin", "-dump", "rest.bin")
    assert os.path.exists(tmp_path / "ck.bin.prev")
    assert second == full[k:]
    a = np.fromfile(tmp_path / "full.bin", np.uint8)
    b = np.fromfile(tmp_path / "rest.bin", np.uint8)
    assert np.array_equal(a, b)
