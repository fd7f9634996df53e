"""The C ABI boundary on the CPU: exports, the .nml parser, the VTK dump.

No GPU calls: process_args and vtkfile are C99 host code in libhydro_amd.so.
"""
import ctypes as C
import filecmp
import os
import re
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

import hydro_amd
import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
INCLUDE = os.path.join(ROOT, "include")


def declared_functions():
    names = set()
    for f in os.listdir(INCLUDE):
        if not f.endswith(".h"):
            continue
        text = open(os.path.join(INCLUDE, f)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*[A-Za-z_][\w\s\*]*?\b(\w+)\s*\([^;{]*\)\s*;", text, re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_function():
    names = declared_functions()
    assert {"hydro_init", "hydro_godunov", "compute_deltat", "make_boundary", "vtkfile",
            "process_args", "hydro_amd_grid_run"} <= names
    out = subprocess.run(["nm", "-D", "--defined-only", hydro_amd.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = sorted(names - exported)
    assert not missing, missing
    for sym in ("functim", "flopsAri", "flopsSqr", "flopsMin", "flopsTra", "MflopsSUM",
                "nbFLOPS"):
        assert sym in exported, sym


def test_library_loads_and_binds():
    L = hydro_amd.lib()
    for n in declared_functions():
        assert hasattr(L, n), n


def test_hydroparam_layout_matches_reference_header():
    """sizeof / offsets of hydroparam_t as a C compiler lays out the reference
    struct (HydroC99_2DMpi/Src/parametres.h:55-95) and our header."""
    src = r'''
#include <stddef.h>
#include <stdio.h>
#include "hydro_surface.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu\n", sizeof(hydroparam_t), offsetof(hydroparam_t, box),
         offsetof(hydroparam_t, dx), offsetof(hydroparam_t, slope_type),
         offsetof(hydroparam_t, testCase), sizeof(hydrowork_t));
  return 0;
}'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "l.c")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I", INCLUDE, "-o", os.path.join(d, "l"), c], check=True)
        ours = subprocess.run([os.path.join(d, "l")], capture_output=True, text=True).stdout
        refdir = "/root/reference/HydroC/HydroC99_2DMpi/Src"
        if os.path.isdir(refdir):
            open(c, "w").write(src.replace('#include "hydro_surface.h"', '#include "parametres.h"'))
            subprocess.run(["gcc", "-I", refdir, "-o", os.path.join(d, "r"), c], check=True)
            theirs = subprocess.run([os.path.join(d, "r")], capture_output=True,
                                    text=True).stdout
            assert ours == theirs
    H = hydro_amd.HydroParam
    assert C.sizeof(H) == int(ours.split()[0])
    assert H.box.offset == int(ours.split()[1]) and H.dx.offset == int(ours.split()[2])
    assert H.slope_type.offset == int(ours.split()[3])
    assert H.testCase.offset == int(ours.split()[4])


DECKS = sorted(f for f in os.listdir(os.path.join(GOLD, "decks")) if f.endswith(".nml"))


@pytest.mark.parametrize("deck", DECKS)
def test_process_args_on_reference_decks(deck, capfd):
    path = os.path.join(GOLD, "decks", deck)
    H = hydro_amd.process_args(["-i", path])
    capfd.readouterr()
    if oracle.ref_available():
        # the reference's own process_args, through the harness (parse only)
        ref = oracle.Ref()
        r = np.zeros(9)
        iv = np.zeros(14, dtype=np.int32)
        ref.lib.ref_parse(path.encode(), r.ctypes.data_as(C.c_void_p),
                          iv.ctypes.data_as(C.c_void_p))
        capfd.readouterr()
        assert (H.tend, H.dx, H.gamma, H.courant_factor, H.smallc, H.smallr, H.slope_type,
                H.dtoutput) == tuple(r[1:9])
        assert (H.nx, H.ny, H.nxystep, H.nstepmax, H.testCase, H.niter_riemann, H.iorder,
                H.scheme, H.boundary_left, H.boundary_right, H.boundary_down, H.boundary_up,
                H.noutput, H.nvar) == tuple(int(x) for x in iv)
    # values read straight from the deck text
    text = open(path).read()
    for key, field in (("nx", "nx"), ("ny", "ny"), ("courant_factor", "courant_factor"),
                       ("testcase", "testCase")):
        m = re.search(rf"^\s*{key}\s*=\s*([-0-9.eE]+)", text, re.M)
        if m:
            assert getattr(H, field) == type(getattr(H, field))(float(m.group(1)))


def test_process_args_gamma_key_and_defaults(tmp_path, capfd):
    d = tmp_path / "g.nml"
    d.write_text("&MESH\nnx=7\n  ny = 9\n/\n&HYDRO\ngamma=1.6666\nscheme=plmde\n/\n")
    H = hydro_amd.process_args(["-i", str(d)])
    capfd.readouterr()
    assert (H.nx, H.ny, H.gamma, H.scheme) == (7, 9, 1.6666, 2)
    assert (H.courant_factor, H.niter_riemann, H.iorder, H.slope_type) == (0.5, 10, 2, 1.0)
    assert H.nxystep == 7 and H.nstepmax == 2**31 - 1 and H.nvar == 4
    assert tuple(H.box[:4]) == (0, 7, 0, 9)


def test_vtkfile_matches_reference_bytes(tmp_path, capfd):
    """Our vtkfile() writes the reference's files byte for byte (same state)."""
    src = os.path.join(GOLD, "vtk_ref")
    u = np.load(os.path.join(src, "uold_step7.npy"))
    p = oracle.Params(nx=10, ny=6, dx=0.1, courant_factor=0.8, tend=1e30, testcase=0)
    deck = oracle.write_deck(p, str(tmp_path / "v.nml"))
    H = hydro_amd.process_args(["-i", deck])
    H.nxt, H.nyt = 14, 10
    H.imin, H.jmin, H.imax, H.jmax = 0, 0, 14, 10
    cwd = os.getcwd()
    try:
        os.chdir(tmp_path)
        hydro_amd.vtkfile(1, H, u)  # step 1 directory/pvd shape
        hydro_amd.vtkfile(2, H, u)  # state of step 7 = the reference's dump #2
    finally:
        os.chdir(cwd)
    capfd.readouterr()
    for rel in ("Dep/0000/02/Hydro_00000_0002.vtr", "Dep/0000/02/Hydro_0002.pvtr",
                "Hydro.pvd", "Dep/0000/01/Hydro_0001.pvtr"):
        assert filecmp.cmp(tmp_path / rel, os.path.join(src, rel), shallow=False), rel
    shutil.rmtree(tmp_path / "Dep")
