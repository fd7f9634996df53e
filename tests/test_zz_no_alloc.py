"""The allocation-free compute_deltat (hydro_amd_compute_deltat_noalloc,
scratch in a file-scope device array) against the entry point it varies,
hydro_amd_grid_compute_deltat, and the oracle (compute_deltat,
HydroC99_2DMpi/Src/compute_deltat.c:166-200): the same dt bit for bit, on two
calls in a row, the second on a smaller grid."""
import numpy as np
import pytest

import hydro_amd
import oracle
from helpers import DeviceArray, bits_differ, to_oracle


def _state(nx, ny, bx, by, e, steps):
    p = hydro_amd.Params(nx=nx, ny=ny, dx=1.0 / nx, courant_factor=0.8, tend=1e30)
    with hydro_amd.Grid(p) as g:
        g.init_blast(bx, by, e)
        g.run(steps, log=False)
        u = g.download()
        dt_grid = g.compute_deltat()
    return p, u, dt_grid


@pytest.mark.gpu
def test_compute_deltat_noalloc_matches_grid_entry_point():
    port = oracle.Port()
    cases = [(300, 200, 40, 150, 3e4, 7), (97, 70, 3, 5, 1e3, 4)]  # then a smaller grid
    for nx, ny, bx, by, e, steps in cases:
        print(f"noalloc case {nx}x{ny}, {steps} steps", flush=True)
        p, u, dt_grid = _state(nx, ny, bx, by, e, steps)
        du, ddt = DeviceArray(u.nbytes), DeviceArray(8)
        try:
            du.upload(u)
            ddt.upload(np.array([-1.0]))
            hydro_amd.compute_deltat_noalloc(p, du.ptr, ddt.ptr)
            dt = ddt.download(np.float64, 1)
        finally:
            du.free()
            ddt.free()
        assert bits_differ(dt, np.array([dt_grid])) == 0, (dt[0], dt_grid)
        dt_oracle = port.compute_deltat(to_oracle(p), u)
        assert bits_differ(dt, np.array([dt_oracle])) == 0, (dt[0], dt_oracle)


@pytest.mark.gpu
def test_compute_deltat_noalloc_rejects_bad_arguments():
    d = DeviceArray(8)
    try:
        for bad in (hydro_amd.Params(nx=0, ny=4, dx=1.0),
                    hydro_amd.Params(nx=8, ny=8, dx=1.0, precision=hydro_amd.MIXED),
                    hydro_amd.Params(nx=1 << 20, ny=1 << 21, dx=1.0)):
            with pytest.raises(hydro_amd.HydroError):
                hydro_amd.compute_deltat_noalloc(bad, d.ptr, d.ptr)
    finally:
        d.free()
