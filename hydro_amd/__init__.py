"""hydro_amd -- MI355X-native HydroC99 split Godunov/MUSCL step.

Python mirror of the C ABI in ``include/hydro_amd.h`` (ctypes over the in-tree
``hydro_amd/libhydro_amd.so``).  All arithmetic runs in the HIP kernels of that
library; this module only moves parameters and host arrays.  If the library is
missing, importing the bindings raises -- there is no CPU fallback.

Arrays use the reference layout: ``uold[4, ny+4, nx+4]`` float64 (slot 0 rho,
1 rho*u, 2 rho*v, 3 E, two ghost layers), as in
HydroC99_2DMpi/Src/parametres.h:102-107.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import os
import subprocess

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libhydro_amd.so")

FP64 = 0
MIXED = 1
FP64_FAST = 2  # FP64, not bit for bit: within a per-field relative L1 of 1e-12

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def build(jobs: int = 8) -> None:
    """Compile libhydro_amd.so and bin/hydro in place (hipcc, gfx950)."""
    subprocess.run(["make", "-s", "-C", PKG_DIR, f"-j{jobs}"], check=True)


class HydroError(RuntimeError):
    pass


class _CParams(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("dx", C.c_double), ("gamma", C.c_double),
                ("courant_factor", C.c_double), ("smallc", C.c_double), ("smallr", C.c_double),
                ("slope_type", C.c_double), ("niter_riemann", C.c_int), ("iorder", C.c_int),
                ("scheme", C.c_int), ("boundary_left", C.c_int), ("boundary_right", C.c_int),
                ("boundary_down", C.c_int), ("boundary_up", C.c_int), ("tend", C.c_double),
                ("nstepmax", C.c_int), ("precision", C.c_int)]


@dataclasses.dataclass
class Params:
    """hydro_amd_params; defaults are HydroC99_2DMpi/Src/parametres.c:64-109."""
    nx: int = 20
    ny: int = 20
    dx: float = 1.0
    gamma: float = 1.4
    courant_factor: float = 0.5
    smallc: float = 1e-10
    smallr: float = 1e-10
    slope_type: float = 1.0
    niter_riemann: int = 10
    iorder: int = 2
    scheme: int = 1
    boundary_left: int = 1
    boundary_right: int = 1
    boundary_down: int = 1
    boundary_up: int = 1
    tend: float = 0.0
    nstepmax: int = 2**31 - 1
    precision: int = FP64

    def c(self) -> _CParams:
        return _CParams(*[getattr(self, f[0]) for f in _CParams._fields_])

    def shape(self):
        return (4, self.ny + 4, self.nx + 4)


class HydroParam(C.Structure):
    """hydroparam_t of include/hydro_surface.h (HydroC99_2DMpi/Src/parametres.h:55-95)."""
    _fields_ = [("prt", C.c_int), ("t", C.c_double), ("tend", C.c_double), ("nstep", C.c_int),
                ("nstepmax", C.c_int), ("noutput", C.c_int), ("dtoutput", C.c_double),
                ("imin", C.c_int), ("imax", C.c_int), ("jmin", C.c_int), ("jmax", C.c_int),
                ("nx", C.c_int), ("ny", C.c_int), ("nxt", C.c_int), ("nyt", C.c_int),
                ("nxyt", C.c_int), ("nxystep", C.c_int), ("nproc", C.c_int), ("mype", C.c_int),
                ("globnx", C.c_int), ("globny", C.c_int), ("box", C.c_int * 8),
                ("nvar", C.c_int), ("dx", C.c_double), ("gamma", C.c_double),
                ("courant_factor", C.c_double), ("smallc", C.c_double), ("smallr", C.c_double),
                ("niter_riemann", C.c_int), ("iorder", C.c_int), ("slope_type", C.c_double),
                ("scheme", C.c_int), ("boundary_right", C.c_int), ("boundary_left", C.c_int),
                ("boundary_down", C.c_int), ("boundary_up", C.c_int), ("testCase", C.c_int)]


class HydroVar(C.Structure):
    _fields_ = [("uold", C.POINTER(C.c_double))]


def process_args(args) -> HydroParam:
    """The library's process_args (a C99 host function; no GPU needed)."""
    argv = (C.c_char_p * (len(args) + 1))(*([b"hydro"] + [a.encode() for a in args]))
    H = HydroParam()
    lib().process_args(len(args) + 1, argv, C.byref(H))
    return H


def vtkfile(step: int, H: HydroParam, uold: np.ndarray) -> None:
    """The library's vtkfile on a host array (writes ./Dep/..., ./Hydro.pvd)."""
    u = np.ascontiguousarray(uold, dtype=np.float64)
    hv = HydroVar(u.ctypes.data_as(C.POINTER(C.c_double)))
    lib().vtkfile(step, H, C.byref(hv))


_LIB = None


def lib() -> C.CDLL:
    """The loaded native library (raises if it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise HydroError(f"{LIB_PATH} is missing: build it with hydro_amd.build() "
                         "(make -C hydro_amd); there is no CPU fallback")
    L = C.CDLL(LIB_PATH)
    G = C.c_void_p
    P = C.POINTER(_CParams)
    L.hydro_amd_last_error.restype = C.c_char_p
    L.hydro_amd_grid_create.argtypes = [P]
    L.hydro_amd_grid_create.restype = G
    L.hydro_amd_grid_destroy.argtypes = [G]
    L.hydro_amd_grid_destroy.restype = None
    for name, args in [("hydro_amd_grid_init", [G, C.c_int]),
                       ("hydro_amd_grid_init_blast", [G, C.c_int, C.c_int, C.c_double]),
                       ("hydro_amd_grid_upload", [G, _dp]),
                       ("hydro_amd_grid_download", [G, _dp]),
                       ("hydro_amd_grid_image", [G, C.c_int, C.c_void_p, C.c_longlong,
                                                 C.POINTER(C.c_int), C.POINTER(C.c_int)]),
                       ("hydro_amd_grid_godunov", [G, C.c_int, C.c_double]),
                       ("hydro_amd_grid_make_boundary", [G, C.c_int]),
                       ("hydro_amd_grid_compute_deltat", [G, C.POINTER(C.c_double)]),
                       ("hydro_amd_grid_run", [G, C.c_int, C.c_void_p]),
                       ("hydro_amd_grid_run_async", [G, C.c_int]),
                       ("hydro_amd_grid_sync", [G]),
                       ("hydro_amd_grid_state", [G, C.POINTER(C.c_double), C.POINTER(C.c_int),
                                                 C.POINTER(C.c_double)]),
                       ("hydro_amd_grid_set_state", [G, C.c_double, C.c_int, C.c_double]),
                       ("hydro_amd_grid_set_limits", [G, C.c_double, C.c_int]),
                       ("hydro_amd_grid_set_tune", [G, C.c_int, C.c_int, C.c_int, C.c_int]),
                       ("hydro_amd_grid_set_profile", [G, C.c_int]),
                       ("hydro_amd_grid_autotune", [G, C.c_int, C.c_void_p]),
                       ("hydro_amd_grid_profile", [G, _dp, C.POINTER(C.c_longlong)])]:
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = C.c_int
    L.hydro_amd_grid_bytes.argtypes = [G]
    L.hydro_amd_grid_bytes.restype = C.c_longlong
    B = C.c_void_p
    L.hydro_amd_batch_create.argtypes = [C.c_int, C.c_longlong]
    L.hydro_amd_batch_create.restype = B
    L.hydro_amd_batch_destroy.argtypes = [B]
    L.hydro_amd_batch_destroy.restype = None
    for name, args in [("hydro_amd_batch_add", [B, P]),
                       ("hydro_amd_batch_remove", [B, C.c_int]),
                       ("hydro_amd_batch_init", [B, C.c_int, C.c_int]),
                       ("hydro_amd_batch_init_blast", [B, C.c_int, C.c_int, C.c_int, C.c_double]),
                       ("hydro_amd_batch_upload", [B, C.c_int, _dp]),
                       ("hydro_amd_batch_download", [B, C.c_int, _dp]),
                       ("hydro_amd_batch_run", [B, C.c_int]),
                       ("hydro_amd_batch_run_async", [B, C.c_int]),
                       ("hydro_amd_batch_sync", [B]),
                       ("hydro_amd_batch_query", [B, C.c_void_p, C.c_void_p, C.c_void_p,
                                                  C.c_void_p, C.c_void_p, C.c_int]),
                       ("hydro_amd_batch_set_tune", [B, C.c_int, C.c_int, C.c_int]),
                       ("hydro_amd_batch_autotune", [B, C.c_int, C.c_void_p])]:
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = C.c_int
    for name in ("hydro_amd_batch_bytes_needed", "hydro_amd_batch_bytes_used"):
        getattr(L, name).restype = C.c_longlong
    L.hydro_amd_batch_bytes_needed.argtypes = [B, C.c_int, C.c_int]
    L.hydro_amd_batch_bytes_used.argtypes = [B]
    S = C.c_void_p
    L.hydro_amd_now.restype = C.c_double
    L.hydro_amd_copy_probe.argtypes = [C.c_longlong, C.c_int]
    L.hydro_amd_copy_probe.restype = C.c_double
    L.hydro_amd_stream_probe.argtypes = [C.c_longlong, C.c_int]
    L.hydro_amd_stream_probe.restype = C.c_double
    L.hydro_amd_server_create.argtypes = [C.c_int, C.c_longlong, C.c_int, C.c_int]
    L.hydro_amd_server_create.restype = S
    L.hydro_amd_server_destroy.argtypes = [S]
    L.hydro_amd_server_destroy.restype = None
    for name, args in [("hydro_amd_server_submit", [S, P, C.c_int, C.c_int, C.c_double]),
                       ("hydro_amd_server_submit_state", [S, P, _dp]),
                       ("hydro_amd_server_pump", [S]),
                       ("hydro_amd_server_counts", [S, C.c_void_p, C.c_void_p, C.c_void_p,
                                                    C.c_void_p]),
                       ("hydro_amd_server_info", [S, C.c_int, C.c_void_p, C.c_void_p,
                                                  C.c_void_p]),
                       ("hydro_amd_server_result", [S, C.c_int, _dp]),
                       ("hydro_amd_server_release", [S, C.c_int]),
                       ("hydro_amd_server_reserve_results", [S, C.c_longlong]),
                       ("hydro_amd_server_result_precision", [S, C.c_int]),
                       ("hydro_amd_server_reruns", [S])]:
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = C.c_int
    L.hydro_amd_compute_deltat_noalloc.argtypes = [P, C.c_void_p, C.c_void_p, C.c_void_p]
    L.hydro_amd_compute_deltat_noalloc.restype = C.c_int
    L.process_args.argtypes = [C.c_int, C.POINTER(C.c_char_p), C.POINTER(HydroParam)]
    L.process_args.restype = None
    L.vtkfile.argtypes = [C.c_int, HydroParam, C.POINTER(HydroVar)]
    L.vtkfile.restype = None
    _LIB = L
    return L


def last_error() -> str:
    return lib().hydro_amd_last_error().decode()


def _check(rc: int) -> int:
    if rc < 0:
        raise HydroError(last_error())
    return rc


class Grid:
    """One grid on the GPU (hydro_amd_grid_*)."""

    def __init__(self, p: Params):
        self.p = p
        self._h = lib().hydro_amd_grid_create(C.byref(p.c()))
        if not self._h:
            raise HydroError(last_error())

    def close(self) -> None:
        if self._h:
            lib().hydro_amd_grid_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # state
    def init(self, testcase: int) -> None:
        _check(lib().hydro_amd_grid_init(self._h, testcase))

    def init_blast(self, bx: int, by: int, e: float) -> None:
        _check(lib().hydro_amd_grid_init_blast(self._h, bx, by, e))

    def upload(self, u: np.ndarray) -> None:
        u = np.ascontiguousarray(u, dtype=np.float64)
        if u.shape != self.p.shape():
            raise ValueError(f"uold shape {u.shape} != {self.p.shape()}")
        _check(lib().hydro_amd_grid_upload(self._h, u))

    def download(self) -> np.ndarray:
        u = np.empty(self.p.shape())
        _check(lib().hydro_amd_grid_download(self._h, u))
        return u

    def image(self, shrink_size: int = 1000) -> np.ndarray:
        """Density image of the current state, computed on the GPU
        (HydroC99_2Domp/Src/Image.c:111-166): uint8 [h, w, 3]."""
        w, h = C.c_int(), C.c_int()
        _check(lib().hydro_amd_grid_image(self._h, shrink_size, None, 0, C.byref(w), C.byref(h)))
        out = np.zeros(max(3 * w.value * h.value, 1), np.uint8)
        _check(lib().hydro_amd_grid_image(self._h, shrink_size, out.ctypes.data, out.size,
                                          C.byref(w), C.byref(h)))
        return out[:3 * w.value * h.value].reshape(h.value, w.value, 3)

    # the reference call surface on this grid
    def godunov(self, idim: int, dt: float) -> None:
        _check(lib().hydro_amd_grid_godunov(self._h, idim, dt))

    def make_boundary(self, idim: int) -> None:
        _check(lib().hydro_amd_grid_make_boundary(self._h, idim))

    def compute_deltat(self) -> float:
        d = C.c_double(0.0)
        _check(lib().hydro_amd_grid_compute_deltat(self._h, C.byref(d)))
        return d.value

    # device time loop
    def run(self, max_steps: int, log: bool = True):
        """Run up to max_steps steps on the GPU; returns the dt of each step."""
        buf = np.zeros(max(max_steps, 1)) if log else None
        ptr = buf.ctypes.data_as(C.c_void_p) if log else None
        n = _check(lib().hydro_amd_grid_run(self._h, max_steps, ptr))
        return buf[:n].copy() if log else n

    def run_async(self, max_steps: int) -> None:
        _check(lib().hydro_amd_grid_run_async(self._h, max_steps))

    def sync(self) -> None:
        _check(lib().hydro_amd_grid_sync(self._h))

    def state(self):
        t, n, dt = C.c_double(), C.c_int(), C.c_double()
        _check(lib().hydro_amd_grid_state(self._h, C.byref(t), C.byref(n), C.byref(dt)))
        return t.value, n.value, dt.value

    def set_state(self, t: float, nstep: int, dt: float) -> None:
        _check(lib().hydro_amd_grid_set_state(self._h, t, nstep, dt))

    def set_limits(self, tend: float, nstepmax: int) -> None:
        _check(lib().hydro_amd_grid_set_limits(self._h, tend, nstepmax))

    def set_tune(self, yseg: int = 64, ywaves: int = 1, xwaves: int = 4, xiter: int = 8) -> None:
        _check(lib().hydro_amd_grid_set_tune(self._h, yseg, ywaves, xwaves, xiter))

    def autotune(self, reps: int = 3) -> dict:
        """Pick the fastest sweep launch shapes for this grid (state unchanged)."""
        best = (C.c_int * 4)()
        _check(lib().hydro_amd_grid_autotune(self._h, reps, C.cast(best, C.c_void_p)))
        return dict(zip(("yseg", "ywaves", "xwaves", "xiter"), list(best)))

    def set_profile(self, on: bool) -> None:
        _check(lib().hydro_amd_grid_set_profile(self._h, int(on)))

    def profile(self):
        """{kernel: (total_ms, launches)} for x_sweep, y_sweep, boundary, courant."""
        ms = np.zeros(4)
        n = (C.c_longlong * 4)()
        _check(lib().hydro_amd_grid_profile(self._h, ms, n))
        return {k: (float(ms[i]), int(n[i]))
                for i, k in enumerate(("x_sweep", "y_sweep", "boundary", "courant"))}

    @property
    def nbytes(self) -> int:
        return int(lib().hydro_amd_grid_bytes(self._h))


class Batch:
    """Many independent grids advanced per launch (hydro_amd_batch_*)."""

    def __init__(self, precision: int = FP64, budget_bytes: int = 0):
        self.precision = precision
        self._h = lib().hydro_amd_batch_create(precision, int(budget_bytes))
        if not self._h:
            raise HydroError(last_error())
        self.params = {}

    def close(self) -> None:
        if self._h:
            lib().hydro_amd_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def add(self, p: Params) -> int:
        """Admit an instance; raises HydroError (e.g. over budget)."""
        i = lib().hydro_amd_batch_add(self._h, C.byref(p.c()))
        if i < 0:
            raise HydroError(last_error())
        self.params[i] = p
        return i

    def try_add(self, p: Params):
        """Admit an instance, or return None when it does not fit the budget."""
        i = lib().hydro_amd_batch_add(self._h, C.byref(p.c()))
        if i < 0:
            return None
        self.params[i] = p
        return i

    def remove(self, i: int) -> None:
        _check(lib().hydro_amd_batch_remove(self._h, i))
        self.params.pop(i, None)

    def bytes_needed(self, nx: int, ny: int) -> int:
        return int(lib().hydro_amd_batch_bytes_needed(self._h, nx, ny))

    @property
    def bytes_used(self) -> int:
        return int(lib().hydro_amd_batch_bytes_used(self._h))

    def init(self, i: int, testcase: int) -> None:
        _check(lib().hydro_amd_batch_init(self._h, i, testcase))

    def init_blast(self, i: int, bx: int, by: int, e: float) -> None:
        _check(lib().hydro_amd_batch_init_blast(self._h, i, bx, by, e))

    def upload(self, i: int, u: np.ndarray) -> None:
        _check(lib().hydro_amd_batch_upload(self._h, i, np.ascontiguousarray(u, np.float64)))

    def download(self, i: int) -> np.ndarray:
        u = np.empty(self.params[i].shape())
        _check(lib().hydro_amd_batch_download(self._h, i, u))
        return u

    def run(self, steps: int) -> None:
        _check(lib().hydro_amd_batch_run(self._h, steps))

    def run_async(self, steps: int) -> None:
        _check(lib().hydro_amd_batch_run_async(self._h, steps))

    def sync(self) -> None:
        _check(lib().hydro_amd_batch_sync(self._h))

    def query(self):
        """{id: (t, nstep, dt, state)} for every live instance; state 1 = more
        steps to run, 0 = ended, -1 = stopped at a non-positive/non-finite dt."""
        cap = max(len(self.params), 1)
        ids = np.zeros(cap, np.int32)
        t = np.zeros(cap)
        n = np.zeros(cap, np.int32)
        dt = np.zeros(cap)
        act = np.zeros(cap, np.int32)
        ptr = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
        k = _check(lib().hydro_amd_batch_query(self._h, ptr(ids), ptr(t), ptr(n), ptr(dt),
                                               ptr(act), cap))
        return {int(ids[j]): (float(t[j]), int(n[j]), float(dt[j]), int(act[j]))
                for j in range(min(k, cap))}

    def set_tune(self, waves: int, xiter: int, yseg: int) -> None:
        _check(lib().hydro_amd_batch_set_tune(self._h, waves, xiter, yseg))

    def autotune(self, reps: int = 2) -> dict:
        best = (C.c_int * 3)()
        _check(lib().hydro_amd_batch_autotune(self._h, reps, C.cast(best, C.c_void_p)))
        return dict(zip(("waves", "xiter", "yseg"), list(best)))


def copy_probe(nbytes: int, reps: int = 10) -> float:
    """GB/s (read + written) of a plain copy kernel over nbytes."""
    v = lib().hydro_amd_copy_probe(int(nbytes), reps)
    if v < 0:
        raise HydroError(last_error())
    return float(v)


def stream_probe(cells: int, reps: int = 10) -> float:
    """GB/s (64 bytes per cell) of a copy in the sweeps' access shape: four
    arrays of `cells` doubles in, four out, non-temporal loads and stores."""
    v = lib().hydro_amd_stream_probe(int(cells), reps)
    if v < 0:
        raise HydroError(last_error())
    return float(v)


def compute_deltat_noalloc(p: Params, d_uold: int, d_dt: int, stream: int = 0) -> None:
    """hydro_amd_compute_deltat_noalloc: dt of the FP64 state at device address
    d_uold ([4, ny+4, nx+4] float64) written to the device double at d_dt, queued
    on `stream` (a hipStream_t handle, 0 = the null stream) without waiting;
    allocates no device memory."""
    _check(lib().hydro_amd_compute_deltat_noalloc(C.byref(p.c()), C.c_void_p(d_uold),
                                                  C.c_void_p(d_dt), C.c_void_p(stream or None)))


def now() -> float:
    """The library's clock (CLOCK_MONOTONIC seconds), used for request times."""
    return float(lib().hydro_amd_now())


class Server:
    """Continuous-batching request loop (hydro_amd_server_*)."""

    QUEUED, RUNNING, DONE, FAILED = 0, 1, 2, 3

    def __init__(self, precision: int = MIXED, budget_bytes: int = 200 * 10**9,
                 chunk_steps: int = 20, keep_results: bool = False):
        self._h = lib().hydro_amd_server_create(precision, int(budget_bytes), chunk_steps,
                                                int(keep_results))
        if not self._h:
            raise HydroError(last_error())
        self.params = {}

    def close(self) -> None:
        if self._h:
            lib().hydro_amd_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, p: Params, bx: int, by: int, e: float) -> int:
        r = _check(lib().hydro_amd_server_submit(self._h, C.byref(p.c()), bx, by, e))
        self.params[r] = p
        return r

    def submit_state(self, p: Params, u: np.ndarray) -> int:
        """Queue a run from a whole initial state [4, ny+4, nx+4] (float64)."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        if u.shape != p.shape():
            raise ValueError(f"uold shape {u.shape} != {p.shape()}")
        r = _check(lib().hydro_amd_server_submit_state(self._h, C.byref(p.c()), u))
        self.params[r] = p
        return r

    def pump(self) -> int:
        return _check(lib().hydro_amd_server_pump(self._h))

    def counts(self):
        v = [C.c_int() for _ in range(4)]
        _check(lib().hydro_amd_server_counts(self._h, *[C.cast(C.byref(x), C.c_void_p)
                                                         for x in v]))
        return dict(zip(("queued", "running", "done", "failed"), [x.value for x in v]))

    def info(self, r: int):
        """(state, submitted, admitted, finished, t, nstep)."""
        times = (C.c_double * 3)()
        t = C.c_double()
        n = C.c_int()
        st = lib().hydro_amd_server_info(self._h, r, C.cast(times, C.c_void_p),
                                         C.cast(C.byref(t), C.c_void_p),
                                         C.cast(C.byref(n), C.c_void_p))
        if st < 0:
            raise HydroError(last_error())
        return st, times[0], times[1], times[2], t.value, n.value

    def result(self, r: int) -> np.ndarray:
        u = np.empty(self.params[r].shape())
        _check(lib().hydro_amd_server_result(self._h, r, u))
        return u

    def release(self, r: int) -> None:
        _check(lib().hydro_amd_server_release(self._h, r))

    def result_precision(self, r: int) -> int:
        """The precision request r's result was computed in (FP64 for a mixed
        run that broke down and was run again)."""
        return _check(lib().hydro_amd_server_result_precision(self._h, r))

    @property
    def reruns(self) -> int:
        return _check(lib().hydro_amd_server_reruns(self._h))

    def reserve_results(self, nbytes: int) -> None:
        """Pin host memory for kept results (copied by DMA beside the kernels)."""
        _check(lib().hydro_amd_server_reserve_results(self._h, int(nbytes)))
