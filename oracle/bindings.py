"""ctypes bindings for the oracle libraries (test infrastructure only)."""
from __future__ import annotations

import ctypes as C
import dataclasses
import os
import subprocess
import tempfile

import numpy as np

ORACLE_DIR = os.path.dirname(os.path.abspath(__file__))
PORT_SO = os.path.join(ORACLE_DIR, "_port", "libhydro_oracle.so")
PORT_FMA_SO = os.path.join(ORACLE_DIR, "_port", "libhydro_oracle_fma.so")
REF_SO = os.path.join(ORACLE_DIR, "_ref", "libhydroref.so")
REF_SP_SO = os.path.join(ORACLE_DIR, "_ref", "libhydroref_sp.so")  # -DPREC_SP
REF_OMP = os.path.join(ORACLE_DIR, "_ref", "hydro_omp")
REF_D = os.path.join(ORACLE_DIR, "_ref", "hydro_d")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build(ref: bool = True) -> None:
    """Compile the port (always) and the reference build (when its sources exist)."""
    target = [] if ref else ["port"]
    subprocess.run(["make", "-s", "-C", ORACLE_DIR] + target, check=True)


def port_available() -> bool:
    return os.path.exists(PORT_SO)


def ref_available() -> bool:
    return os.path.exists(REF_SO)


@dataclasses.dataclass
class Params:
    """Run parameters; defaults are HydroC99_2DMpi/Src/parametres.c:64-109."""
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
    testcase: int = 0
    nxystep: int = -1

    @property
    def nxt(self) -> int:
        return self.nx + 4

    @property
    def nyt(self) -> int:
        return self.ny + 4

    def shape(self):
        return (4, self.nyt, self.nxt)


class _HoParams(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("dx", C.c_double), ("gamma", C.c_double),
                ("courant_factor", C.c_double), ("smallc", C.c_double), ("smallr", C.c_double),
                ("slope_type", C.c_double), ("niter_riemann", C.c_int), ("iorder", C.c_int),
                ("scheme", C.c_int), ("boundary_left", C.c_int), ("boundary_right", C.c_int),
                ("boundary_down", C.c_int), ("boundary_up", C.c_int)]


def _ho(p: Params) -> _HoParams:
    return _HoParams(p.nx, p.ny, p.dx, p.gamma, p.courant_factor, p.smallc, p.smallr,
                     p.slope_type, p.niter_riemann, p.iorder, p.scheme, p.boundary_left,
                     p.boundary_right, p.boundary_down, p.boundary_up)


def deck_text(p: Params) -> str:
    """A .nml deck in the reference's key=value format (parametres.c:162-257)."""
    scheme = {1: "muscl", 2: "plmde", 3: "collela"}[p.scheme]
    lines = ["&RUN", f"tend={p.tend!r}", f"nstepmax={p.nstepmax}", "/", "&MESH",
             f"nx={p.nx}", f"ny={p.ny}"]
    if p.nxystep > 0:
        lines.append(f"nxystep={p.nxystep}")
    lines += [f"dx={p.dx!r}", f"boundary_left={p.boundary_left}",
              f"boundary_right={p.boundary_right}", f"boundary_down={p.boundary_down}",
              f"boundary_up={p.boundary_up}", f"testcase={p.testcase}", "/", "&HYDRO",
              f"courant_factor={p.courant_factor!r}", f"niter_riemann={p.niter_riemann}",
              f"iorder={p.iorder}", f"slope_type={p.slope_type!r}", f"scheme={scheme}",
              f"smallr={p.smallr!r}", f"smallc={p.smallc!r}"]
    if p.gamma != 1.4:
        lines.append(f"gamma={p.gamma!r}")
    lines.append("/")
    return "\n".join(lines) + "\n"


def write_deck(p: Params, path: str | None = None) -> str:
    if path is None:
        fd, path = tempfile.mkstemp(suffix=".nml")
        os.close(fd)
    with open(path, "w") as f:
        f.write(deck_text(p))
    return path


class Port:
    """This repository's C99 restatement (oracle/hydro_oracle.c).  fma=True
    loads its build with FMA contraction: an equally valid FP64 rounding of
    the same algorithm, used to measure how far rounding alone moves a result
    (not bit-identical to the reference)."""

    def __init__(self, fma: bool = False):
        path = PORT_FMA_SO if fma else PORT_SO
        if not os.path.exists(path):
            build(ref=False)
        lib = C.CDLL(path)
        P = C.POINTER(_HoParams)
        lib.ho_init.argtypes = [P, C.c_int, _dp]
        lib.ho_init_blast.argtypes = [P, C.c_int, C.c_int, C.c_double, _dp]
        lib.ho_make_boundary.argtypes = [C.c_int, P, _dp]
        lib.ho_compute_deltat.argtypes = [P, _dp]
        lib.ho_compute_deltat.restype = C.c_double
        lib.ho_godunov.argtypes = [C.c_int, C.c_double, P, _dp]
        lib.ho_run.argtypes = [P, _dp, C.c_double, C.c_int, C.c_int, C.POINTER(C.c_double),
                               C.POINTER(C.c_int), C.POINTER(C.c_double), _dp]
        lib.ho_run.restype = C.c_int
        lib.ho_constoprim.argtypes = [C.c_int, _dp, _dp, _dp, C.c_double]
        lib.ho_eos.argtypes = [C.c_int, _dp, _dp, _dp, C.c_double, C.c_double]
        lib.ho_slope.argtypes = [C.c_int, _dp, _dp, C.c_double]
        lib.ho_trace.argtypes = [C.c_int, _dp, _dp, _dp, C.c_double, C.c_int, _dp, _dp]
        lib.ho_qleftright.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp]
        lib.ho_riemann.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _ip, C.c_double, C.c_double,
                                   C.c_double, C.c_int]
        lib.ho_cmpflx.argtypes = [C.c_int, C.c_int, _dp, _dp, C.c_double]
        lib.ho_image.argtypes = [C.c_int, C.c_int, _dp, C.c_int, C.c_void_p,
                                 C.POINTER(C.c_int), C.POINTER(C.c_int)]
        lib.ho_image.restype = C.c_int
        self.lib = lib

    # whole grid
    def init(self, p: Params, testcase: int | None = None) -> np.ndarray:
        u = np.zeros(p.shape())
        self.lib.ho_init(_ho(p), p.testcase if testcase is None else testcase, u)
        return u

    def init_blast(self, p: Params, bx: int, by: int, e_blast: float) -> np.ndarray:
        u = np.zeros(p.shape())
        self.lib.ho_init_blast(_ho(p), bx, by, e_blast, u)
        return u

    def make_boundary(self, idim: int, p: Params, u: np.ndarray) -> None:
        self.lib.ho_make_boundary(idim, _ho(p), u)

    def compute_deltat(self, p: Params, u: np.ndarray) -> float:
        return self.lib.ho_compute_deltat(_ho(p), u)

    def godunov(self, idim: int, dt: float, p: Params, u: np.ndarray) -> None:
        self.lib.ho_godunov(idim, dt, _ho(p), u)

    def run(self, p: Params, u: np.ndarray, max_steps: int, state=None):
        """Advance u in place; returns (dt_log, state) with state = (t, nstep, dt)."""
        t, n, dt = (0.0, 0, 0.0) if state is None else state
        ct, cn, cdt = C.c_double(t), C.c_int(n), C.c_double(dt)
        log = np.zeros(max(max_steps, 1))
        k = self.lib.ho_run(_ho(p), u, p.tend, p.nstepmax, max_steps, C.byref(ct), C.byref(cn),
                            C.byref(cdt), log)
        return log[:k].copy(), (ct.value, cn.value, cdt.value)

    def image(self, u: np.ndarray, shrink_size: int = 1000) -> np.ndarray:
        """Density image of a [4, ny+4, nx+4] state: uint8 [h, w, 3]."""
        nx, ny = u.shape[2] - 4, u.shape[1] - 4
        rho = np.ascontiguousarray(u[0], dtype=np.float64)
        w, h = C.c_int(), C.c_int()
        self.lib.ho_image(nx, ny, rho, shrink_size, None, C.byref(w), C.byref(h))
        out = np.zeros(max(3 * w.value * h.value, 1), np.uint8)
        self.lib.ho_image(nx, ny, rho, shrink_size, out.ctypes.data, C.byref(w), C.byref(h))
        return out[:3 * w.value * h.value].reshape(h.value, w.value, 3)

    # stages over one line, SoA [4][n]
    def constoprim(self, u, smallr):
        n = u.shape[1]
        q = np.zeros((4, n)); e = np.zeros(n)
        self.lib.ho_constoprim(n, np.ascontiguousarray(u), q, e, smallr)
        return q, e

    def eos(self, e, q, smallc, gamma):
        n = e.shape[0]
        q = q.copy(); c = np.zeros(n)
        self.lib.ho_eos(n, np.ascontiguousarray(e), q, c, smallc, gamma)
        return q, c

    def slope(self, q, slope_type):
        n = q.shape[1]
        dq = np.zeros((4, n))
        self.lib.ho_slope(n, np.ascontiguousarray(q), dq, slope_type)
        return dq

    def trace(self, q, dq, c, dtdx, scheme=1):
        n = q.shape[1]
        qxm = np.zeros((4, n)); qxp = np.zeros((4, n))
        self.lib.ho_trace(n, np.ascontiguousarray(q), np.ascontiguousarray(dq),
                          np.ascontiguousarray(c), dtdx, scheme, qxm, qxp)
        return qxm, qxp

    def qleftright(self, nface, qxm, qxp):
        n = qxm.shape[1]
        ql = np.zeros((4, n)); qr = np.zeros((4, n))
        self.lib.ho_qleftright(nface, n, np.ascontiguousarray(qxm), np.ascontiguousarray(qxp),
                               ql, qr)
        return ql, qr

    def riemann(self, ql, qr, smallr, smallc, gamma, niter, nface=None):
        n = ql.shape[1]
        nface = n if nface is None else nface
        qg = np.zeros((4, n)); sg = np.zeros(n, dtype=np.int32)
        self.lib.ho_riemann(nface, n, np.ascontiguousarray(ql), np.ascontiguousarray(qr), qg, sg,
                            smallr, smallc, gamma, niter)
        return qg, sg

    def cmpflx(self, qg, gamma, nface=None):
        n = qg.shape[1]
        nface = n if nface is None else nface
        fl = np.zeros((4, n))
        self.lib.ho_cmpflx(nface, n, np.ascontiguousarray(qg), fl, gamma)
        return fl


class Ref:
    """The reference HydroC99_2DMpi sources, compiled in place (oracle/_ref).

    The reference keeps its state in globals, so one deck is open per process
    at a time (per build); ``open``/``close`` bracket it.  sp=True loads the
    reference's all-FP32 build (-DPREC_SP: real_t = float,
    HydroC99_2DMpi/Src/parametres.h:41-45); uold, dt and t cross as FP64
    either way, and its stage functions are not bound."""

    _libs = {}

    def __init__(self, sp: bool = False):
        path = REF_SP_SO if sp else REF_SO
        if not os.path.exists(path):
            raise FileNotFoundError(path + " not built (needs the reference sources)")
        if path not in Ref._libs:
            lib = C.CDLL(path)
            lib.ref_open.argtypes = [C.c_char_p, _ip]
            lib.ref_open.restype = C.c_int
            lib.ref_get_reals.argtypes = [_dp]
            lib.ref_get_ints.argtypes = [_ip]
            lib.ref_set_gamma.argtypes = [C.c_double]
            lib.ref_uold_len.restype = C.c_long
            lib.ref_get_uold.argtypes = [_dp]
            lib.ref_set_uold.argtypes = [_dp]
            lib.ref_compute_deltat.argtypes = [C.POINTER(C.c_double)]
            lib.ref_godunov.argtypes = [C.c_int, C.c_double]
            lib.ref_make_boundary.argtypes = [C.c_int]
            lib.ref_steps.argtypes = [C.c_int, _dp, _dp]
            lib.ref_steps.restype = C.c_int
            lib.ref_vtkfile.argtypes = [C.c_int]
            if sp:
                Ref._libs[path] = lib
        if path not in Ref._libs:
            lib = Ref._lib_stages(lib)
            Ref._libs[path] = lib
        self.lib = Ref._libs[path]
        self.dims = None

    @staticmethod
    def _lib_stages(lib):
        # the reference's own stage functions (HydroC99_2DMpi/Src/*.h), used
        # with slices = Hstep = 1 so the work arrays are SoA [nvar][Hnxyt]
        lib.constoprim.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int,
                                   _dp, _dp, _dp]
        lib.equation_of_state.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                          C.c_double, C.c_int, C.c_int, _dp, _dp, _dp]
        lib.slope.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, _dp,
                              _dp]
        lib.trace.argtypes = [C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                              C.c_int, _dp, _dp, _dp, _dp, _dp]
        lib.qleftright.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                   C.c_int, _dp, _dp, _dp, _dp]
        lib.cmpflx.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, _dp,
                               _dp]
        lib.ref_riemann_line.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double,
                                         C.c_double, C.c_int, _dp, _dp, _dp, _ip]
        return lib

    # time loop on a deck
    def open(self, p: Params, gamma: float | None = None):
        path = write_deck(p)
        dims = np.zeros(8, dtype=np.int32)
        try:
            rc = self.lib.ref_open(path.encode(), dims)
        finally:
            os.unlink(path)
        if rc != 0:
            raise RuntimeError("reference harness already has an open deck")
        if gamma is not None:
            self.lib.ref_set_gamma(gamma)
        self.dims = dims
        return dims

    def close(self):
        self.lib.ref_close()

    def get_uold(self) -> np.ndarray:
        nx, ny, nxt, nyt = self.dims[:4]
        u = np.zeros((4, nyt, nxt))
        self.lib.ref_get_uold(u)
        return u

    def set_uold(self, u: np.ndarray) -> None:
        self.lib.ref_set_uold(np.ascontiguousarray(u, dtype=np.float64))

    def compute_deltat(self) -> float:
        d = C.c_double(0.0)
        self.lib.ref_compute_deltat(C.byref(d))
        return d.value

    def godunov(self, idim: int, dt: float) -> None:
        self.lib.ref_godunov(idim, dt)

    def make_boundary(self, idim: int) -> None:
        self.lib.ref_make_boundary(idim)

    def steps(self, k: int):
        dts = np.zeros(max(k, 1)); ts = np.zeros(max(k, 1))
        n = self.lib.ref_steps(k, dts, ts)
        return dts[:n].copy(), ts[:n].copy()

    # stage functions, one line: arrays SoA [4][n]
    def constoprim(self, u, smallr):
        n = u.shape[1]
        q = np.zeros((4, n)); e = np.zeros(n)
        self.lib.constoprim(n, n, 4, smallr, 1, 1, np.ascontiguousarray(u), q, e)
        return q, e

    def eos(self, e, q, smallc, gamma):
        n = e.shape[0]
        q = q.copy(); c = np.zeros(n)
        self.lib.equation_of_state(0, n, n, 4, smallc, gamma, 1, 1, np.ascontiguousarray(e), q,
                                   c)
        return q, c

    def slope(self, q, slope_type):
        n = q.shape[1]
        dq = np.zeros((4, n))
        self.lib.slope(n, 4, n, slope_type, 1, 1, np.ascontiguousarray(q), dq)
        return dq

    def trace(self, q, dq, c, dtdx, scheme=1):
        n = q.shape[1]
        qxm = np.zeros((4, n)); qxp = np.zeros((4, n))
        self.lib.trace(dtdx, n, scheme, 4, n, 1, 1, np.ascontiguousarray(q),
                       np.ascontiguousarray(dq), np.ascontiguousarray(c), qxm, qxp)
        return qxm, qxp

    def qleftright(self, nface, qxm, qxp):
        """idim=1 with Hnx = nface-1 so bmax = nface (qleftright.c:58-62)."""
        n = qxm.shape[1]
        ql = np.zeros((4, n)); qr = np.zeros((4, n))
        self.lib.qleftright(1, nface - 1, nface - 1, n, 4, 1, 1, np.ascontiguousarray(qxm),
                            np.ascontiguousarray(qxp), ql, qr)
        return ql, qr

    def riemann(self, ql, qr, smallr, smallc, gamma, niter, nface=None):
        n = ql.shape[1]
        nface = n if nface is None else nface
        qg = np.zeros((4, n)); sg = np.zeros(n, dtype=np.int32)
        self.lib.ref_riemann_line(nface, n, smallr, smallc, gamma, niter,
                                  np.ascontiguousarray(ql), np.ascontiguousarray(qr), qg, sg)
        return qg, sg

    def cmpflx(self, qg, gamma, nface=None):
        n = qg.shape[1]
        nface = n if nface is None else nface
        fl = np.zeros((4, n))
        self.lib.cmpflx(nface, n, 4, gamma, 1, 1, np.ascontiguousarray(qg), fl)
        return fl
