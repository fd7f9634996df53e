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


class DeviceArray:
    """Device memory from the HIP runtime the library itself links (for tests
    of entry points that take device pointers)."""

    _hip = None

    @classmethod
    def hip(cls):
        if cls._hip is None:
            import ctypes as C
            hydro_amd.lib()  # loads libamdhip64.so.7 as the library's dependency
            h = C.CDLL("libamdhip64.so.7")  # the instance already loaded
            h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            h.hipFree.argtypes = [C.c_void_p]
            h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            for f in (h.hipMalloc, h.hipFree, h.hipMemcpy, h.hipDeviceSynchronize):
                f.restype = C.c_int
            cls._hip = h
        return cls._hip

    def __init__(self, nbytes: int):
        import ctypes as C
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        assert self.hip().hipMalloc(C.byref(p), self.nbytes) == 0, "hipMalloc failed"
        self.ptr = p.value

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert self.hip().hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, 1) == 0  # host->device

    def download(self, dtype, count: int) -> np.ndarray:
        out = np.empty(count, dtype)
        assert out.nbytes <= self.nbytes
        assert self.hip().hipDeviceSynchronize() == 0
        assert self.hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) == 0  # device->host
        return out

    def free(self) -> None:
        if self.ptr:
            self.hip().hipFree(self.ptr)
            self.ptr = None
