"""The slope limiter of slope.c:71-90 with slope_type 1, as the fast and mixed
forms compute it on the GPU (hydro_numerics.hpp, slope1_med): the median of
the two one-sided differences and zero.

The reference computes dc = (dl + dr) / 2, sg = dc > 0 ? 1 : -1, keep =
dl * dr > 0 and sg * fmin(keep * fmin(|dl|, |dr|), |dc|).  Where dl and dr
share a sign, fl(|dl + dr|) / 2 >= min(|dl|, |dr|) because rounding is
monotone, so the limiter reduces to median(dl, dr, 0).  These checks hold the
two forms equal, in FP64 and FP32, on random and edge-case differences:
same value (a zero of either sign counts as zero), except where dl * dr
underflows to zero, where the reference's slope is 0 and the median the
smaller difference (below sqrt of the smallest subnormal).
"""
import numpy as np
import pytest


def reference_limiter(dl, dr):
    """slope.c:80-88, slope_type 1, in the arrays' precision and the
    reference's order of operations."""
    t = dl.dtype.type
    dc = t(0.5) * (dl + dr)
    sg = np.where(dc > 0, t(1), t(-1))
    keep = np.where(dl * dr <= 0, t(0), t(1))
    t1 = np.fmin(np.abs(dl), np.abs(dr))
    return sg * np.fmin(keep * t1, np.abs(dc))


def median_limiter(dl, dr):
    """max(min(dl, dr), min(max(dl, dr), 0)): the FP64 form's four
    min/max instructions (v_med3_f32 computes the same in FP32)."""
    z = dl.dtype.type(0)
    return np.fmax(np.fmin(dl, dr), np.fmin(np.fmax(dl, dr), z))


def cases(dtype, n=200000, seed=20261203):
    rng = np.random.default_rng(seed)
    info = np.finfo(dtype)
    mags = [rng.uniform(-1, 1, n),                                  # O(1)
            rng.standard_normal(n) * 10.0 ** rng.integers(-30, 30, n),  # wide range
            np.sign(rng.standard_normal(n)) * np.exp(rng.uniform(-20, 20, n))]
    dl = np.concatenate([m.astype(dtype) for m in mags])
    dr = np.concatenate([np.roll(m, 1).astype(dtype) for m in mags])
    # equal, opposite, zeros of both signs, one-ulp neighbours, extremes
    edge = np.array([0.0, -0.0, 1.0, -1.0, 3.0, info.tiny, -info.tiny, info.max / 4,
                     -info.max / 4, 1.0 + info.eps, 1.0 - info.eps / 2], dtype=dtype)
    a, b = np.meshgrid(edge, edge)
    return np.concatenate([dl, a.ravel()]), np.concatenate([dr, b.ravel()])


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_median_form_equals_reference_limiter(dtype):
    dl, dr = cases(dtype)
    with np.errstate(over="ignore", under="ignore"):
        ref = reference_limiter(dl, dr)
        med = median_limiter(dl, dr)
        underflow = (dl * dr == 0) & (dl != 0) & (dr != 0)
    ok = ~underflow
    assert np.array_equal(ref[ok], med[ok]), np.flatnonzero(ref[ok] != med[ok])[:10]
    # where the product underflows the two differ by less than the smaller
    # difference, itself below sqrt of the smallest subnormal
    bound = np.sqrt(np.finfo(dtype).smallest_subnormal.astype(np.float64) * 4)
    assert np.all(np.abs(med[underflow].astype(np.float64)) <= bound)
    assert np.all(ref[underflow] == 0)


def test_median_form_cases_by_hand():
    d = np.array
    # same sign: the smaller difference with its sign
    assert median_limiter(d([2.0]), d([3.0]))[0] == 2.0
    assert median_limiter(d([-2.0]), d([-0.5]))[0] == -0.5
    # opposite signs or a zero: flat
    assert median_limiter(d([2.0]), d([-3.0]))[0] == 0.0
    assert median_limiter(d([0.0]), d([5.0]))[0] == 0.0
    # and the reference agrees
    for a, b in ((2.0, 3.0), (-2.0, -0.5), (2.0, -3.0), (0.0, 5.0), (1e-200, 1e-300)):
        assert reference_limiter(d([a]), d([b]))[0] == median_limiter(d([a]), d([b]))[0] or (
            a * b == 0.0)
