// Per-cell and per-face numerics of HydroC99's split Godunov/MUSCL step, as
// inline device functions shared by every sweep kernel (single grids, batched
// grids, FP64 and mixed precision).
//
// Parity rules (FP64 form, R = double): every expression keeps the reference's
// evaluation order (C's left-to-right association), the library is compiled
// with -ffp-contract=off so no a*b+c is fused, and '/' and sqrt are the
// correctly rounded hardware sequences.  With those, each function returns the
// same bits as the reference routine cited beside it.
//
// Mixed precision (R = float): storage, primitives, slopes, traces, wave
// speeds and fluxes are float; the Riemann solver's Newton pressure
// iteration runs in double (riemann_flux), the update accumulates in double
// (update1) and the CFL maximum and dt are double in the callers.
//
// FP64 fast form (R = double, form kFast): FP64 throughout, but not bit for
// bit: quotients and square roots come from the hardware estimates refined to
// within about an ulp (shared reciprocals, no correctly rounded sequences),
// the Newton iteration is rearranged to one division per step, and the unit
// that holds these kernels is compiled with multiply-add contraction.  Its
// results stay within a few ulp per operation of the reference's.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

namespace hydro {

// Arithmetic forms of the FP64 sweeps (the FD template argument of the sweep
// bodies): kIeee -- the IEEE operations alone; kGated -- the range-gated
// divisions (FastDiv below), bit-identical to kIeee; kFast -- the FP64 fast
// form (estimate-based quotients and roots, see prim_fast / riemann_fast).
// The FP32-storage sweeps take kIeee (their arithmetic is chosen by R).
constexpr int kIeee = 0;
constexpr int kGated = 1;
constexpr int kFast = 2;

// Physical / scheme constants, precomputed per launch in the compute type.
template <typename R>
struct Phys {
  R gamma, gm1, smallr, smallc, smallp, entho, slope_type;
  R slope_inv;    // 1 / slope_type (the FP64 fast form's slopes)
  double smallr_inv;  // 1 / smallr (the FP64 fast form's star state)
  double gamma_d, gamma6_d, smallp_d, smallpp_d;  // Newton constants (double)
  int niter;      // niter_riemann
  int iorder;     // 1: first order (dq = 0)
  int scheme;     // 1 muscl, 2 plmde, 3 collela
  int fastdiv;    // FP64 only: the parameters allow the range-gated divisions
};

template <typename R>
__host__ __device__ inline Phys<R> make_phys(double gamma, double smallr, double smallc,
                                             double slope_type, int niter, int iorder,
                                             int scheme) {
  Phys<R> p;
  // equation_of_state.c:60 / riemann.c:97-99 / cmpflx.c:60
  p.gamma = (R) gamma;
  p.gm1 = (R) gamma - (R) 1;
  p.smallr = (R) smallr;
  p.smallc = (R) smallc;
  p.smallp = ((R) smallc * (R) smallc) / (R) gamma;
  p.entho = (R) 1 / ((R) gamma - (R) 1);
  p.slope_type = (R) slope_type;
  p.slope_inv = (R) (1.0 / slope_type);
  p.smallr_inv = 1.0 / smallr;
  p.gamma_d = gamma;
  p.gamma6_d = (gamma + 1.0) / (2.0 * gamma);
  p.smallp_d = (smallc * smallc) / gamma;
  p.smallpp_d = smallr * p.smallp_d;
  p.niter = niter;
  p.iorder = iorder;
  p.scheme = scheme;
  // Bounds the gated FP64 divisions rely on (see FastDiv): gamma in
  // [1.01, 3], smallr in [2^-100, 2^30], smallc in [2^-50, 2^30].
  p.fastdiv = sizeof(R) == 8 && gamma >= 1.01 && gamma <= 3.0 && smallr >= 0x1p-100 &&
              smallr <= 0x1p30 && smallc >= 0x1p-50 && smallc <= 0x1p30;
  return p;
}

template <typename R>
struct Q4 {  // rho, normal velocity, transverse velocity, pressure
  R r, u, v, p;
};

template <typename R>
struct Prim {
  Q4<R> q;
  R c;     // sound speed
  R rinv;  // FP64 with FAST / kFast: the refined reciprocal of q.r
  R cinv;  // kFast: 1 / c
};

template <typename R>
__device__ __forceinline__ R maxm(R x, R y) { return x > y ? x : y; }  // utils.h:47 MAX

// True when p holds on every active lane of the wave (a wave-uniform value).
__device__ __forceinline__ bool wave_all(bool p) { return __builtin_amdgcn_ballot_w64(!p) == 0; }

// Correctly rounded FP64 square root, fast form.  The compiler lowers sqrt(x)
// to: x' = x * 2^256 if x < 2^-767; the v_rsq_f64 estimate g of x'; two
// coupled Newton-Raphson refinements of x'*g and g/2 (sqrt_core below, op for
// op); the result scaled back by 2^-128 if x was scaled; and x' itself
// returned for +-0 and +inf.  For x in [2^-767, DBL_MAX] both scalings are by
// 2^0 and the select is not taken, so sqrt(x) is sqrt_core(x) bit for bit.
// sqrt_fast takes the core alone when every active lane of the wave is in
// that range -- the normal case: the arguments here are gamma*p*rho-like
// products floored by smallp and smallr -- and the full sequence otherwise.
__device__ __forceinline__ double sqrt_core(double x) {
  const double g = __builtin_amdgcn_rsq(x);
  double s = x * g;
  double h = g * 0.5;
  const double r = __builtin_fma(-h, s, 0.5);
  s = __builtin_fma(s, r, s);
  h = __builtin_fma(h, r, h);
  double d = __builtin_fma(-s, s, x);
  s = __builtin_fma(d, h, s);
  d = __builtin_fma(-s, s, x);
  return __builtin_fma(d, h, s);
}
__device__ __forceinline__ bool sqrt_core_exact(double x) {
  // high word in [hi(2^-767), hi(DBL_MAX)]: 2^-767 <= x <= DBL_MAX, x not NaN
  const unsigned hi = (unsigned) __double2hiint(x);
  return hi - 0x10000000u <= 0x7fefffffu - 0x10000000u;
}
__device__ __forceinline__ double sqrt_fast(double x) {
  if (wave_all(sqrt_core_exact(x))) return sqrt_core(x);
  return sqrt(x);
}
__device__ __forceinline__ void sqrt_fast2(double x, double y, double& sx, double& sy) {
  if (wave_all(sqrt_core_exact(x) && sqrt_core_exact(y))) {
    sx = sqrt_core(x);
    sy = sqrt_core(y);
  } else {
    sx = sqrt(x);
    sy = sqrt(y);
  }
}

// ---- FP64 division without the range fix-ups ---------------------------------
//
// The compiler's FP64 a / b is: b' = v_div_scale(b), a' = v_div_scale(a), the
// v_rcp_f64 estimate of b' refined by two Newton-Raphson steps, q = a' * r,
// one residual correction (v_div_fmas: an FMA, rescaled when v_div_scale set
// its flag) and v_div_fixup (special operands, sign).  v_div_scale changes
// nothing and leaves the flag clear unless an operand is denormal or extreme:
// |a| < 2^-969, a/b outside the normal range, exponent(a) - exponent(b) >= 768
// or |b| >= 2^1022.  Away from those, a / b is the same sequence without the
// two scalings (FastDiv below, bit for bit: checked on the GPU for 1.1e8
// operand pairs with frexp exponents in [-600, 600], none differing).  Zeros,
// infinities and NaNs are settled by v_div_fixup from a and b alone, whatever
// the core computed, so they agree too.  The refined reciprocal depends on b
// alone: divisions by one b share it (rcp_of once, qdiv per numerator).
//
// The sweeps take FastDiv only where a wave-uniform test has bounded every
// operand: a cell's conserved values (cell_in), a face's floored Riemann
// inputs (in riemann_flux) and the Newton pressure at each iteration (pst_in),
// with the launch parameters bounded (Phys::fastdiv).  The exponent bounds
// carried through prim_cell, trace_cell and riemann_flux from those tests
// keep every division operand far from the conditions above (the tightest
// margin, the Newton step's relative change d / (pst + smallpp), stays above
// 2^-950); the test failing on any lane sends the whole wave down the IEEE
// path (IeeeDiv), which is the reference's arithmetic.
struct Rcp {
  double b, r;
};
__device__ __forceinline__ Rcp rcp_of(double b) {
  const double r0 = __builtin_amdgcn_rcp(b);
  const double r1 = __builtin_fma(r0, __builtin_fma(-b, r0, 1.0), r0);
  return Rcp{b, __builtin_fma(r1, __builtin_fma(-b, r1, 1.0), r1)};
}
__device__ __forceinline__ double qdiv(double a, const Rcp& d) {
  const double q = a * d.r;
  return __builtin_amdgcn_div_fixup(__builtin_fma(__builtin_fma(-d.b, q, a), d.r, q), d.b, a);
}
// qdiv without the v_div_fixup, for a numerator that is never -0.  For
// finite, nonzero, in-range operands the fix-up returns the core's quotient
// with the sign sign(a) * sign(b), which the core's nonzero result already
// has; for a = +0 (b > 0) the core gives +0 itself (+0 * r, then
// fma(-b, +0, +0) = +0, fma(+0, r, +0) = +0).  Only a = -0, where the
// fix-up returns -0 and the core +0, needs it.
__device__ __forceinline__ double qdiv_nz(double a, const Rcp& d) {
  const double q = a * d.r;
  return __builtin_fma(__builtin_fma(-d.b, q, a), d.r, q);
}

// Division / square-root policies of the FP64 numerics.  sqrt1/sqrt2 of
// FastDiv take the sqrt_core sequence directly: callers use them only for
// arguments shown positive, finite and at least 2^-767.
struct IeeeDiv {
  struct R {
    double b;
  };
  static __device__ __forceinline__ R rcp(double b) { return R{b}; }
  static __device__ __forceinline__ double q(double a, const R& d) { return a / d.b; }
  static __device__ __forceinline__ double div(double a, double b) { return a / b; }
  // (numerators that are never -0: see qdiv_nz)
  static __device__ __forceinline__ double qn(double a, const R& d) { return a / d.b; }
  static __device__ __forceinline__ double divn(double a, double b) { return a / b; }
  static __device__ __forceinline__ double sqrt1(double x) { return sqrt_fast(x); }
  static __device__ __forceinline__ void sqrt2(double x, double y, double& sx, double& sy) {
    sqrt_fast2(x, y, sx, sy);
  }
};
struct FastDiv {
  using R = Rcp;
  static __device__ __forceinline__ R rcp(double b) { return rcp_of(b); }
  static __device__ __forceinline__ double q(double a, const R& d) { return qdiv(a, d); }
  static __device__ __forceinline__ double div(double a, double b) { return qdiv(a, rcp_of(b)); }
  static __device__ __forceinline__ double qn(double a, const R& d) { return qdiv_nz(a, d); }
  static __device__ __forceinline__ double divn(double a, double b) {
    return qdiv_nz(a, rcp_of(b));
  }
  static __device__ __forceinline__ double sqrt1(double x) { return sqrt_core(x); }
  static __device__ __forceinline__ void sqrt2(double x, double y, double& sx, double& sy) {
    sx = sqrt_core(x);
    sy = sqrt_core(y);
  }
};

// Operand tests (per lane; callers take wave_all of them).  Comparisons fail
// for NaN, so infinities and NaNs never pass.
//   in_pos(x): x in [2^-41, 2^40)              (floored densities, pressures)
//   in_zero(x): x = +-0 or |x| in [2^-41, 2^40) (velocities, momenta, energy)
//   pst_in(x): x = +-0 or |x| in [2^-151, 2^150) (the Newton pressure)
__device__ __forceinline__ bool in_pos(double x) { return x >= 0x1p-41 && x < 0x1p40; }
__device__ __forceinline__ bool in_zero(double x) {
  const double a = fabs(x);
  return a < 0x1p40 && (a >= 0x1p-41 || a == 0.0);
}
__device__ __forceinline__ bool pst_in(double x) {
  const double a = fabs(x);
  return a < 0x1p150 && (a >= 0x1p-151 || a == 0.0);
}

// A cell whose primitives and trace may take FastDiv: density (after the
// smallr floor), momenta and energy as in_pos / in_zero.
template <typename S>
__device__ __forceinline__ bool cell_in(S U0, S U1, S U2, S U3, const Phys<S>& ph) {
  if constexpr (sizeof(S) == 8) {
    return ph.fastdiv && in_pos(maxm(U0, ph.smallr)) && in_zero(U1) && in_zero(U2) &&
           in_zero(U3);
  } else {
    return false;
  }
}

// Division and square root.  FP64: the IEEE operations themselves (what the
// reference computes).  FP32 (mixed precision only): the hardware reciprocal
// and square root (v_rcp_f32 / v_sqrt_f32, about 1 ulp each) instead of the
// correctly rounded multi-instruction sequences.
__device__ __forceinline__ double xdiv(double a, double b) { return a / b; }
__device__ __forceinline__ float xdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
// c + a/b: in FP64 exactly as written (the quotient rounded, then the sum);
// in FP32 the quotient a * rcp(b) fuses into the sum (one v_fma_f32).
__device__ __forceinline__ double add_div(double c, double a, double b) { return c + a / b; }
__device__ __forceinline__ float add_div(float c, float a, float b) {
  return __builtin_fmaf(a, __builtin_amdgcn_rcpf(b), c);
}
__device__ __forceinline__ double xsqrt(double a) { return sqrt_fast(a); }
__device__ __forceinline__ float xsqrt(float a) { return __builtin_amdgcn_sqrtf(a); }
__device__ __forceinline__ void xsqrt2(double a, double b, double& sa, double& sb) {
  sqrt_fast2(a, b, sa, sb);
}
__device__ __forceinline__ void xsqrt2(float a, float b, float& sa, float& sb) {
  sa = __builtin_amdgcn_sqrtf(a);
  sb = __builtin_amdgcn_sqrtf(b);
}

// FP64 division and square roots of the mixed form's Newton iteration: the
// hardware estimates v_rcp_f64 / v_rsq_f64 refined by one Newton-Raphson (or
// Goldschmidt) step each, relative error about 1e-15 -- far below the
// iteration's 1e-6 tolerance and the FP32 rounding of its result.
__device__ __forceinline__ double rcp_nr(double b) {
  const double y = __builtin_amdgcn_rcp(b);
  return __builtin_fma(__builtin_fma(-b, y, 1.0), y, y);
}
__device__ __forceinline__ double div_nr(double a, double b) { return a * rcp_nr(b); }
// sqrt(x) and 1/sqrt(x) for x > 0 from one v_rsq_f64 estimate g: with
// s = x g and the residual r = 1/2 - s g/2, one Goldschmidt step gives
// s + s r and g + g r.
__device__ __forceinline__ void sqrt_rsq_nr(double x, double& s, double& is) {
  const double g = __builtin_amdgcn_rsq(x);
  const double s0 = x * g;
  const double r = __builtin_fma(-(g * 0.5), s0, 0.5);
  s = __builtin_fma(s0, r, s0);
  is = __builtin_fma(g, r, g);
}
// constoprim.c:62-78 + equation_of_state.c:64-81 for one cell.  U is in sweep
// slot order (slot 1 = normal momentum).  FAST (FP64, for cells that passed
// cell_in): FastDiv, the four divisions by rho sharing one reciprocal; the
// sound-speed argument gamma*p/rho is then in [2^-183, 2^163].  UABS: the
// caller uses u and v only through |u|, |v| and their squares, so their
// quotients may lose the sign of a zero (qn).
template <typename R, bool FAST = false, bool UABS = false>
__device__ __forceinline__ Prim<R> prim_cell(R U0, R U1, R U2, R U3, const Phys<R>& ph) {
  Prim<R> o;
  o.rinv = (R) 0;
  o.cinv = (R) 0;
  const R half = (R) 0.5;
  R r = maxm(U0, ph.smallr);
  R u, v, e, p;
  if constexpr (sizeof(R) == 8) {
    using P = std::conditional_t<FAST, FastDiv, IeeeDiv>;
    const auto rr = P::rcp(r);
    if constexpr (FAST) o.rinv = rr.r;
    if constexpr (UABS) {
      u = P::qn(U1, rr);
      v = P::qn(U2, rr);
    } else {
      u = P::q(U1, rr);
      v = P::q(U2, rr);
    }
    R eken = half * (u * u + v * v);
    // (a +-0 quotient, whose sign qn may lose, meets eken: if eken = +0, e is
    // +-0 and p = max(gm1*r*e, r*smallp) = r*smallp for either sign)
    e = P::qn(U3, rr) - eken;
    p = ph.gm1 * r * e;
    p = maxm(p, r * ph.smallp);
    o.c = P::sqrt1(P::qn(ph.gamma * p, rr));  // gamma*p > 0
  } else {
    // FP32 (mixed form): one reciprocal of rho for the quotients, and c and
    // 1/c (the trace's) from one v_rsq_f32 estimate
    const R ir = __builtin_amdgcn_rcpf(r);
    u = U1 * ir;
    v = U2 * ir;
    R eken = half * (u * u + v * v);
    e = __builtin_fmaf(U3, ir, -eken);
    p = ph.gm1 * r * e;
    p = maxm(p, r * ph.smallp);
    const R x = ph.gamma * p * ir;  // > 0
    o.cinv = __builtin_amdgcn_rsqf(x);
    o.c = x * o.cinv;
    o.rinv = ir;
  }
  o.q.r = r;
  o.q.u = u;
  o.q.v = v;
  o.q.p = p;
  return o;
}

// Wave-uniform: may every cell of this wave take prim_cell / trace_cell with
// FAST (FP64 only)?
template <typename S>
__device__ __forceinline__ bool wave_cell_in(S U0, S U1, S U2, S U3, const Phys<S>& ph) {
  if constexpr (sizeof(S) == 8) {
    return wave_all(cell_in<S>(U0, U1, U2, U3, ph));
  } else {
    return false;
  }
}

// ---- FP64 fast form (kFast) ---------------------------------------------------
//
// Quotients and square roots from the hardware estimates v_rcp_f64 and
// v_rsq_f64 (relative error about 2^-24, measured on MI355X) refined by
// Newton-Raphson or coupled Goldschmidt steps instead of the correctly
// rounded sequences: two steps (rcp2, sqrt_rsq2, sqrt_1ulp) leave a result
// within about an ulp, and divisions by one denominator share its
// reciprocal.  Inside the Newton pressure iteration one step (rcp_nr,
// sqrt_rsq_nr: about 2^-48) is enough except in the last iteration the
// limit allows: an error in an iterate is damped by the next Newton step, so
// p* carries only the error of its last correction, below 1e-6 of p* when
// the face converged, and a face stopped by the limit takes that last step
// with the two-step estimates.

// 1/b within about an ulp: the estimate and two Newton-Raphson steps.
__device__ __forceinline__ double rcp2(double b) {
  const double y0 = __builtin_amdgcn_rcp(b);
  const double y1 = __builtin_fma(__builtin_fma(-b, y0, 1.0), y0, y0);
  return __builtin_fma(__builtin_fma(-b, y1, 1.0), y1, y1);
}
// sqrt(x) and 1/sqrt(x) for x > 0, both within about an ulp: two coupled
// Goldschmidt steps from one v_rsq_f64 estimate g (s = x g, h = g / 2,
// r = 1/2 - s h, then s += s r, h += h r).
__device__ __forceinline__ void sqrt_rsq2(double x, double& s, double& is) {
  const double g = __builtin_amdgcn_rsq(x);
  const double s0 = x * g, h0 = 0.5 * g;
  const double r0 = __builtin_fma(-h0, s0, 0.5);
  const double s1 = __builtin_fma(s0, r0, s0), h1 = __builtin_fma(h0, r0, h0);
  const double r1 = __builtin_fma(-h1, s1, 0.5);
  s = __builtin_fma(s1, r1, s1);
  is = 2.0 * __builtin_fma(h1, r1, h1);
}
// sqrt(x) for x >= 0 within about an ulp: one coupled step, then a residual
// correction (sqrt(+0) gives NaN here; callers clamp the result with fmax,
// which returns the other operand for a NaN).
__device__ __forceinline__ double sqrt_1ulp(double x) {
  const double g = __builtin_amdgcn_rsq(x);
  const double s0 = x * g, h0 = 0.5 * g;
  const double r0 = __builtin_fma(-h0, s0, 0.5);
  const double s1 = __builtin_fma(s0, r0, s0), h1 = __builtin_fma(h0, r0, h0);
  return __builtin_fma(__builtin_fma(-s1, s1, x), h1, s1);
}

// constoprim.c:62-78 + equation_of_state.c:64-81 for one cell, fast form:
// one reciprocal of rho for the four quotients, and c and 1/c (for the
// trace) from one estimate.
__device__ __forceinline__ Prim<double> prim_fast(double U0, double U1, double U2, double U3,
                                                  const Phys<double>& ph) {
  Prim<double> o;
  // (fmax: one v_max_f64; the same values as the reference's MAX here, whose
  // second operand is never NaN)
  const double r = fmax(U0, ph.smallr);
  const double ir = rcp2(r);
  const double u = U1 * ir, v = U2 * ir;
  const double eken = 0.5 * (u * u + v * v);
  const double e = U3 * ir - eken;
  double p = ph.gm1 * r * e;
  p = fmax(p, r * ph.smallp);
  sqrt_rsq2(ph.gamma * p * ir, o.c, o.cinv);  // gamma p / rho > 0
  o.q = Q4<double>{r, u, v, p};
  o.rinv = ir;
  return o;
}

// Courant numbers of one cell in grid slot order (compute_deltat.c:71-88 +
// courantOnXY.c:115-121): returns c + |vx| and c + |vy|.  FD (the form):
// kGated -- the primitives take FastDiv when every active lane's cell passes
// cell_in; kFast -- prim_fast.
template <typename R, int FD = kIeee>
__device__ __forceinline__ void courant_cell(R U0, R U1, R U2, R U3, const Phys<R>& ph, R& cx,
                                             R& cy) {
  Prim<R> o;
  if constexpr (FD == kFast && sizeof(R) == 8) {
    o = prim_fast(U0, U1, U2, U3, ph);
  } else {
    o = FD == kGated && wave_cell_in<R>(U0, U1, U2, U3, ph)
            ? prim_cell<R, true, true>(U0, U1, U2, U3, ph)
            : prim_cell<R, false>(U0, U1, U2, U3, ph);
  }
  cx = o.c + fabs(o.q.u);
  cy = o.c + fabs(o.q.v);
}

// slope.c:71-90 with slope_type == 1 as a median.  There dc = (dl + dr) / 2,
// and where dl and dr have one sign fl(|dl + dr|) / 2 >= min(|dl|, |dr|)
// (rounding is monotone), so the limiter sg * fmin(keep * min(|dl|, |dr|),
// |dc|) is min(|dl|, |dr|) with their sign, and zero where their signs differ
// or one is zero: median(dl, dr, 0).  The same value as the reference's
// except for the sign of a zero slope, and where dl * dr underflows to 0
// (both below ~1e-162 in FP64, ~4e-23 in FP32), where the reference's slope
// is 0 and this one the smaller of the two.  Used by the mixed and fast forms
// only; the FP64 form keeps the reference's operations.
__device__ __forceinline__ float slope1_med(float dl, float dr) {
  return __builtin_amdgcn_fmed3f(dl, dr, 0.0f);
}
__device__ __forceinline__ double slope1_med(double dl, double dr) {
  return fmax(fmin(dl, dr), fmin(fmax(dl, dr), 0.0));
}

// slope.c:71-90 (non-OLDSTYLE limiter) for one variable.  ST1 specialises
// slope_type == 1, where st*x == x and x/st == x exactly (in FP32, the
// median form above).
template <typename R, bool ST1>
__device__ __forceinline__ R slope1(R qm, R q0, R qp, R st) {
  const R half = (R) 0.5;
  R dl, dr, dc;
  if constexpr (ST1 && sizeof(R) == 4) return slope1_med(q0 - qm, qp - q0);
  if (ST1) {
    dl = q0 - qm;
    dr = qp - q0;
    dc = half * (dl + dr);
  } else {
    dl = st * (q0 - qm);
    dr = st * (qp - q0);
    dc = xdiv(half * (dl + dr), st);
  }
  R sg = (dc > (R) 0) ? (R) 1 : (R) -1;
  R keep = ((dl * dr) <= (R) 0) ? (R) 0 : (R) 1;
  R t1 = fmin(fabs(dl), fabs(dr));
  return sg * fmin(keep * t1, fabs(dc));
}

// slope.c:71-90 in the FP64 fast form: for slope_type 1 the median form
// (slope1_med); otherwise the same limiter with the sign copied from dc (a
// zero slope keeps the sign of a zero dc, where the reference's sg = -1 gives
// -0) and dc scaled by 1/slope_type.
template <bool ST1>
__device__ __forceinline__ double slope1_fast(double qm, double q0, double qp, double st,
                                              double ist) {
  if constexpr (ST1) {
    return slope1_med(q0 - qm, qp - q0);
  } else {
    const double dl = st * (q0 - qm);
    const double dr = st * (qp - q0);
    const double dc = (0.5 * (dl + dr)) * ist;
    const double m = (dl * dr <= 0.0) ? 0.0 : fmin(fmin(fabs(dl), fabs(dr)), fabs(dc));
    return copysign(m, dc);
  }
}

template <bool ST1>
__device__ __forceinline__ Q4<double> slope4_fast(const Q4<double>& qm, const Q4<double>& q0,
                                                  const Q4<double>& qp, const Phys<double>& ph) {
  const double st = ph.slope_type, ist = ph.slope_inv;
  return Q4<double>{slope1_fast<ST1>(qm.r, q0.r, qp.r, st, ist),
                    slope1_fast<ST1>(qm.u, q0.u, qp.u, st, ist),
                    slope1_fast<ST1>(qm.v, q0.v, qp.v, st, ist),
                    slope1_fast<ST1>(qm.p, q0.p, qp.p, st, ist)};
}

template <typename R, bool ST1>
__device__ __forceinline__ Q4<R> slope4(const Q4<R>& qm, const Q4<R>& q0, const Q4<R>& qp,
                                        R st) {
  Q4<R> d;
  d.r = slope1<R, ST1>(qm.r, q0.r, qp.r, st);
  d.u = slope1<R, ST1>(qm.u, q0.u, qp.u, st);
  d.v = slope1<R, ST1>(qm.v, q0.v, qp.v, st);
  d.p = slope1<R, ST1>(qm.p, q0.p, qp.p, st);
  return d;
}

// The limited slopes of a cell in form FD (slope4, or slope4_fast for kFast).
template <typename R, bool ST1, int FD>
__device__ __forceinline__ Q4<R> slope_form(const Q4<R>& qm, const Q4<R>& q0, const Q4<R>& qp,
                                            const Phys<R>& ph) {
  if constexpr (FD == kFast && sizeof(R) == 8) return slope4_fast<ST1>(qm, q0, qp, ph);
  else return slope4<R, ST1>(qm, q0, qp, ph.slope_type);
}

// Trace constants of trace.c:63-80.
template <typename R>
struct TraceK {
  R dtdx, zerol, zeror, project;
};

template <typename R>
__device__ __forceinline__ TraceK<R> make_tracek(R dtdx, int scheme) {
  TraceK<R> k;
  k.dtdx = dtdx;
  k.zerol = (R) 0;
  k.zeror = (R) 0;
  k.project = (R) 0;
  if (scheme == 1) {
    k.zerol = xdiv(-(R) 100, dtdx);
    k.zeror = xdiv((R) 100, dtdx);
    k.project = (R) 1;
  } else if (scheme == 2) {
    k.project = (R) 1;
  }
  return k;
}

// trace.c:85-147 after its four quotients: the MUSCL-Hancock predictor of
// one cell -> qxm (the state at the cell's right face, seen from inside) and
// qxp (its left face).  rOcc = rho/c, OrOcc = c/rho, dprcc = dp/(rho c),
// a0r = drho - dp/c^2, csq = c^2.
template <typename R>
__device__ __forceinline__ void trace_tail(const Q4<R>& q, const Q4<R>& dq, R cc, R csq, R rOcc,
                                           R OrOcc, R dprcc, R a0r, const TraceK<R>& k,
                                           Q4<R>& qxm, Q4<R>& qxp) {
  const R half = (R) 0.5;
  const R one = (R) 1;
  R am = half * (dprcc - dq.u) * rOcc;
  R ap = half * (dprcc + dq.u) * rOcc;
  R a0v = dq.v;
  R upcc = q.u + cc;
  R umcc = q.u - cc;
  // The predictor factors (s * dtdx +- one: the reference's product then sum,
  // the FP64 unit contracting nothing), clamped to
  // +-project where s has crossed zeror / zerol.  With umcc <= u <= upcc no
  // clamp applies on a lane with upcc < zeror and umcc > zerol, so the
  // selects run only when some lane of the wave fails that test (every lane
  // under schemes 2 and 3, where zeror = zerol = 0; a lane whose test fails
  // for a NaN takes the selects, which keep the unclamped NaN).
  R spmR = umcc * k.dtdx + one, sppR = upcc * k.dtdx + one, sp0R = q.u * k.dtdx + one;
  R spmL = umcc * k.dtdx - one, sppL = upcc * k.dtdx - one, sp0L = q.u * k.dtdx - one;
  if (!wave_all(upcc < k.zeror && umcc > k.zerol)) {
    spmR = (umcc >= k.zeror) ? k.project : spmR;
    sppR = (upcc >= k.zeror) ? k.project : sppR;
    sp0R = (q.u >= k.zeror) ? k.project : sp0R;
    spmL = (umcc <= k.zerol) ? -k.project : spmL;
    sppL = (upcc <= k.zerol) ? -k.project : sppL;
    sp0L = (q.u <= k.zerol) ? -k.project : sp0L;
  }
  {
    const R spm = spmR, spp = sppR, sp0 = sp0R;
    R apr = -half * spp * ap;
    R amr = -half * spm * am;
    R azr = -half * sp0 * a0r;
    R azv = -half * sp0 * a0v;
    qxp.r = q.r + (apr + amr + azr);
    qxp.u = q.u + (apr - amr) * OrOcc;
    qxp.v = q.v + (azv);
    qxp.p = q.p + (apr + amr) * csq;
  }
  {
    const R spm = spmL, spp = sppL, sp0 = sp0L;
    R apl = -half * spp * ap;
    R aml = -half * spm * am;
    R azl = -half * sp0 * a0r;
    R azv = -half * sp0 * a0v;
    qxm.r = q.r + (apl + aml + azl);
    qxm.u = q.u + (apl - aml) * OrOcc;
    qxm.v = q.v + (azv);
    qxm.p = q.p + (apl + aml) * csq;
  }
}

// trace.c:85-147 for one cell in the FP64 form (quotients as the form
// computes them, then trace_tail; the fast and mixed forms take trace_fast).
// FAST: FastDiv, for cells whose own and neighbour values (the slope inputs)
// passed cell_in.  RINV: rinv is prim_cell's reciprocal of q.r, reused for
// cc / q.r.
template <typename R, bool FAST = false, bool RINV = false>
__device__ __forceinline__ void trace_cell(const Q4<R>& q, const Q4<R>& dq, R cc,
                                           const TraceK<R>& k, Q4<R>& qxm, Q4<R>& qxp,
                                           R rinv = (R) 0) {
  static_assert(sizeof(R) == 8, "the FP32 form takes trace_fast");
  const R csq = cc * cc;
  using P = std::conditional_t<FAST, FastDiv, IeeeDiv>;
  const R rOcc = P::divn(q.r, cc);  // rho, c > 0
  R OrOcc;
  if constexpr (FAST && RINV) OrOcc = qdiv_nz(cc, Rcp{q.r, rinv});
  else OrOcc = P::divn(cc, q.r);
  const R dprcc = P::div(dq.p, q.r * cc);
  const R a0r = dq.r - P::div(dq.p, csq);
  trace_tail<R>(q, dq, cc, csq, rOcc, OrOcc, dprcc, a0r, k, qxm, qxp);
}

// ---- FP64 Riemann solver (the FP64 form) ------------------------------------

// One Newton step of riemann.c:160-201 for one face; returns true when the
// face has converged (the reference's goon = 0).  P: IeeeDiv, or FastDiv when
// pst, the floored inputs and the launch parameters are in range.
template <typename P>
__device__ __forceinline__ bool newton_step(double& pst, double pl, double pr, double cl,
                                            double cr, double ul, double ur,
                                            const typename P::R& rpl, const typename P::R& rpr,
                                            double g6, double smallpp) {
  // Numerators: pst - p is +0 when pst == p (never -0, p > 0); the wave
  // speeds and q's are positive; rel (which may be +-0 for d = +-0) is only
  // compared by magnitude.
  double wwl, wwr;
  P::sqrt2(cl * (1.0 + P::qn(g6 * (pst - pl), rpl)), cr * (1.0 + P::qn(g6 * (pst - pr), rpr)),
           wwl, wwr);
  const double swwl = wwl * wwl;
  const double qlw = P::divn(2.0 * wwl * swwl, swwl + cl);
  const double qrw = P::divn(2.0 * wwr * (wwr * wwr), (wwr * wwr) + cr);
  const double usl = ul - P::divn(pst - pl, wwl);
  const double usr = ur + P::divn(pst - pr, wwr);
  const double t = (P::divn(qrw * qlw, qrw + qlw) * (usl - usr));
  const double d = (t > (-pst)) ? t : (-pst);
  pst += d;
  const double rel = P::divn(d, pst + smallpp);
  const double arel = (rel > 0) ? rel : -rel;
  return !(arel > 1e-6);
}

// Star state, sampling at x/t = 0 and flux (riemann.c:211-276, cmpflx.c:67-90)
// for the converged pressure pst.  With FastDiv the square roots take the
// core alone (cstar's argument, 0 when pst is, returns +0 directly).
template <typename P>
__device__ __forceinline__ Q4<double> star_flux(double pst, const Q4<double>& ql,
                                                const Q4<double>& qr, double rl, double ul,
                                                double pl, double rr, double ur, double pr,
                                                double cl, double cr, const typename P::R& rpl,
                                                const typename P::R& rpr,
                                                const Phys<double>& ph) {
  const double half = 0.5, one = 1.0, g6 = ph.gamma6_d;
  const double ps = pst;
  double wr, wl;
  // Numerators that are never -0 (qn / divn): differences p - ps and ps - p
  // (p > 0), gamma*po, wo, the fabs'd gamma*ps / rstar, and spout + spin
  // (neither term is -0: co, cstar >= smallc > 0 and ushock = wo/ro - sg*uo
  // with wo/ro > 0).  rstar's outer quotient keeps the fix-up (its
  // denominator is not bounded away from 0).
  P::sqrt2(cr * (one + P::qn(g6 * (ps - pr), rpr)), cl * (one + P::qn(g6 * (ps - pl), rpl)), wr,
           wl);
  // ((ul + (pl-ps)/wl) + ur) - (pr-ps)/wr, the reference's left-to-right sum
  const double ustar = half * (((ul + P::divn(pl - ps, wl)) + ur) - P::divn(pr - ps, wr));
  const bool left = ustar > 0.0;
  const double sg = left ? 1.0 : -1.0;
  const double ro = left ? rl : rr;
  const double uo = left ? ul : ur;
  const double po = left ? pl : pr;
  const double wo = left ? wl : wr;
  double rstar = P::div(ro, one + P::divn(ro * (po - ps), wo * wo));
  rstar = fmax(rstar, ph.smallr);
  const auto rro = P::rcp(ro);
  double co, cstar;  // (riemann.c computes co before rstar; neither depends on the other)
  if constexpr (std::is_same_v<P, FastDiv>) {
    // both arguments bounded; cstar's is +0 when ps is (sqrt(+0) = +0)
    const double xs = fabs(P::divn(ph.gamma * ps, rstar));
    co = sqrt_core(fabs(P::qn(ph.gamma * po, rro)));
    cstar = xs == 0.0 ? xs : sqrt_core(xs);
  } else {
    sqrt_fast2(fabs(P::q(ph.gamma * po, rro)), fabs(P::div(ph.gamma * ps, rstar)), co, cstar);
  }
  co = fmax(ph.smallc, co);
  cstar = fmax(ph.smallc, cstar);
  double spout = co - sg * uo;
  double spin = cstar - sg * ustar;
  const double ushock = P::qn(wo, rro) - sg * uo;
  if (ps >= po) {
    spin = ushock;
    spout = ushock;
  }
  const double scr = fmax(spout - spin, ph.smallc + fabs(spout + spin));
  double frac = (one + P::divn(spout + spin, scr)) * half;
  frac = fmax(0.0, fmin(one, frac));
  double gr, gu, gp;
  if (spout < 0.0) {
    gr = ro;
    gu = uo;
    gp = po;
  } else if (spin > 0.0) {
    gr = rstar;
    gu = ustar;
    gp = ps;
  } else {
    gr = frac * rstar + (one - frac) * ro;
    gu = frac * ustar + (one - frac) * uo;
    gp = frac * ps + (one - frac) * po;
  }
  const double gv = left ? ql.v : qr.v;
  Q4<double> f;
  const double m = gr * gu;
  f.r = m;
  f.u = m * gu + gp;
  f.v = m * gv;
  const double ekin = half * gr * (gu * gu + gv * gv);
  const double etot = gp * ph.entho + ekin;
  f.p = gu * (etot + gp);
  return f;
}

// First guess, Newton iteration and star state of one face in FP64.  Each
// block takes FastDiv when its operands are bounded -- the floored inputs
// passed their tests (ff) and, for a Newton step and the star state, the
// wave's pressures pass pst_in -- and IeeeDiv otherwise.
template <bool G>
__device__ __forceinline__ Q4<double> riemann_d(const Q4<double>& ql, const Q4<double>& qr,
                                                double rl, double ul, double pl, double rr,
                                                double ur, double pr, double cl, double cr,
                                                bool ff, const Phys<double>& ph) {
  ff = G && ff;  // G false: the IEEE code alone
  const double g6 = ph.gamma6_d;
  double wl0, wr0, pst;
  Rcp rpl{pl, 0.0}, rpr{pr, 0.0};
  if (G && ff) {
    FastDiv::sqrt2(cl, cr, wl0, wr0);
    // numerator: a positive sum plus a product, never -0
    pst = fmax(FastDiv::divn((wr0 * pl + wl0 * pr) + wl0 * wr0 * (ul - ur), wl0 + wr0), 0.0);
    rpl = rcp_of(pl);
    rpr = rcp_of(pr);
  } else {
    IeeeDiv::sqrt2(cl, cr, wl0, wr0);
    pst = fmax(((wr0 * pl + wl0 * pr) + wl0 * wr0 * (ul - ur)) / (wl0 + wr0), 0.0);
  }
  for (int it = 0; it < ph.niter; it++) {
    bool done;
    if (G && ff && wave_all(pst_in(pst)))
      done = newton_step<FastDiv>(pst, pl, pr, cl, cr, ul, ur, rpl, rpr, g6, ph.smallpp_d);
    else
      done = newton_step<IeeeDiv>(pst, pl, pr, cl, cr, ul, ur, IeeeDiv::R{pl}, IeeeDiv::R{pr},
                                  g6, ph.smallpp_d);
    if (done) break;
  }
  if (G && ff && wave_all(pst_in(pst)))
    return star_flux<FastDiv>(pst, ql, qr, rl, ul, pl, rr, ur, pr, cl, cr, rpl, rpr, ph);
  return star_flux<IeeeDiv>(pst, ql, qr, rl, ul, pl, rr, ur, pr, cl, cr, IeeeDiv::R{pl},
                            IeeeDiv::R{pr}, ph);
}

// trace.c:85-147 in the fast and mixed forms: the four quotients from 1/rho
// (ir) and 1/c (ic), and the predictor's products regrouped: -half * sp *
// (half * x * rOcc) is sp * (x * (-rho / 4c)), and each face's sums fold into
// multiply-adds.
template <typename R>
__device__ __forceinline__ void trace_fast(const Q4<R>& q, const Q4<R>& dq, R cc, R ic, R ir,
                                           const TraceK<R>& k, Q4<R>& qxm, Q4<R>& qxp) {
  const R one = (R) 1;
  const R csq = cc * cc;
  const R dpic = dq.p * ic;
  const R dprcc = dpic * ir;              // dp / (rho c)
  const R a0r = dq.r - dpic * ic;         // drho - dp / c^2
  const R OrOcc = cc * ir;                // c / rho
  const R qr4 = (R) -0.25 * (q.r * ic);   // -rho / (4c)
  const R nap = (dprcc + dq.u) * qr4;     // -ap / 2
  const R nam = (dprcc - dq.u) * qr4;     // -am / 2
  const R na0 = (R) -0.5 * a0r, nav = (R) -0.5 * dq.v;
  const R upcc = q.u + cc, umcc = q.u - cc;
  R spmR = umcc * k.dtdx + one, sppR = upcc * k.dtdx + one, sp0R = q.u * k.dtdx + one;
  R spmL = umcc * k.dtdx - one, sppL = upcc * k.dtdx - one, sp0L = q.u * k.dtdx - one;
  if (!wave_all(upcc < k.zeror && umcc > k.zerol)) {  // as in trace_tail
    spmR = (umcc >= k.zeror) ? k.project : spmR;
    sppR = (upcc >= k.zeror) ? k.project : sppR;
    sp0R = (q.u >= k.zeror) ? k.project : sp0R;
    spmL = (umcc <= k.zerol) ? -k.project : spmL;
    sppL = (upcc <= k.zerol) ? -k.project : sppL;
    sp0L = (q.u <= k.zerol) ? -k.project : sp0L;
  }
  {
    const R amr = spmR * nam;
    const R t = sppR * nap + amr;  // apr + amr
    qxp.r = q.r + (t + sp0R * na0);
    qxp.u = q.u + (sppR * nap - amr) * OrOcc;
    qxp.v = q.v + sp0R * nav;
    qxp.p = q.p + t * csq;
  }
  {
    const R aml = spmL * nam;
    const R t = sppL * nap + aml;
    qxm.r = q.r + (t + sp0L * na0);
    qxm.u = q.u + (sppL * nap - aml) * OrOcc;
    qxm.v = q.v + sp0L * nav;
    qxm.p = q.p + t * csq;
  }
}

// riemann.c:128-276 + cmpflx.c:67-90 for one face, fast form.  The first
// guess and the Newton iteration are the reference's recurrence with fewer
// divisions (as in the mixed form, here in FP64 throughout): 1/pl and 1/pr
// are loop invariants, 1/w comes from the estimate that gives w, with
// q = a / b (a = 2 w^3, b = w^2 + c) qr ql / (qr + ql) = al ar / (al br +
// ar bl) is one division, and |d / (p* + smallpp)| > 1e-6 is tested as
// |d| > 1e-6 (p* + smallpp).  In the star state, sqrt(gamma po / ro) is
// sqrt(gamma po ro) / ro with sqrt(gamma po ro) from the first guess; with
// rstar = ro / X, one reciprocal of ro X gives 1/ro, 1/X and 1/rstar (1/smallr
// where the floor holds); wo / ro and (spout + spin) / scr are products with
// reciprocals.
__device__ __forceinline__ Q4<double> riemann_fast(const Q4<double>& ql, const Q4<double>& qr,
                                                   const Phys<double>& ph) {
  const double half = 0.5, one = 1.0;
  const double rl = fmax(ql.r, ph.smallr);
  const double ul = ql.u;
  const double pl = fmax(ql.p, rl * ph.smallp);
  const double rr = fmax(qr.r, ph.smallr);
  const double ur = qr.u;
  const double pr = fmax(qr.p, rr * ph.smallp);
  const double cl = ph.gamma * pl * rl;
  const double cr = ph.gamma * pr * rr;

  // first guess (riemann.c:145-150); wl0 = sqrt(gamma pl rl) serves the star
  // state's sound speed too, so it is refined to an ulp
  double wl0, iwl0, wr0, iwr0;
  sqrt_rsq2(cl, wl0, iwl0);
  sqrt_rsq2(cr, wr0, iwr0);
  double pst = fmax(((wr0 * pl + wl0 * pr) + wl0 * wr0 * (ul - ur)) * rcp2(wl0 + wr0), 0.0);

  // Newton iteration (riemann.c:160-201).  The last iteration the limit
  // allows takes the estimates refined to an ulp: when a face stops there
  // unconverged, that step's correction is not damped by a later one.
  const double g6 = ph.gamma6_d;
  const double ipl = rcp2(pl), ipr = rcp2(pr);
  const double gpl = g6 * ipl, gpr = g6 * ipr;
  for (int it = 0; it < ph.niter; it++) {
    const bool last = it + 1 == ph.niter;  // wave-uniform
    const double xl = cl * (1.0 + (pst - pl) * gpl), xr = cr * (1.0 + (pst - pr) * gpr);
    double wwl, iwl, wwr, iwr;
    if (!last) {
      sqrt_rsq_nr(xl, wwl, iwl);
      sqrt_rsq_nr(xr, wwr, iwr);
    } else {
      sqrt_rsq2(xl, wwl, iwl);
      sqrt_rsq2(xr, wwr, iwr);
    }
    const double swl = wwl * wwl, swr = wwr * wwr;
    const double al = 2.0 * wwl * swl, bl = swl + cl;
    const double ar = 2.0 * wwr * swr, br = swr + cr;
    const double usl = ul - (pst - pl) * iwl;
    const double usr = ur + (pst - pr) * iwr;
    const double den = al * br + ar * bl;
    const double rden = last ? rcp2(den) : rcp_nr(den);
    const double t = ((al * ar) * rden) * (usl - usr);
    const double d = fmax(t, -pst);  // (t > -pst ? t : -pst for every pst that is not NaN)
    pst += d;
    if (!(fabs(d) > 1e-6 * (pst + ph.smallpp_d))) break;
  }

  // star state and sampling at x/t = 0 (riemann.c:211-276)
  const double ps = pst;
  double wl, iwl, wr, iwr;
  sqrt_rsq2(cr * (one + (ps - pr) * gpr), wr, iwr);
  sqrt_rsq2(cl * (one + (ps - pl) * gpl), wl, iwl);
  // ((ul + (pl-ps)/wl) + ur) - (pr-ps)/wr, the reference's left-to-right sum
  const double ustar = half * (((ul + (pl - ps) * iwl) + ur) - (pr - ps) * iwr);
  const bool left = ustar > 0.0;
  const double sg = left ? 1.0 : -1.0;
  const double ro = left ? rl : rr;
  const double uo = left ? ul : ur;
  const double po = left ? pl : pr;
  const double wo = left ? wl : wr;
  const double iwo = left ? iwl : iwr;
  const double wo0 = left ? wl0 : wr0;
  // rstar = ro / X: one reciprocal of ro X gives 1/ro, 1/X and 1/rstar
  const double X = one + ro * (po - ps) * (iwo * iwo);
  const double y = rcp2(ro * X);
  const double iro = X * y;
  const double rs0 = ro * (ro * y);  // ro / X
  const double rstar = fmax(rs0, ph.smallr);
  const double irstar = rs0 > ph.smallr ? X * iro : ph.smallr_inv;
  double co = wo0 * iro;  // sqrt(gamma po / ro)
  double cstar = sqrt_1ulp(ph.gamma * ps * irstar);
  co = fmax(ph.smallc, co);
  cstar = fmax(ph.smallc, cstar);
  double spout = co - sg * uo;
  double spin = cstar - sg * ustar;
  const double ushock = wo * iro - sg * uo;
  if (ps >= po) {
    spin = ushock;
    spout = ushock;
  }
  const double scr = fmax(spout - spin, ph.smallc + fabs(spout + spin));
  double frac = (one + (spout + spin) * rcp2(scr)) * half;
  frac = fmax(0.0, fmin(one, frac));
  double gr, gu, gp;
  if (spout < 0.0) {
    gr = ro;
    gu = uo;
    gp = po;
  } else if (spin > 0.0) {
    gr = rstar;
    gu = ustar;
    gp = ps;
  } else {
    gr = frac * rstar + (one - frac) * ro;
    gu = frac * ustar + (one - frac) * uo;
    gp = frac * ps + (one - frac) * po;
  }
  const double gv = left ? ql.v : qr.v;
  Q4<double> f;
  const double m = gr * gu;
  f.r = m;
  f.u = m * gu + gp;
  f.v = m * gv;
  const double ekin = half * gr * (gu * gu + gv * gv);
  const double etot = gp * ph.entho + ekin;
  f.p = gu * (etot + gp);
  return f;
}

// The mixed-precision form (R = float) of riemann_flux: FP32 inputs, first
// guess and star state, the Newton iteration in FP64 with the refined
// hardware estimates (div_nr, rcp_nr, sqrt_rsq_nr).
template <typename R>
__device__ __forceinline__ Q4<R> riemann_flux_mixed(const Q4<R>& ql, const Q4<R>& qr,
                                                    const Phys<R>& ph) {
  const R half = (R) 0.5;
  const R one = (R) 1;
  R rl = fmax(ql.r, ph.smallr);
  R ul = ql.u;
  R pl = fmax(ql.p, rl * ph.smallp);
  R rr = fmax(qr.r, ph.smallr);
  R ur = qr.u;
  R pr = fmax(qr.p, rr * ph.smallp);
  R cl = ph.gamma * pl * rl;
  R cr = ph.gamma * pr * rr;

  // First guess and Newton iteration for p* (riemann.c:145-201), in double.
  // The iteration is the reference's recurrence rearranged to fewer
  // divisions: 1/pl and 1/pr are loop invariants; 1/w comes from the same
  // reciprocal square root as w; with ql = al/bl and qr = ar/br (al = 2 wl^3,
  // bl = wl^2 + cl), qr ql / (qr + ql) = al ar / (al br + ar bl) is one
  // division; |d / (p* + smallpp)| > 1e-6 is tested as |d| > 1e-6 (p* +
  // smallpp) (p* >= 0, smallpp > 0).
  // The first guess (riemann.c:145-150) is the iteration's starting point
  // and is formed in FP32: a Newton step squares its relative error (1e-7 ->
  // 1e-14), so the FP64 iterate it reaches is as accurate as from an FP64
  // guess.
  const double dpl = pl, dpr = pr, dcl = cl, dcr = cr, dul = ul, dur = ur;
  R wl0, wr0;
  xsqrt2(cl, cr, wl0, wr0);
  double pst = fmax((double) xdiv((wr0 * pl + wl0 * pr) + wl0 * wr0 * (ul - ur), wl0 + wr0), 0.0);
  const double g6 = ph.gamma6_d;
  const double gpl = g6 * rcp_nr(dpl), gpr = g6 * rcp_nr(dpr);
  for (int it = 0; it < ph.niter; it++) {
    double wwl, iwl, wwr, iwr;
    sqrt_rsq_nr(dcl * (1.0 + (pst - dpl) * gpl), wwl, iwl);
    sqrt_rsq_nr(dcr * (1.0 + (pst - dpr) * gpr), wwr, iwr);
    const double swl = wwl * wwl, swr = wwr * wwr;
    const double al = 2.0 * wwl * swl, bl = swl + dcl;
    const double ar = 2.0 * wwr * swr, br = swr + dcr;
    const double usl = dul - (pst - dpl) * iwl;
    const double usr = dur + (pst - dpr) * iwr;
    const double t = div_nr(al * ar, al * br + ar * bl) * (usl - usr);
    const double d = (t > (-pst)) ? t : (-pst);
    pst += d;
    if (!(fabs(d) > 1e-6 * (pst + ph.smallpp_d))) break;
  }

  // Star state and sampling at x/t = 0 (riemann.c:211-276), with the fast
  // form's shared quotients: w and 1/w from one v_rsq_f32 estimate; with
  // rstar = ro / X, one reciprocal of ro X gives 1/ro, 1/X and 1/rstar (1/smallr
  // where the floor holds); sqrt(gamma po / ro) is sqrt(gamma po ro) / ro,
  // the numerator the first guess's wl0 or wr0.
  // (g6/pl, g6/pr shared with the Newton iteration)
  R ps = (R) pst;
  const R xr = cr * __builtin_fmaf(ps - pr, (R) gpr, one);
  const R xl = cl * __builtin_fmaf(ps - pl, (R) gpl, one);
  const R iwr = __builtin_amdgcn_rsqf(xr), iwl = __builtin_amdgcn_rsqf(xl);
  const R wr = xr * iwr, wl = xl * iwl;
  // ((ul + (pl-ps)/wl) + ur) - (pr-ps)/wr, the reference's left-to-right sum
  R ustar = half * __builtin_fmaf(-(pr - ps), iwr, __builtin_fmaf(pl - ps, iwl, ul) + ur);
  bool left = ustar > (R) 0;
  R sg = left ? (R) 1 : (R) -1;
  R ro = left ? rl : rr;
  R uo = left ? ul : ur;
  R po = left ? pl : pr;
  R wo = left ? wl : wr;
  R iwo = left ? iwl : iwr;
  const R wo0 = left ? wl0 : wr0;
  const R X = __builtin_fmaf(ro * (po - ps), iwo * iwo, one);
  const R y = __builtin_amdgcn_rcpf(ro * X);
  const R iro = X * y;
  const R rs0 = ro * (ro * y);  // ro / X
  const R rstar = fmax(rs0, ph.smallr);
  const R irstar = rs0 > ph.smallr ? X * iro : (R) ph.smallr_inv;
  R co = fmax(ph.smallc, wo0 * iro);
  R cstar = fmax(ph.smallc, xsqrt(fabs(ph.gamma * ps * irstar)));
  R spout = co - sg * uo;
  R spin = cstar - sg * ustar;
  R ushock = wo * iro - sg * uo;
  if (ps >= po) {
    spin = ushock;
    spout = ushock;
  }
  R scr = fmax(spout - spin, ph.smallc + fabs(spout + spin));
  R frac = add_div(one, spout + spin, scr) * half;
  frac = fmax((R) 0, fmin(one, frac));
  R gr, gu, gp;
  if (spout < (R) 0) {
    gr = ro;
    gu = uo;
    gp = po;
  } else if (spin > (R) 0) {
    gr = rstar;
    gu = ustar;
    gp = ps;
  } else {
    gr = frac * rstar + (one - frac) * ro;
    gu = frac * ustar + (one - frac) * uo;
    gp = frac * ps + (one - frac) * po;
  }
  R gv = left ? ql.v : qr.v;

  // cmpflx.c:67-90
  Q4<R> f;
  R m = gr * gu;
  f.r = m;
  f.u = m * gu + gp;
  f.v = m * gv;
  R ekin = half * gr * (gu * gu + gv * gv);
  R etot = gp * ph.entho + ekin;
  f.p = gu * (etot + gp);
  return f;
}

// riemann.c:128-276 for one face (qleft = ql, qright = qr) followed by
// cmpflx.c:67-90: returns the Godunov flux through the face.  FD (FP64, the
// form): kGated -- the range-gated divisions may be used; kIeee -- the IEEE
// code alone; kFast -- riemann_fast.
template <typename R, int FD = kIeee>
__device__ __forceinline__ Q4<R> riemann_flux(const Q4<R>& ql, const Q4<R>& qr,
                                              const Phys<R>& ph) {
  if constexpr (sizeof(R) == 8 && FD == kFast) {
    return riemann_fast(ql, qr, ph);
  } else if constexpr (sizeof(R) == 8) {
    const double rl = fmax(ql.r, ph.smallr);
    const double ul = ql.u;
    const double pl = fmax(ql.p, rl * ph.smallp);
    const double rr = fmax(qr.r, ph.smallr);
    const double ur = qr.u;
    const double pr = fmax(qr.p, rr * ph.smallp);
    const double cl = ph.gamma * pl * rl;
    const double cr = ph.gamma * pr * rr;
    if constexpr (FD == kGated) {
      // the face's floored densities and pressures in_pos, velocities in_zero
      const bool fast = ph.fastdiv && wave_all(in_pos(rl) && in_pos(pl) && in_pos(rr) &&
                                               in_pos(pr) && in_zero(ul) && in_zero(ur));
      return riemann_d<true>(ql, qr, rl, ul, pl, rr, ur, pr, cl, cr, fast, ph);
    } else {
      return riemann_d<false>(ql, qr, rl, ul, pl, rr, ur, pr, cl, cr, false, ph);
    }
  } else {
    return riemann_flux_mixed(ql, qr, ph);
  }
}


// conservar.c:147-197: U_i + (F_{i-1/2} - F_{i+1/2}) * dtdx, accumulated in
// double (exactly the reference expression when R = double).  The fluxes
// come in double: a flux is converted once (exactly) and serves both cells
// of its face.
template <typename R>
__device__ __forceinline__ double update1(R u, double fl, double fr, double dtdx) {
  return (double) u + (fl - fr) * dtdx;
}
template <typename R>
__device__ __forceinline__ Q4<double> to_d4(const Q4<R>& v) {
  return Q4<double>{(double) v.r, (double) v.u, (double) v.v, (double) v.p};
}

}  // namespace hydro
