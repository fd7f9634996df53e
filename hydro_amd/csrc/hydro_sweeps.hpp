// Task-level device code of the fused sweeps, shared by the single-grid
// kernels (hydro_kernels.hip) and the batched kernels (hydro_batch.hip), so
// both run exactly the same arithmetic.
//
// * x sweep: one 64-lane wave owns 64 consecutive cells of a row and updates
//   the middle 60 (kXSeg).  Each lane turns its cell into primitives;
//   neighbours come from DPP wave rotations (v_mov_b32_dpp wave_ror/wave_rol),
//   so slope, trace, Riemann solve, flux and update all stay in registers.
// * y sweep: one lane owns one column and marches down a segment of rows with
//   a three-cell window (primitives in registers; the carried trace state
//   and flux in lane-private LDS slots), two rows per trip, loading each row
//   one row ahead of its use.  A wave covers 64 adjacent columns, so every
//   load/store is one contiguous row piece.
// * Both read one state buffer and write the other (ping-pong).
// * CFL maxima of the written state can be folded into the epilogue: reduced
//   wave -> workgroup -> one 64-bit atomic max per workgroup into one of a
//   set of slots; the bit pattern of a non-negative double orders like its
//   value, so the integer max is the exact floating max.  (A NaN Courant
//   number orders above +inf and takes its slot; the reference's sequential
//   MAX is order-dependent there, so dt on states holding NaN has no single
//   reference value.)
#pragma once
#include "hydro_kernels.hpp"
#include "hydro_numerics.hpp"

namespace hydro {

// ---- grid stores of the y sweep ------------------------------------------------

// The y march writes each cell of its column strip once, in 64-column row
// pieces, and a whole state is far larger than the caches: its stores carry
// the non-temporal hint (global_store ... nt), so the written lines stream out
// instead of displacing the rows the march is still to read (measured on the
// 8192x8192 blast: y sweep 3 % faster in the FP64 fast form; the x sweeps,
// whose row pieces overlap by their halos, and the loads are faster without
// the hint).  The values are unchanged.
template <typename S>
__device__ __forceinline__ void st_stream(S* p, S v) {
  __builtin_nontemporal_store(v, p);
}

// ---- lane exchange -----------------------------------------------------------

// lane l receives the value of lane l-1 (DPP wave_ror:1, so lane 0 receives
// lane 63's); from_right: lane l+1's (wave_rol:1, lane 63 receives lane 0's).
// Every source lane is valid, so each 32-bit half is one v_mov_b32_dpp into a
// fresh register.  The wrapped values reach only lanes 0-1 and 62-63 of a
// segment, which are halo lanes whose results are never stored.
__device__ __forceinline__ double from_left(double x) {
  int lo = __double2loint(x), hi = __double2hiint(x);
  lo = __builtin_amdgcn_mov_dpp(lo, 0x13C, 0xf, 0xf, false);
  hi = __builtin_amdgcn_mov_dpp(hi, 0x13C, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double from_right(double x) {
  int lo = __double2loint(x), hi = __double2hiint(x);
  lo = __builtin_amdgcn_mov_dpp(lo, 0x134, 0xf, 0xf, false);
  hi = __builtin_amdgcn_mov_dpp(hi, 0x134, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float from_left(float x) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x13C, 0xf, 0xf, false));
}
__device__ __forceinline__ float from_right(float x) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x134, 0xf, 0xf, false));
}
template <typename R>
__device__ __forceinline__ Q4<R> from_left4(const Q4<R>& a) {
  return Q4<R>{from_left(a.r), from_left(a.u), from_left(a.v), from_left(a.p)};
}
template <typename R>
__device__ __forceinline__ Q4<R> from_right4(const Q4<R>& a) {
  return Q4<R>{from_right(a.r), from_right(a.u), from_right(a.v), from_right(a.p)};
}

// ---- CFL maximum -------------------------------------------------------------

__device__ __forceinline__ double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = maxm(v, __shfl_xor(v, o));
  return v;
}

// Workgroup max of (cx, cy), then one atomic max each into slot `slot` of
// cour[0..nslot) / cour[nslot..2 nslot).  Every thread of the workgroup must
// call it (it has barriers); it can be called again by the same workgroup.
__device__ __forceinline__ void commit_courant(double cx, double cy, uint64_t* cour, int nslot,
                                               int slot) {
  __shared__ double sx[16], sy[16];
  cx = wave_max(cx);
  cy = wave_max(cy);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sx[w] = cx;
    sy[w] = cy;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = blockDim.x >> 6;
    double mx = sx[0], my = sy[0];
    for (int k = 1; k < nw; k++) {
      mx = maxm(mx, sx[k]);
      my = maxm(my, sy[k]);
    }
    atomicMax((unsigned long long*) &cour[slot], (unsigned long long) __double_as_longlong(mx));
    atomicMax((unsigned long long*) &cour[nslot + slot],
              (unsigned long long) __double_as_longlong(my));
  }
  __syncthreads();
}

// Max over nslot slot pairs (clearing them), then compute_deltat.c:197:
// dt = cfl*dx / MAX(cx, MAX(cy, smallc)).  Called by one whole wave; every
// lane returns the same dt.
__device__ __forceinline__ double wave_reduce_dt(uint64_t* cour, int nslot, double cf, double dx,
                                                 double smallc) {
  double cx = 0.0, cy = 0.0;
  const int lane = threadIdx.x & 63;
  for (int s = lane; s < nslot; s += 64) {
    cx = maxm(cx, __longlong_as_double((long long) cour[s]));
    cy = maxm(cy, __longlong_as_double((long long) cour[nslot + s]));
    cour[s] = 0;
    cour[nslot + s] = 0;
  }
  cx = wave_max(cx);
  cy = wave_max(cy);
  return cf * dx / maxm(cx, maxm(cy, smallc));
}

// ---- x sweep: one 60-cell segment of one row per wave ----------------------

// The wave's index in its workgroup, as a wave-uniform (scalar) value.
__device__ __forceinline__ int wave_id() {
  return __builtin_amdgcn_readfirstlane((int) (threadIdx.x >> 6));
}

// Row segments s0, s0 + stride, s0 + 2*stride, ... of a grid with nseg
// segments per row, as (row, segment) pairs: one division up front, then
// carries.  Every value is wave-uniform; callers keep nseg * rows < 2^31.
struct SegWalk {
  int row, seg;
  __device__ __forceinline__ SegWalk(int s0, int nseg) {
    row = s0 / nseg;
    seg = s0 - row * nseg;
  }
  __device__ __forceinline__ void advance(int stride, int nseg) {
    seg += stride;
    while (seg >= nseg) {
      seg -= nseg;
      row++;
    }
  }
};

// FD, the arithmetic form (hydro_numerics.hpp): kGated -- the range-gated
// FP64 divisions may be used (FP64 launches with enough waves to keep the
// SIMDs busy); kIeee -- the IEEE code alone; kFast -- the FP64 fast form.
template <typename S, bool ST1, bool COUR, int FD>
__device__ __forceinline__ void x_segment(const GridGeom& g, const Phys<S>& ph,
                                          const TraceK<S>& tk, double dtdx,
                                          const GridView<S>& in, const GridView<S>& out,
                                          long jrow, int seg, double& cx, double& cy) {
  const int lane = threadIdx.x & 63;
  const long j = jrow + 2;
  const int i = kXSeg * seg + lane;  // padded column of this lane
  const int ic = i < g.nx + 4 ? i : g.nx + 3;
  const long o = j * g.pitch + ic;

  const S U0 = in.p[0][o], U1 = in.p[1][o], U2 = in.p[2][o], U3 = in.p[3][o];
  Prim<S> pc;
  bool fc = false;
  if constexpr (FD == kFast) {
    pc = prim_fast(U0, U1, U2, U3, ph);
  } else {
    // every cell of the wave (the slope neighbours included) in range: FastDiv
    fc = FD == kGated && wave_cell_in<S>(U0, U1, U2, U3, ph);
    pc = fc ? prim_cell<S, true>(U0, U1, U2, U3, ph) : prim_cell<S, false>(U0, U1, U2, U3, ph);
  }
  Q4<S> dq = {(S) 0, (S) 0, (S) 0, (S) 0};
  if (ph.iorder != 1) {
    const Q4<S> qm = from_left4(pc.q), qp = from_right4(pc.q);
    dq = slope_form<S, ST1, FD>(qm, pc.q, qp, ph);
  }
  Q4<S> qxm, qxp;
  if constexpr (FD == kFast || sizeof(S) == 4)
    trace_fast(pc.q, dq, pc.c, pc.cinv, pc.rinv, tk, qxm, qxp);
  else if (fc) trace_cell<S, true, true>(pc.q, dq, pc.c, tk, qxm, qxp, pc.rinv);
  else trace_cell<S, false>(pc.q, dq, pc.c, tk, qxm, qxp);
  const Q4<S> qr = from_right4(qxp);
  // flux through the face between this lane's cell and the next one
  const Q4<double> fr = to_d4(riemann_flux<S, FD>(qxm, qr, ph));
  const Q4<double> fl = from_left4(fr);

  const bool wr = lane >= 2 && lane <= kXSeg + 1 && i <= g.nx + 1;
  const double n0 = update1<S>(U0, fl.r, fr.r, dtdx);
  const double n1 = update1<S>(U1, fl.u, fr.u, dtdx);
  const double n2 = update1<S>(U2, fl.v, fr.v, dtdx);
  const double n3 = update1<S>(U3, fl.p, fr.p, dtdx);
  if (wr) {
    out.p[0][o] = (S) n0;
    out.p[1][o] = (S) n1;
    out.p[2][o] = (S) n2;
    out.p[3][o] = (S) n3;
    if (COUR) {
      S ax, ay;
      courant_cell<S, FD>((S) n0, (S) n1, (S) n2, (S) n3, ph, ax, ay);
      cx = maxm(cx, (double) ax);
      cy = maxm(cy, (double) ay);
    }
  }
}

// ---- x sweep, workgroup form: W waves share one row piece through LDS ------

// LDS of one workgroup: three exchange buffers (primitives, right-face states,
// fluxes), each four planes of 64*W + 2 values.  Slot t + 1 holds thread t's
// value; slots 0 and 64*W + 1 repeat the end threads' values, so the halo
// threads read plausible states (their results are never stored).
template <typename S, int W>
struct XLds {
  S q[4][64 * W + 2], x[4][64 * W + 2];
  double f[4][64 * W + 2];  // fluxes, converted once for the update
  int in[W];  // per wave: all its cells passed cell_in
};

template <int W, typename S>
__device__ __forceinline__ void lds_put(S (&b)[4][64 * W + 2], int t, const Q4<S>& v) {
  b[0][t + 1] = v.r;
  b[1][t + 1] = v.u;
  b[2][t + 1] = v.v;
  b[3][t + 1] = v.p;
  if (t == 0 || t == 64 * W - 1) {
    const int e = t == 0 ? 0 : 64 * W + 1;
    b[0][e] = v.r;
    b[1][e] = v.u;
    b[2][e] = v.v;
    b[3][e] = v.p;
  }
}
template <int W, typename S>
__device__ __forceinline__ Q4<S> lds_get(const S (&b)[4][64 * W + 2], int slot) {
  return Q4<S>{b[0][slot], b[1][slot], b[2][slot], b[3][slot]};
}

// One thread's cell of a row piece once loaded: its place, conserved values
// and primitives (`fc`: kGated, the wave's cells passed cell_in).
template <typename S>
struct XCell {
  long o;
  int i;
  S U0, U1, U2, U3;
  Prim<S> pc;
  bool fc;
};

// The workgroup's 64*W threads own 64*W consecutive cells of a row (padded
// columns (64*W - 4)*seg + t) and update the middle 64*W - 4: the same
// arithmetic as x_segment, with the neighbour values passed through LDS (no
// VALU lane moves) and the 4-cell halo paid once per workgroup instead of once
// per wave.  The workgroup walks `npieces` consecutive row pieces from piece
// `first`, software-pipelined: a piece's fluxes and the next piece's
// primitives cross at one barrier (A), where the piece's update then runs
// beside the next piece's slopes and traces, and the right-face states at
// another (B): two barriers per piece.  Every LDS buffer is rewritten only
// after a barrier that follows its last read: q (and the gated form's
// per-wave flags `in`) for piece k+1 after B of piece k, read between A and B
// of piece k+1; x before B, read between B and the next A; f between B and
// the next A, read between that A and the next B.  Every thread of the
// workgroup calls it with the same arguments: it has barriers.
template <typename S, bool ST1, bool COUR, int W, int FD>
__device__ __forceinline__ void x_blocks(const GridGeom& g, const Phys<S>& ph,
                                         const TraceK<S>& tk, double dtdx,
                                         const GridView<S>& in, const GridView<S>& out,
                                         int first, int npieces, int nseg, XLds<S, W>& sh,
                                         double& cx, double& cy) {
  constexpr int N = 64 * W;
  // the FP64 form's range-gated code has no registers to spare for it
  constexpr bool kPrefetch = FD == kFast || sizeof(S) == 4;
  const int t = threadIdx.x;
  SegWalk w(first, nseg);
  if (npieces <= 0 || w.row >= g.ny) return;  // workgroup-uniform
  // a piece's cell: its loads, then its primitives (published)
  auto fetch = [&](const SegWalk& ww) {
    XCell<S> c;
    c.i = (N - 4) * ww.seg + t;
    const int ic = c.i < g.nx + 4 ? c.i : g.nx + 3;
    c.o = ((long) ww.row + 2) * g.pitch + ic;
    c.U0 = in.p[0][c.o];
    c.U1 = in.p[1][c.o];
    c.U2 = in.p[2][c.o];
    c.U3 = in.p[3][c.o];
    return c;
  };
  auto prims = [&](XCell<S>& c) {
    c.fc = false;
    // kGated: FastDiv for the primitives when the wave's cells are in range,
    // for the traces when every wave's are (the slopes read neighbours of
    // other waves)
    if constexpr (FD == kFast) {
      c.pc = prim_fast(c.U0, c.U1, c.U2, c.U3, ph);
    } else {
      c.fc = FD == kGated && wave_cell_in<S>(c.U0, c.U1, c.U2, c.U3, ph);
      c.pc = c.fc ? prim_cell<S, true>(c.U0, c.U1, c.U2, c.U3, ph)
                  : prim_cell<S, false>(c.U0, c.U1, c.U2, c.U3, ph);
    }
    if (ph.iorder != 1) {  // launch-uniform
      lds_put<W>(sh.q, t, c.pc.q);
      if constexpr (FD == kGated) {
        if ((t & 63) == 0) sh.in[t >> 6] = c.fc;
      }
    }
  };
  // update a piece from its own right-face flux and the left neighbour's
  auto update = [&](const XCell<S>& c, const Q4<double>& fr) {
    const Q4<double> fl = lds_get<W>(sh.f, t);
    const bool wr = t >= 2 && t <= N - 3 && c.i <= g.nx + 1;
    const double n0 = update1<S>(c.U0, fl.r, fr.r, dtdx);
    const double n1 = update1<S>(c.U1, fl.u, fr.u, dtdx);
    const double n2 = update1<S>(c.U2, fl.v, fr.v, dtdx);
    const double n3 = update1<S>(c.U3, fl.p, fr.p, dtdx);
    if (wr) {
      out.p[0][c.o] = (S) n0;
      out.p[1][c.o] = (S) n1;
      out.p[2][c.o] = (S) n2;
      out.p[3][c.o] = (S) n3;
      if (COUR) {
        S ax, ay;
        courant_cell<S, FD>((S) n0, (S) n1, (S) n2, (S) n3, ph, ax, ay);
        cx = maxm(cx, (double) ax);
        cy = maxm(cy, (double) ay);
      }
    }
  };
  XCell<S> cur = fetch(w), prev, next;
  prims(cur);
  Q4<double> fprev;
  bool have_prev = false;
  for (int k = 0;;) {
    __syncthreads();  // A: cur's primitives, prev's fluxes
    if (have_prev) update(prev, fprev);
    Q4<S> dq = {(S) 0, (S) 0, (S) 0, (S) 0};
    bool ft = cur.fc;
    if (ph.iorder != 1) {
      if constexpr (FD == kGated) {
        int all = 1;
        for (int v = 0; v < W; v++) all &= sh.in[v];
        ft = __builtin_amdgcn_readfirstlane(all) != 0;  // the same in every lane
      }
      dq = slope_form<S, ST1, FD>(lds_get<W>(sh.q, t), cur.pc.q, lds_get<W>(sh.q, t + 2), ph);
    }
    Q4<S> qxm, qxp;
    if constexpr (FD == kFast || sizeof(S) == 4) {
      trace_fast(cur.pc.q, dq, cur.pc.c, cur.pc.cinv, cur.pc.rinv, tk, qxm, qxp);
    } else if constexpr (FD == kGated) {
      if (ft) trace_cell<S, true, true>(cur.pc.q, dq, cur.pc.c, tk, qxm, qxp, cur.pc.rinv);
      else trace_cell<S, false>(cur.pc.q, dq, cur.pc.c, tk, qxm, qxp);  // (ft implies fc)
    } else {
      trace_cell<S>(cur.pc.q, dq, cur.pc.c, tk, qxm, qxp);
    }
    lds_put<W>(sh.x, t, qxp);
    __syncthreads();  // B: the right-face states
    // (fast and mixed forms) the next piece's loads go out before this
    // piece's Riemann solves
    SegWalk wn = w;
    wn.advance(1, nseg);
    const bool more = k + 1 < npieces && wn.row < g.ny;  // workgroup-uniform
    if constexpr (kPrefetch) {
      if (more) next = fetch(wn);
    }
    const Q4<S> qr = lds_get<W>(sh.x, t + 2);
    // flux through the face between this thread's cell and the next one
    fprev = to_d4(riemann_flux<S, FD>(qxm, qr, ph));
    lds_put<W>(sh.f, t, fprev);
    prev = cur;
    have_prev = true;
    if (!more) break;
    w = wn;
    k++;
    if constexpr (kPrefetch) cur = next;
    else cur = fetch(w);
    prims(cur);
  }
  __syncthreads();  // A: the last piece's fluxes
  update(prev, fprev);
}

// ---- y sweep: one column, rows [j0, j1) ----------------------------------------

template <typename S>
struct Cell {
  S u0, u1, u2, u3;  // conserved, sweep slot order
};

// Lane-private LDS slots of one wave of the y march: the flux through the
// face above the row being updated, that row's lower-face trace state, and
// the conserved values of the rows awaiting their update (slot j % 3 for
// row j), carried from one row to the next through LDS rather than
// registers (the march's registers then fit four waves per SIMD instead of
// three; the conserved values are not read from memory twice).  Only the
// owning lane reads or writes its slots, so no barrier is needed.
template <typename S>
struct YStash {
  double f[4][64];  // converted once for the update
  S xm[4][64];
  S u[3][4][64];
};

template <typename S>
__device__ __forceinline__ void stash_put(S (&b)[4][64], int lane, const Q4<S>& v) {
  b[0][lane] = v.r;
  b[1][lane] = v.u;
  b[2][lane] = v.v;
  b[3][lane] = v.p;
}
template <typename S>
__device__ __forceinline__ Q4<S> stash_get(const S (&b)[4][64], int lane) {
  return Q4<S>{b[0][lane], b[1][lane], b[2][lane], b[3][lane]};
}

template <typename S>
__device__ __forceinline__ void stash_put_cell(S (&b)[4][64], int lane, const Cell<S>& c) {
  b[0][lane] = c.u0;
  b[1][lane] = c.u1;
  b[2][lane] = c.u2;
  b[3][lane] = c.u3;
}
template <typename S>
__device__ __forceinline__ Cell<S> stash_get_cell(const S (&b)[4][64], int lane) {
  return Cell<S>{b[0][lane], b[1][lane], b[2][lane], b[3][lane]};
}

template <typename S>
__device__ __forceinline__ Cell<S> load_y(const GridView<S>& in, long o) {
  // gather with the IU/IV swap of conservar.c:106-109
  return Cell<S>{in.p[0][o], in.p[2][o], in.p[1][o], in.p[3][o]};
}

template <typename S, bool ST1, bool COUR, int FD>
__device__ __forceinline__ void y_march(const GridGeom& g, const Phys<S>& ph,
                                        const TraceK<S>& tk, double dtdx,
                                        const GridView<S>& in, const GridView<S>& out,
                                        bool colok, long i, int j0, int j1, YStash<S>* ys,
                                        double& cx, double& cy) {
  const long P = g.pitch;
  const int lane = threadIdx.x & 63;
  const bool o1 = ph.iorder == 1;

  // prologue: cells j0-2 .. j0+1
  Cell<S> ua = load_y(in, (long) (j0 - 2) * P + i);
  Cell<S> ub = load_y(in, (long) (j0 - 1) * P + i);
  Cell<S> uc = load_y(in, (long) j0 * P + i);
  Cell<S> ud = load_y(in, (long) (j0 + 1) * P + i);
  // rows j0 and j0+1 await their updates
  stash_put_cell(ys->u[j0 % 3], lane, uc);
  stash_put_cell(ys->u[(j0 + 1) % 3], lane, ud);
  // kGated: FastDiv per row when the wave's cells of the row are in range (a
  // trace also needs the rows above and below it in range)
  auto prim = [&](const Cell<S>& u, bool f) {
    if constexpr (FD == kFast) return prim_fast(u.u0, u.u1, u.u2, u.u3, ph);
    else
      return f ? prim_cell<S, true>(u.u0, u.u1, u.u2, u.u3, ph)
               : prim_cell<S, false>(u.u0, u.u1, u.u2, u.u3, ph);
  };
  auto trace = [&](const Prim<S>& pq, const Q4<S>& dq, bool f, Q4<S>& qxm, Q4<S>& qxp) {
    if constexpr (FD == kFast || sizeof(S) == 4)
      trace_fast(pq.q, dq, pq.c, pq.cinv, pq.rinv, tk, qxm, qxp);
    else if (f) trace_cell<S, true>(pq.q, dq, pq.c, tk, qxm, qxp);
    else trace_cell<S, false>(pq.q, dq, pq.c, tk, qxm, qxp);
  };
  auto cell_ok = [&](const Cell<S>& u) {
    return FD == kGated && wave_cell_in<S>(u.u0, u.u1, u.u2, u.u3, ph);
  };
  const bool fa = cell_ok(ua);
  const bool fb = cell_ok(ub);
  bool fC = cell_ok(uc);
  bool fD = cell_ok(ud);
  Prim<S> pa = prim(ua, fa);
  Prim<S> pb = prim(ub, fb);
  Prim<S> pcc = prim(uc, fC);
  Prim<S> pd = prim(ud, fD);
  const Q4<S> z = {(S) 0, (S) 0, (S) 0, (S) 0};
  Q4<S> qxm_b, qxp_b, qxm_c, qxp_c;
  trace(pb, o1 ? z : slope_form<S, ST1, FD>(pa.q, pb.q, pcc.q, ph), fa && fb && fC, qxm_b,
        qxp_b);
  trace(pcc, o1 ? z : slope_form<S, ST1, FD>(pb.q, pcc.q, pd.q, ph), fb && fC && fD, qxm_c,
        qxp_c);
  stash_put(ys->f, lane, to_d4(riemann_flux<S, FD>(qxm_b, qxp_c, ph)));  // face (j0-1 | j0)
  stash_put(ys->xm, lane, qxm_c);

  // window: cell C = row j-2 (being updated), D = row j-1; carried in LDS:
  // qxm of C, the flux through the face below C and the conserved values of
  // C and D
  Q4<S> qC = pcc.q, qD = pd.q;
  S cD = pd.c, rD = pd.rinv, iD = pd.cinv;
  // the next row's loads are issued one row ahead of their use
  Cell<S> uN = load_y(in, (long) (j0 + 2) * P + i);
  // One row of the march: take row j, prefetch row j+1, update row j-2,
  // shift the window.
  auto row = [&](int j) {
    // (the compiler must read the slots the previous row wrote, not forward
    // the values in registers)
    asm volatile("" ::: "memory");
    const long oc = (long) (j - 2) * P + i;
    const Cell<S> uE = uN;
    const int sj = j % 3;  // wave-uniform
    stash_put_cell(ys->u[sj], lane, uE);
    const int jn = j + 1 <= j1 + 1 ? j + 1 : j1 + 1;  // last row: a harmless reload
    uN = load_y(in, (long) jn * P + i);
    const bool fE = cell_ok(uE);
    const Prim<S> pE = prim(uE, fE);
    Q4<S> qxmD, qxpD;
    {
      const Q4<S> dqD = o1 ? z : slope_form<S, ST1, FD>(qC, qD, pE.q, ph);
      if constexpr (FD == kFast || sizeof(S) == 4) trace_fast(qD, dqD, cD, iD, rD, tk, qxmD, qxpD);
      else if (fC && fD && fE) trace_cell<S, true, true>(qD, dqD, cD, tk, qxmD, qxpD, rD);
      else trace_cell<S, false>(qD, dqD, cD, tk, qxmD, qxpD);
    }
    const Q4<S> qxmC = stash_get(ys->xm, lane);
    stash_put(ys->xm, lane, qxmD);
    const Q4<double> fcur = to_d4(riemann_flux<S, FD>(qxmC, qxpD, ph));  // face (j-2 | j-1)
    const Q4<double> fprev = stash_get(ys->f, lane);
    stash_put(ys->f, lane, fcur);
    // (read C's conserved values only now: loaded early they would hold
    // registers through the Riemann solve)
    asm volatile("" ::: "memory");
    const Cell<S> uC = stash_get_cell(ys->u[sj == 2 ? 0 : sj + 1], lane);  // row j-2
    const double n0 = update1<S>(uC.u0, fprev.r, fcur.r, dtdx);
    const double n1 = update1<S>(uC.u1, fprev.u, fcur.u, dtdx);
    const double n2 = update1<S>(uC.u2, fprev.v, fcur.v, dtdx);
    const double n3 = update1<S>(uC.u3, fprev.p, fcur.p, dtdx);
    if (colok) {
      // scatter back with the swap undone (conservar.c:177-180)
      st_stream(&out.p[0][oc], (S) n0);
      st_stream(&out.p[2][oc], (S) n1);
      st_stream(&out.p[1][oc], (S) n2);
      st_stream(&out.p[3][oc], (S) n3);
      if (COUR) {
        S ax, ay;
        courant_cell<S, FD>((S) n0, (S) n2, (S) n1, (S) n3, ph, ax, ay);
        cx = maxm(cx, (double) ax);
        cy = maxm(cy, (double) ay);
      }
    }
    qC = qD;
    qD = pE.q;
    cD = pE.c;
    rD = pE.rinv;
    iD = pE.cinv;
    fC = fD;
    fD = fE;
  };
  // two rows per trip: the window shifts by two, so its registers map back
  // onto themselves without copies
  int j = j0 + 2;
  for (; j + 1 <= j1 + 1; j += 2) {
    row(j);
    row(j + 1);
  }
  if (j <= j1 + 1) row(j);
}

// ---- ghost cells -------------------------------------------------------------

// make_boundary.c:237-289 (x) and 381-433 (y) for one interior line k and one
// variable v: the four ghost cells in the reference's order (low side 0, 1,
// then high side n+2, n+3), so sources that are themselves ghosts (grids
// narrower than 3 cells, the absorbing high-x source n+2) read exactly what
// the reference reads.  Types: 1 reflect (normal momentum negated), 2 absorb,
// any other positive value periodic, <= 0 untouched.  Values go to src and,
// when dst is non-null, to dst (the other ping-pong buffer).
template <typename S>
__device__ __forceinline__ void boundary_line(int idim, const GridGeom& g, int bl, int br,
                                              int bd, int bu, S* sp, S* dp, int v, int k) {
  const int nn = (idim == 1) ? g.nx : g.ny;
  const int normal = (idim == 1) ? 1 : 2;
  const long line = k + 2;
  for (int gp = 0; gp < 4; gp++) {
    const int lo = gp < 2;
    const int typ = (idim == 1) ? (lo ? bl : br) : (lo ? bd : bu);
    if (typ <= 0) continue;
    const int gidx = lo ? gp : nn + gp;  // 0, 1, nn+2, nn+3
    int sidx;
    S sign = (S) 1;
    if (typ == 1) {
      sidx = lo ? 3 - gidx : 2 * nn + 3 - gidx;
      if (v == normal) sign = (S) -1;
    } else if (typ == 2) {
      // reference source cells: left 2, right nn+2, down 4, up ny+1
      if (idim == 1) sidx = lo ? 2 : nn + 2;
      else sidx = lo ? 4 : nn + 1;
    } else {
      sidx = lo ? nn + gidx : gidx - nn;
    }
    long od, os;
    if (idim == 1) {
      od = line * g.pitch + gidx;
      os = line * g.pitch + sidx;
    } else {
      od = (long) gidx * g.pitch + line;
      os = (long) sidx * g.pitch + line;
    }
    const S val = sp[os] * sign;
    sp[od] = val;
    if (dp) dp[od] = val;
  }
}

// ---- initial state ---------------------------------------------------------------

// hydro_funcs.c:79-128: interior rho=1, u=v=0, E=1e-5, ghosts 0, E=eb at padded
// column bx for rows [by0, by1).  Element t of (nx+4)*(ny+4).
template <typename S>
__device__ __forceinline__ void init_cell(const GridGeom& g, GridView<S> a, GridView<S> b,
                                          long t, int bx, int by0, int by1, S eb) {
  const long nxt = g.nx + 4, nyt = g.ny + 4;
  const long j = t / nxt, i = t - j * nxt;
  const bool in = i >= 2 && i < nxt - 2 && j >= 2 && j < nyt - 2;
  S v0 = in ? (S) 1 : (S) 0, v3 = in ? (S) 1e-5 : (S) 0;
  if (i == bx && j >= by0 && j < by1) v3 = eb;
  const long o = j * g.pitch + i;
  a.p[0][o] = v0;
  a.p[1][o] = (S) 0;
  a.p[2][o] = (S) 0;
  a.p[3][o] = v3;
  if (b.p[0]) {
    b.p[0][o] = v0;
    b.p[1][o] = (S) 0;
    b.p[2][o] = (S) 0;
    b.p[3][o] = v3;
  }
}

}  // namespace hydro
