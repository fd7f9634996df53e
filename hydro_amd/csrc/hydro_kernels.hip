// HIP kernels for gfx950 (MI355X): fused x / y sweeps of HydroC99's split
// Godunov/MUSCL step, the ghost-cell fill, the CFL reduction and the
// device-side time-loop control.
//
// Design (one kernel per sweep direction, no intermediate arrays in HBM):
//
// * x sweep: one 64-lane wave owns 64 consecutive cells of a row and updates
//   the middle 60.  Each lane turns its cell into primitives; neighbours come
//   from DPP wave shifts (v_mov_b32_dpp wave_shl/wave_shr), so slope, trace,
//   Riemann solve, flux and update all happen in registers.  Row segments are
//   read and written with unit stride.
// * y sweep: one lane owns one column and marches down a segment of rows with
//   a three-cell register window (primitives, one carried trace state, one
//   carried flux).  A wave covers 64 adjacent columns, so every load/store is
//   a contiguous 512 B (FP64) row piece.
// * Both sweeps read one state buffer and write the other (ping-pong), so
//   segments never race on shared halo cells.
// * The second sweep of a step can fold the CFL maximum of the state it has
//   just written into its epilogue (the next step's compute_deltat), reduced
//   wave -> workgroup -> one 64-bit atomic max per workgroup into one of
//   kCourSlots slots; the bit pattern of a non-negative double orders like its
//   value, so the integer max is the exact floating max.
#include "hydro_kernels.hpp"
#include "hydro_numerics.hpp"

namespace hydro {

// ---- lane exchange -----------------------------------------------------------

// lane l receives the value of lane l-1 (DPP wave_shr:1); lane 0 keeps its own.
__device__ __forceinline__ double from_left(double x) {
  int lo = __double2loint(x), hi = __double2hiint(x);
  lo = __builtin_amdgcn_update_dpp(lo, lo, 0x138, 0xf, 0xf, false);
  hi = __builtin_amdgcn_update_dpp(hi, hi, 0x138, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}
// lane l receives the value of lane l+1 (DPP wave_shl:1); lane 63 keeps its own.
__device__ __forceinline__ double from_right(double x) {
  int lo = __double2loint(x), hi = __double2hiint(x);
  lo = __builtin_amdgcn_update_dpp(lo, lo, 0x130, 0xf, 0xf, false);
  hi = __builtin_amdgcn_update_dpp(hi, hi, 0x130, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float from_left(float x) {
  int b = __float_as_int(x);
  b = __builtin_amdgcn_update_dpp(b, b, 0x138, 0xf, 0xf, false);
  return __int_as_float(b);
}
__device__ __forceinline__ float from_right(float x) {
  int b = __float_as_int(x);
  b = __builtin_amdgcn_update_dpp(b, b, 0x130, 0xf, 0xf, false);
  return __int_as_float(b);
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

// Every thread of the workgroup must call this (it has a barrier).
__device__ __forceinline__ void commit_courant(double cx, double cy, uint64_t* cour) {
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
    const int slot = blockIdx.x % kCourSlots;
    atomicMax((unsigned long long*) &cour[slot], (unsigned long long) __double_as_longlong(mx));
    atomicMax((unsigned long long*) &cour[kCourSlots + slot],
              (unsigned long long) __double_as_longlong(my));
  }
}

// ---- x sweep ---------------------------------------------------------------

template <typename S, bool ST1, bool COUR>
__global__ __launch_bounds__(256) void k_sweep_x(GridGeom g, Phys<S> ph, double dx,
                                                 double dt_val, const RunCtl* ctl,
                                                 GridView<S> in, GridView<S> out,
                                                 uint64_t* cour, int nseg, int xiter) {
  double dt = dt_val;
  if (ctl) {
    if (!ctl->active) return;  // grid-uniform
    dt = ctl->dt;
  }
  const double dtdx = dt / dx;  // hydro_godunov.c:101
  const TraceK<S> tk = make_tracek<S>((S) dtdx, ph.scheme);
  const int lane = threadIdx.x & 63;
  const int wpb = blockDim.x >> 6;
  const long ntask = (long) nseg * g.ny;
  double cx = 0.0, cy = 0.0;
  // A workgroup walks wpb * xiter consecutive row segments, wpb at a time.
  for (int it = 0; it < xiter; it++) {
    const long task = ((long) blockIdx.x * xiter + it) * wpb + (threadIdx.x >> 6);
    const bool ok = task < ntask;
    const long jrow = ok ? task / nseg : 0;
    const int seg = ok ? (int) (task - jrow * nseg) : 0;
    const long j = jrow + 2;
    const int i = kXSeg * seg + lane;  // padded column of this lane
    const int ic = i < g.nx + 4 ? i : g.nx + 3;
    const long o = j * g.pitch + ic;

    const S U0 = in.p[0][o], U1 = in.p[1][o], U2 = in.p[2][o], U3 = in.p[3][o];
    const Prim<S> pc = prim_cell<S>(U0, U1, U2, U3, ph);
    Q4<S> dq = {(S) 0, (S) 0, (S) 0, (S) 0};
    if (ph.iorder != 1) {
      const Q4<S> qm = from_left4(pc.q), qp = from_right4(pc.q);
      dq = slope4<S, ST1>(qm, pc.q, qp, ph.slope_type);
    }
    Q4<S> qxm, qxp;
    trace_cell<S>(pc.q, dq, pc.c, tk, qxm, qxp);
    const Q4<S> qr = from_right4(qxp);
    // flux through the face between this lane's cell and the next one
    const Q4<S> fr = riemann_flux<S>(qxm, qr, ph);
    const Q4<S> fl = from_left4(fr);

    const bool wr = ok && lane >= 2 && lane <= kXSeg + 1 && i <= g.nx + 1;
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
        courant_cell<S>((S) n0, (S) n1, (S) n2, (S) n3, ph, ax, ay);
        cx = maxm(cx, (double) ax);
        cy = maxm(cy, (double) ay);
      }
    }
  }
  if (COUR) commit_courant(cx, cy, cour);
}

// ---- y sweep ---------------------------------------------------------------

template <typename S>
struct Cell {
  S u0, u1, u2, u3;  // conserved, sweep slot order
};

template <typename S>
__device__ __forceinline__ Cell<S> load_y(const GridView<S>& in, long o) {
  // gather with the IU/IV swap of conservar.c:106-109
  return Cell<S>{in.p[0][o], in.p[2][o], in.p[1][o], in.p[3][o]};
}

template <typename S, bool ST1, bool COUR>
__global__ __launch_bounds__(256) void k_sweep_y(GridGeom g, Phys<S> ph, double dx,
                                                 double dt_val, const RunCtl* ctl,
                                                 GridView<S> in, GridView<S> out,
                                                 uint64_t* cour, int yseg, int ncolblk) {
  double dt = dt_val;
  if (ctl) {
    if (!ctl->active) return;
    dt = ctl->dt;
  }
  const double dtdx = dt / dx;
  const TraceK<S> tk = make_tracek<S>((S) dtdx, ph.scheme);
  const int cb = blockIdx.x % ncolblk;
  const int sgi = blockIdx.x / ncolblk;
  const int col = cb * blockDim.x + threadIdx.x;
  const bool colok = col < g.nx;
  const long i = 2 + (colok ? col : g.nx - 1);
  const int j0 = 2 + sgi * yseg;
  const int j1 = min(j0 + yseg, g.ny + 2);  // rows [j0, j1) are updated
  const long P = g.pitch;
  const S st = ph.slope_type;
  const bool o1 = ph.iorder == 1;

  // prologue: cells j0-2 .. j0+1
  Cell<S> ua = load_y(in, (long) (j0 - 2) * P + i);
  Cell<S> ub = load_y(in, (long) (j0 - 1) * P + i);
  Cell<S> uc = load_y(in, (long) j0 * P + i);
  Cell<S> ud = load_y(in, (long) (j0 + 1) * P + i);
  Prim<S> pa = prim_cell<S>(ua.u0, ua.u1, ua.u2, ua.u3, ph);
  Prim<S> pb = prim_cell<S>(ub.u0, ub.u1, ub.u2, ub.u3, ph);
  Prim<S> pcc = prim_cell<S>(uc.u0, uc.u1, uc.u2, uc.u3, ph);
  Prim<S> pd = prim_cell<S>(ud.u0, ud.u1, ud.u2, ud.u3, ph);
  const Q4<S> z = {(S) 0, (S) 0, (S) 0, (S) 0};
  Q4<S> qxm_b, qxp_b, qxm_c, qxp_c;
  trace_cell<S>(pb.q, o1 ? z : slope4<S, ST1>(pa.q, pb.q, pcc.q, st), pb.c, tk, qxm_b, qxp_b);
  trace_cell<S>(pcc.q, o1 ? z : slope4<S, ST1>(pb.q, pcc.q, pd.q, st), pcc.c, tk, qxm_c, qxp_c);
  Q4<S> fprev = riemann_flux<S>(qxm_b, qxp_c, ph);  // face (j0-1 | j0)

  // window: cell C = row j-2 (being updated), D = row j-1; carried: qxm of C,
  // flux through the face below C
  Q4<S> qC = pcc.q, qD = pd.q;
  S cD = pd.c;
  Q4<S> qxmC = qxm_c;
  Cell<S> uC = uc, uD = ud;
  double cx = 0.0, cy = 0.0;
  for (int j = j0 + 2; j <= j1 + 1; j++) {
    const long oj = (long) j * P + i;
    const Cell<S> uE = load_y(in, oj);
    const Prim<S> pE = prim_cell<S>(uE.u0, uE.u1, uE.u2, uE.u3, ph);
    Q4<S> qxmD, qxpD;
    trace_cell<S>(qD, o1 ? z : slope4<S, ST1>(qC, qD, pE.q, st), cD, tk, qxmD, qxpD);
    const Q4<S> fcur = riemann_flux<S>(qxmC, qxpD, ph);  // face (j-2 | j-1)
    const double n0 = update1<S>(uC.u0, fprev.r, fcur.r, dtdx);
    const double n1 = update1<S>(uC.u1, fprev.u, fcur.u, dtdx);
    const double n2 = update1<S>(uC.u2, fprev.v, fcur.v, dtdx);
    const double n3 = update1<S>(uC.u3, fprev.p, fcur.p, dtdx);
    const long oc = (long) (j - 2) * P + i;
    if (colok) {
      // scatter back with the swap undone (conservar.c:177-180)
      out.p[0][oc] = (S) n0;
      out.p[2][oc] = (S) n1;
      out.p[1][oc] = (S) n2;
      out.p[3][oc] = (S) n3;
      if (COUR) {
        S ax, ay;
        courant_cell<S>((S) n0, (S) n2, (S) n1, (S) n3, ph, ax, ay);
        cx = maxm(cx, (double) ax);
        cy = maxm(cy, (double) ay);
      }
    }
    fprev = fcur;
    qxmC = qxmD;
    qC = qD;
    qD = pE.q;
    cD = pE.c;
    uC = uD;
    uD = uE;
  }
  if (COUR) commit_courant(cx, cy, cour);
}

// ---- standalone CFL maximum -------------------------------------------------

template <typename S>
__global__ __launch_bounds__(256) void k_courant(GridGeom g, Phys<S> ph, GridView<S> in,
                                                 uint64_t* cour, const RunCtl* ctl) {
  if (ctl && !ctl->active) return;
  double cx = 0.0, cy = 0.0;
  const long n = (long) g.nx * g.ny;
  for (long t = (long) blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (long) gridDim.x * blockDim.x) {
    const long jj = t / g.nx;
    const long ii = t - jj * g.nx;
    const long o = (jj + 2) * g.pitch + ii + 2;
    S ax, ay;
    courant_cell<S>(in.p[0][o], in.p[1][o], in.p[2][o], in.p[3][o], ph, ax, ay);
    cx = maxm(cx, (double) ax);
    cy = maxm(cy, (double) ay);
  }
  commit_courant(cx, cy, cour);
}

// ---- ghost cells -------------------------------------------------------------

// make_boundary.c:237-289 (x) and 381-433 (y): ghosts of direction idim along
// interior lines, sources chosen as the reference does for types 1 (reflect,
// normal momentum negated), 2 (absorb) and any other positive value
// (periodic).  One thread per (line, variable) fills the four ghost cells in
// the reference's order (low side 0, 1 then high side n+2, n+3), so sources
// that are themselves ghosts (grids narrower than 3 cells, the absorbing
// high-x source n+2) read exactly what the reference reads.  Values go to src
// and, when set, to dst2 (the other ping-pong buffer).
template <typename S>
__global__ __launch_bounds__(256) void k_boundary(int idim, GridGeom g, int bl, int br, int bd,
                                                  int bu, GridView<S> src, GridView<S> dst2,
                                                  int has_dst2, const RunCtl* ctl) {
  if (ctl && !ctl->active) return;
  const int nl = (idim == 1) ? g.ny : g.nx;
  const long t = (long) blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 4L * nl) return;
  const int k = (int) (t % nl);
  const int v = (int) (t / nl);
  const int nn = (idim == 1) ? g.nx : g.ny;
  const int normal = (idim == 1) ? 1 : 2;
  const long line = k + 2;
  S* const sp = src.p[v];
  S* const dp = dst2.p[v];
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
    if (has_dst2) dp[od] = val;
  }
}

// ---- dt and run control ------------------------------------------------------

// Reduce the slots (clearing them for the next use) into the dt formula of
// compute_deltat.c:197.  Must be called by all 256 threads of one workgroup.
__device__ double reduce_dt(uint64_t* cour, double cf, double dx, double smallc) {
  __shared__ double sx[4], sy[4];
  double cx = 0.0, cy = 0.0;
  for (int s = threadIdx.x; s < kCourSlots; s += blockDim.x) {
    cx = maxm(cx, __longlong_as_double((long long) cour[s]));
    cy = maxm(cy, __longlong_as_double((long long) cour[kCourSlots + s]));
    cour[s] = 0;
    cour[kCourSlots + s] = 0;
  }
  cx = wave_max(cx);
  cy = wave_max(cy);
  if ((threadIdx.x & 63) == 0) {
    sx[threadIdx.x >> 6] = cx;
    sy[threadIdx.x >> 6] = cy;
  }
  __syncthreads();
  double mx = sx[0], my = sy[0];
  for (int k = 1; k < (int) (blockDim.x >> 6); k++) {
    mx = maxm(mx, sx[k]);
    my = maxm(my, sy[k]);
  }
  return cf * dx / maxm(mx, maxm(my, smallc));
}

__global__ __launch_bounds__(256) void k_dt_finalize(uint64_t* cour, double cf, double dx,
                                                     double smallc, double* dt_out) {
  double dt = reduce_dt(cour, cf, dx, smallc);
  if (threadIdx.x == 0) *dt_out = dt;
}

// main.c:244-260: open a step.
__global__ __launch_bounds__(256) void k_ctl_begin(RunCtl* ctl, uint64_t* cour, double cf,
                                                   double dx, double smallc) {
  const double t = ctl->t, tend = ctl->tend;
  const int nstep = ctl->nstep, nstepmax = ctl->nstepmax;
  const bool act = (t < tend) && (nstep < nstepmax);
  __syncthreads();
  if (!act) {
    if (threadIdx.x == 0) ctl->active = 0;
    return;  // block-uniform
  }
  if ((nstep % 2) == 0) {
    double dt = reduce_dt(cour, cf, dx, smallc);
    if (nstep == 0) dt = dt / 2.0;
    if (threadIdx.x == 0) ctl->dt = dt;
  }
  if (threadIdx.x == 0) ctl->active = 1;
}

// main.c:297-298: close a step.
__global__ void k_ctl_end(RunCtl* ctl) {
  if (threadIdx.x != 0 || !ctl->active) return;
  const int n = ctl->nstep;
  const int k = n - ctl->log_base;
  if (ctl->log && k >= 0 && k < ctl->log_cap) ctl->log[k] = ctl->dt;
  ctl->nstep = n + 1;
  ctl->t += ctl->dt;
}

__global__ void k_clear_slots(uint64_t* cour) {
  for (int s = threadIdx.x; s < 2 * kCourSlots; s += blockDim.x) cour[s] = 0;
}

template <typename S>
__global__ __launch_bounds__(256) void k_init(GridGeom g, GridView<S> a, GridView<S> b, int has_b,
                                              int bx, int by0, int by1, S eb) {
  const long nxt = g.nx + 4, nyt = g.ny + 4;
  const long t = (long) blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nxt * nyt) return;
  const long j = t / nxt, i = t - j * nxt;
  const bool in = i >= 2 && i < nxt - 2 && j >= 2 && j < nyt - 2;
  S v0 = in ? (S) 1 : (S) 0, v3 = in ? (S) 1e-5 : (S) 0;
  if (i == bx && j >= by0 && j < by1) v3 = eb;
  const long o = j * g.pitch + i;
  a.p[0][o] = v0; a.p[1][o] = (S) 0; a.p[2][o] = (S) 0; a.p[3][o] = v3;
  if (has_b) { b.p[0][o] = v0; b.p[1][o] = (S) 0; b.p[2][o] = (S) 0; b.p[3][o] = v3; }
}

// ---- launchers -----------------------------------------------------------------

static inline unsigned cdiv(long a, long b) { return (unsigned) ((a + b - 1) / b); }

template <typename S>
void launch_boundary(int idim, const GridGeom& g, const PhysArgs& ph, GridView<S> src,
                     GridView<S> dst2, const RunCtl* ctl, hipStream_t st) {
  const int nl = (idim == 1) ? g.ny : g.nx;
  const long n = 4L * nl;
  const int has2 = dst2.p[0] != nullptr;
  hipLaunchKernelGGL(k_boundary<S>, dim3(cdiv(n, 256)), dim3(256), 0, st, idim, g, ph.bl, ph.br,
                     ph.bd, ph.bu, src, dst2, has2, ctl);
}

template <typename S>
void launch_sweep(int idim, const GridGeom& g, const PhysArgs& pa, GridView<S> in,
                  GridView<S> out, double dt_val, const RunCtl* ctl, uint64_t* cour,
                  const SweepTune& tune, hipStream_t st) {
  const Phys<S> ph = make_phys<S>(pa.gamma, pa.smallr, pa.smallc, pa.slope_type, pa.niter,
                                  pa.iorder, pa.scheme);
  const bool st1 = pa.slope_type == 1.0;
  const bool c = cour != nullptr;
  if (idim == 1) {
    const int nseg = (int) cdiv(g.nx, kXSeg);
    const long waves = (long) nseg * g.ny;
    const int xw = tune.xwaves > 0 ? tune.xwaves : 4;
    int xiter = tune.xiter > 0 ? tune.xiter : 8;
    // keep at least ~4 workgroups per CU before walking several segments
    while (xiter > 1 && cdiv(waves, (long) xw * xiter) < 1024) xiter >>= 1;
    const dim3 grid(cdiv(waves, (long) xw * xiter)), blk(64 * xw);
#define HX(A, B) \
  hipLaunchKernelGGL((k_sweep_x<S, A, B>), grid, blk, 0, st, g, ph, pa.dx, dt_val, ctl, in, out, \
                     cour, nseg, xiter)
    if (st1) { if (c) HX(true, true); else HX(true, false); }
    else { if (c) HX(false, true); else HX(false, false); }
#undef HX
  } else {
    const int yseg = tune.yseg > 0 ? tune.yseg : 64;
    const int bw = 64 * (tune.ywaves > 0 ? tune.ywaves : 4);
    const int ncolblk = (int) cdiv(g.nx, bw);
    const int nsg = (int) cdiv(g.ny, yseg);
    const dim3 grid((unsigned) ncolblk * nsg), blk(bw);
#define HY(A, B) \
  hipLaunchKernelGGL((k_sweep_y<S, A, B>), grid, blk, 0, st, g, ph, pa.dx, dt_val, ctl, in, out, \
                     cour, yseg, ncolblk)
    if (st1) { if (c) HY(true, true); else HY(true, false); }
    else { if (c) HY(false, true); else HY(false, false); }
#undef HY
  }
}

template <typename S>
void launch_courant(const GridGeom& g, const PhysArgs& pa, GridView<S> in, uint64_t* cour,
                    const RunCtl* ctl, hipStream_t st) {
  const Phys<S> ph = make_phys<S>(pa.gamma, pa.smallr, pa.smallc, pa.slope_type, pa.niter,
                                  pa.iorder, pa.scheme);
  const long n = (long) g.nx * g.ny;
  long nb = cdiv(n, 256 * 8);
  if (nb > 4096) nb = 4096;
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(k_courant<S>, dim3((unsigned) nb), dim3(256), 0, st, g, ph, in, cour, ctl);
}

void launch_dt_finalize(uint64_t* cour, double cf, double dx, double smallc, double* dt_out,
                        hipStream_t st) {
  hipLaunchKernelGGL(k_dt_finalize, dim3(1), dim3(256), 0, st, cour, cf, dx, smallc, dt_out);
}

void launch_ctl_begin(RunCtl* ctl, uint64_t* cour, double cf, double dx, double smallc,
                      hipStream_t st) {
  hipLaunchKernelGGL(k_ctl_begin, dim3(1), dim3(256), 0, st, ctl, cour, cf, dx, smallc);
}

void launch_ctl_end(RunCtl* ctl, hipStream_t st) {
  hipLaunchKernelGGL(k_ctl_end, dim3(1), dim3(64), 0, st, ctl);
}

void launch_clear_slots(uint64_t* cour, hipStream_t st) {
  hipLaunchKernelGGL(k_clear_slots, dim3(1), dim3(256), 0, st, cour);
}

template <typename S>
void launch_init(const GridGeom& g, GridView<S> a, GridView<S> b, int bx, int by0, int by1,
                 double e_blast, hipStream_t st) {
  const long n = (long) (g.nx + 4) * (g.ny + 4);
  hipLaunchKernelGGL(k_init<S>, dim3(cdiv(n, 256)), dim3(256), 0, st, g, a, b,
                     b.p[0] != nullptr, bx, by0, by1, (S) e_blast);
}
template void launch_init<double>(const GridGeom&, GridView<double>, GridView<double>, int, int,
                                  int, double, hipStream_t);
template void launch_init<float>(const GridGeom&, GridView<float>, GridView<float>, int, int,
                                 int, double, hipStream_t);

template void launch_boundary<double>(int, const GridGeom&, const PhysArgs&, GridView<double>,
                                      GridView<double>, const RunCtl*, hipStream_t);
template void launch_boundary<float>(int, const GridGeom&, const PhysArgs&, GridView<float>,
                                     GridView<float>, const RunCtl*, hipStream_t);
template void launch_sweep<double>(int, const GridGeom&, const PhysArgs&, GridView<double>,
                                   GridView<double>, double, const RunCtl*, uint64_t*,
                                   const SweepTune&, hipStream_t);
template void launch_sweep<float>(int, const GridGeom&, const PhysArgs&, GridView<float>,
                                  GridView<float>, double, const RunCtl*, uint64_t*,
                                  const SweepTune&, hipStream_t);
template void launch_courant<double>(const GridGeom&, const PhysArgs&, GridView<double>,
                                     uint64_t*, const RunCtl*, hipStream_t);
template void launch_courant<float>(const GridGeom&, const PhysArgs&, GridView<float>,
                                    uint64_t*, const RunCtl*, hipStream_t);

}  // namespace hydro
