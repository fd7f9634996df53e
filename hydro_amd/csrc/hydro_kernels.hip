// HIP kernels for gfx950 (MI355X), one grid per launch: fused x / y sweeps of
// HydroC99's split Godunov/MUSCL step, ghost fill, CFL reduction, initial
// state and the device-side time-loop control.  The per-task bodies live in
// hydro_sweeps.hpp (shared with the batched kernels).
//
// The file is compiled three times.  The FP64 unit (the default,
// HYDRO_BASE_UNIT) holds everything except the FP32-storage and FP64 fast
// sweeps and is built with -ffp-contract=off, so the FP64 path is op-for-op
// the reference's.  The mixed-precision unit (HYDRO_MIXED_UNIT) holds only
// the FP32-storage sweeps and CFL kernels, and the fast unit
// (HYDRO_FAST_UNIT) only the FP64 fast-form sweeps; both are built with
// -ffp-contract=on: those forms carry no bitwise contract with the
// reference, and multiply-adds fused within each source expression cut their
// VALU instruction counts while keeping the single-grid and batched kernels
// (same source) bit-identical.
#include "hydro_kernels.hpp"
#include "hydro_sweeps.hpp"

namespace hydro {

// ---- x sweep ---------------------------------------------------------------

// A workgroup walks wpb * xiter consecutive row segments, wpb at a time.
template <typename S, bool ST1, bool COUR, int FD>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void k_sweep_x(GridGeom g, Phys<S> ph, double dx,
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
  const int wpb = blockDim.x >> 6;
  double cx = 0.0, cy = 0.0;
  // segments (blockIdx.x * xiter + it) * wpb + wave, it = 0 .. xiter-1
  SegWalk w((int) blockIdx.x * xiter * wpb + wave_id(), nseg);
  for (int it = 0; it < xiter && w.row < g.ny; it++) {
    x_segment<S, ST1, COUR, FD>(g, ph, tk, dtdx, in, out, w.row, w.seg, cx, cy);
    w.advance(wpb, nseg);
  }
  if (COUR) commit_courant(cx, cy, cour, kCourSlots, blockIdx.x % kCourSlots);
}

// Workgroup form (W = 2 or 4 waves): a workgroup walks xiter consecutive
// row pieces of 64*W - 4 updated cells, neighbours through LDS (x_blocks).
template <typename S, bool ST1, bool COUR, int W, int FD>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(5))) void k_sweep_xb(GridGeom g, Phys<S> ph, double dx,
                                                     double dt_val, const RunCtl* ctl,
                                                     GridView<S> in, GridView<S> out,
                                                     uint64_t* cour, int nseg, int xiter) {
  __shared__ XLds<S, W> sh;
  double dt = dt_val;
  if (ctl) {
    if (!ctl->active) return;  // grid-uniform
    dt = ctl->dt;
  }
  const double dtdx = dt / dx;  // hydro_godunov.c:101
  const TraceK<S> tk = make_tracek<S>((S) dtdx, ph.scheme);
  double cx = 0.0, cy = 0.0;
  x_blocks<S, ST1, COUR, W, FD>(g, ph, tk, dtdx, in, out, (int) blockIdx.x * xiter, xiter, nseg,
                                sh, cx, cy);
  if (COUR) commit_courant(cx, cy, cour, kCourSlots, blockIdx.x % kCourSlots);
}

// ---- y sweep ---------------------------------------------------------------

// A workgroup covers 64 columns; its waves take consecutive row segments of
// yseg rows each.
template <typename S, bool ST1, bool COUR, int FD>
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
  const int sgi = (blockIdx.x / ncolblk) * (blockDim.x >> 6) + wave_id();
  const int col = cb * 64 + (threadIdx.x & 63);
  const bool colok = col < g.nx;
  const long i = 2 + (colok ? col : g.nx - 1);
  const int j0 = 2 + sgi * yseg;
  double cx = 0.0, cy = 0.0;
  if (j0 < g.ny + 2) {  // wave-uniform
    const int j1 = min(j0 + yseg, g.ny + 2);  // rows [j0, j1) are updated
    extern __shared__ __attribute__((aligned(16))) unsigned char ylds[];
    YStash<S>* ys = reinterpret_cast<YStash<S>*>(ylds) + wave_id();
    y_march<S, ST1, COUR, FD>(g, ph, tk, dtdx, in, out, colok, i, j0, j1, ys, cx, cy);
  }
  if (COUR) commit_courant(cx, cy, cour, kCourSlots, blockIdx.x % kCourSlots);
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
  commit_courant(cx, cy, cour, kCourSlots, blockIdx.x % kCourSlots);
}

// ---- ghost cells -------------------------------------------------------------

// One thread per (line, variable); see boundary_line.
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
  boundary_line<S>(idim, g, bl, br, bd, bu, src.p[v], has_dst2 ? dst2.p[v] : nullptr, v, k);
}

#ifdef HYDRO_BASE_UNIT
// ---- dt and run control ------------------------------------------------------

__global__ __launch_bounds__(64) void k_dt_finalize(uint64_t* cour, double cf, double dx,
                                                    double smallc, double* dt_out) {
  const double dt = wave_reduce_dt(cour, kCourSlots, cf, dx, smallc);
  if (threadIdx.x == 0) *dt_out = dt;
}

// main.c:244-260: open a step.
__global__ __launch_bounds__(64) void k_ctl_begin(RunCtl* ctl, uint64_t* cour, double cf,
                                                  double dx, double smallc) {
  const double t = ctl->t, tend = ctl->tend;
  const int nstep = ctl->nstep, nstepmax = ctl->nstepmax;
  const bool act = (t < tend) && (nstep < nstepmax);
  if (!act) {
    if (threadIdx.x == 0) ctl->active = 0;
    return;  // wave-uniform
  }
  if ((nstep % 2) == 0) {
    double dt = wave_reduce_dt(cour, kCourSlots, cf, dx, smallc);
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
#endif  // HYDRO_BASE_UNIT

template <typename S>
__global__ __launch_bounds__(256) void k_init(GridGeom g, GridView<S> a, GridView<S> b, int bx,
                                              int by0, int by1, S eb) {
  const long t = (long) blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long) (g.nx + 4) * (g.ny + 4)) return;
  init_cell<S>(g, a, b, t, bx, by0, by1, eb);
}

#ifdef HYDRO_BASE_UNIT
// ---- plain copy (measurement reference) ---------------------------------------

// Streams n doubles from a to b; used to calibrate HBM byte counters and as
// the achievable-copy-bandwidth reference.
__global__ __launch_bounds__(256) void k_copy_probe(const double* __restrict__ a,
                                                    double* __restrict__ b, long n) {
  for (long k = (long) blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (long) gridDim.x * blockDim.x)
    b[k] = a[k];
}

void launch_copy_probe(const double* a, double* b, long n, hipStream_t st) {
  hipLaunchKernelGGL(k_copy_probe, dim3(8192), dim3(256), 0, st, a, b, n);
}

// The sweeps' access shape as a plain copy: four arrays of n doubles in, four
// out, one cell per thread, non-temporal loads and stores (the fastest such
// copy measured: 5.7-6.2 TB/s on MI355X for 8196^2 cells, 5.5 TB/s without
// the hints).  Its rate is the bandwidth of the sweeps' bytes with no
// arithmetic and a linear access pattern.
__global__ __launch_bounds__(256) void k_stream_probe(StreamProbe4 p, long n) {
  const long i = (long) blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v[4];
#pragma unroll
  for (int k = 0; k < 4; k++) v[k] = __builtin_nontemporal_load(&p.in[k][i]);
#pragma unroll
  for (int k = 0; k < 4; k++) __builtin_nontemporal_store(v[k], &p.out[k][i]);
}

void launch_stream_probe(const StreamProbe4& p, long n, hipStream_t st) {
  hipLaunchKernelGGL(k_stream_probe, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, st, p, n);
}
#endif  // HYDRO_BASE_UNIT

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

// The sweep launch for one arithmetic form FD (shapes as in launch_sweep).
template <typename S, int FD>
static void launch_sweep_fd(int idim, const GridGeom& g, const Phys<S>& ph, double dx,
                            GridView<S> in, GridView<S> out, double dt_val, const RunCtl* ctl,
                            uint64_t* cour, const SweepTune& tune, bool st1, hipStream_t st) {
  const bool c = cour != nullptr;
#define DISPATCH(F)                         \
  do {                                      \
    if (st1) {                              \
      if (c) F(true, true); else F(true, false);    \
    } else {                                \
      if (c) F(false, true); else F(false, false);  \
    }                                       \
  } while (0)
  if (idim == 1 && (tune.xwaves == 2 || tune.xwaves == 4)) {
    const int W = tune.xwaves;
    const int nseg = (int) cdiv(g.nx, 64 * W - 4);
    const long pieces = (long) nseg * g.ny;
    int xiter = tune.xiter > 0 ? tune.xiter : 8;
    while (xiter > 1 && cdiv(pieces, xiter) < 1024) xiter >>= 1;
    const dim3 grid(cdiv(pieces, xiter)), blk(64 * W);
#define HXB(A, B, WW) \
  hipLaunchKernelGGL((k_sweep_xb<S, A, B, WW, FD>), grid, blk, 0, st, g, ph, dx, dt_val, ctl, \
                     in, out, cour, nseg, xiter)
#define HXW(A, B) do { if (W == 2) HXB(A, B, 2); else HXB(A, B, 4); } while (0)
    DISPATCH(HXW);
#undef HXW
#undef HXB
  } else if (idim == 1) {
    const int nseg = (int) cdiv(g.nx, kXSeg);
    const long waves = (long) nseg * g.ny;
    const int xw = tune.xwaves > 0 ? tune.xwaves : 4;
    int xiter = tune.xiter > 0 ? tune.xiter : 8;
    // keep at least ~4 workgroups per CU before walking several segments
    while (xiter > 1 && cdiv(waves, (long) xw * xiter) < 1024) xiter >>= 1;
    const dim3 grid(cdiv(waves, (long) xw * xiter)), blk(64 * xw);
#define HX(A, B) \
  hipLaunchKernelGGL((k_sweep_x<S, A, B, FD>), grid, blk, 0, st, g, ph, dx, dt_val, ctl, in, \
                     out, cour, nseg, xiter)
    DISPATCH(HX);
#undef HX
  } else {
    int yseg = tune.yseg > 0 ? tune.yseg : 64;
    const int yw = tune.ywaves > 0 ? tune.ywaves : 1;
    const int ncolblk = (int) cdiv(g.nx, 64);
    // a narrow or short grid has few columns to march: shorten the segments
    // (down to 16 rows; each adds a 4-row halo) until there are about two
    // waves per SIMD, rather than leave most of the GPU idle
    while (yseg > 16 && (long) ncolblk * cdiv(g.ny, yseg) < 2048) yseg >>= 1;
    const int nsg = (int) cdiv(g.ny, (long) yseg * yw);
#define HY(A, B)                                                                             \
  hipLaunchKernelGGL((k_sweep_y<S, A, B, FD>), dim3((unsigned) ncolblk * nsg), dim3(64 * yw), \
                     yw * sizeof(YStash<S>), st, g, ph, dx, dt_val, ctl, in, out, cour, yseg,  \
                     ncolblk)
    DISPATCH(HY);
#undef HY
  }
#undef DISPATCH
}

#ifndef HYDRO_FAST_UNIT
template <typename S>
void launch_sweep(int idim, const GridGeom& g, const PhysArgs& pa, GridView<S> in,
                  GridView<S> out, double dt_val, const RunCtl* ctl, uint64_t* cour,
                  const SweepTune& tune, hipStream_t st) {
  const Phys<S> ph = make_phys<S>(pa.gamma, pa.smallr, pa.smallc, pa.slope_type, pa.niter,
                                  pa.iorder, pa.scheme);
  const bool st1 = pa.slope_type == 1.0;
  // The range-gated FP64 divisions (the kGated kernels) pay when the launch
  // keeps the SIMDs busy.  A sweep of a small grid has too few waves for
  // that: each wave's march is latency-bound, and the gates' wave-wide
  // tests, branches and larger code lengthen it, so such launches run the
  // kIeee kernels (the results are the same bits either way).
  constexpr long kFastDivWaves = 4096;  // 4 waves per SIMD on 256 CUs
  if constexpr (sizeof(S) == 8) {
    const bool fd = ph.fastdiv &&
                    (idim == 1 ? (long) cdiv(g.nx, kXSeg) * g.ny
                               : (long) cdiv(g.nx, 64) * cdiv(g.ny, 16)) >= kFastDivWaves;
    if (fd)
      launch_sweep_fd<S, kGated>(idim, g, ph, pa.dx, in, out, dt_val, ctl, cour, tune, st1, st);
    else
      launch_sweep_fd<S, kIeee>(idim, g, ph, pa.dx, in, out, dt_val, ctl, cour, tune, st1, st);
  } else {
    launch_sweep_fd<S, kIeee>(idim, g, ph, pa.dx, in, out, dt_val, ctl, cour, tune, st1, st);
  }
}
#else
void launch_sweep_fast(int idim, const GridGeom& g, const PhysArgs& pa, GridView<double> in,
                       GridView<double> out, double dt_val, const RunCtl* ctl, uint64_t* cour,
                       const SweepTune& tune, hipStream_t st) {
  const Phys<double> ph = make_phys<double>(pa.gamma, pa.smallr, pa.smallc, pa.slope_type,
                                            pa.niter, pa.iorder, pa.scheme);
  launch_sweep_fd<double, kFast>(idim, g, ph, pa.dx, in, out, dt_val, ctl, cour, tune,
                                 pa.slope_type == 1.0, st);
}
#endif

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

#ifdef HYDRO_BASE_UNIT
void launch_dt_finalize(uint64_t* cour, double cf, double dx, double smallc, double* dt_out,
                        hipStream_t st) {
  hipLaunchKernelGGL(k_dt_finalize, dim3(1), dim3(64), 0, st, cour, cf, dx, smallc, dt_out);
}

void launch_ctl_begin(RunCtl* ctl, uint64_t* cour, double cf, double dx, double smallc,
                      hipStream_t st) {
  hipLaunchKernelGGL(k_ctl_begin, dim3(1), dim3(64), 0, st, ctl, cour, cf, dx, smallc);
}

void launch_ctl_end(RunCtl* ctl, hipStream_t st) {
  hipLaunchKernelGGL(k_ctl_end, dim3(1), dim3(64), 0, st, ctl);
}

void launch_clear_slots(uint64_t* cour, hipStream_t st) {
  hipLaunchKernelGGL(k_clear_slots, dim3(1), dim3(256), 0, st, cour);
}
#endif  // HYDRO_BASE_UNIT

template <typename S>
void launch_init(const GridGeom& g, GridView<S> a, GridView<S> b, int bx, int by0, int by1,
                 double e_blast, hipStream_t st) {
  const long n = (long) (g.nx + 4) * (g.ny + 4);
  hipLaunchKernelGGL(k_init<S>, dim3(cdiv(n, 256)), dim3(256), 0, st, g, a, b, bx, by0, by1,
                     (S) e_blast);
}
#ifdef HYDRO_BASE_UNIT
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
template void launch_courant<double>(const GridGeom&, const PhysArgs&, GridView<double>,
                                     uint64_t*, const RunCtl*, hipStream_t);
#elif defined(HYDRO_MIXED_UNIT)
template void launch_sweep<float>(int, const GridGeom&, const PhysArgs&, GridView<float>,
                                  GridView<float>, double, const RunCtl*, uint64_t*,
                                  const SweepTune&, hipStream_t);
template void launch_courant<float>(const GridGeom&, const PhysArgs&, GridView<float>,
                                    uint64_t*, const RunCtl*, hipStream_t);
#endif

}  // namespace hydro
