// Batched kernels: many independent grids (instances) of uneven sizes advanced
// per launch.  The arithmetic per cell is the single-grid code
// (hydro_sweeps.hpp), so every instance's result is bit-identical to running
// it alone.
//
// A step is two phases.  In each phase every instance does one sweep, x or y
// depending on its own step parity (even steps x then y, odd steps y then x,
// HydroC99_2DMpi/Src/main.c:285-291), so a phase's task table mixes x tasks
// of some instances with y tasks of others (the x and y tasks run as separate
// launches, and each list as two: instances with slope_type == 1 first, in
// the specialised variant).  Workgroups are persistent: the grid is sized to
// the GPU, and the workgroups pull task indices from an atomic counter.  A
// task is one workgroup's worth of one instance:
//   x task: wpb * xiter consecutive 60-cell row segments (each wave walks xiter)
//   y task: 64 columns x (wpb * yseg) rows (each wave marches yseg rows; the
//           instance's yseg evens its rows out over its row groups)
// so the CFL reduction of a task never mixes instances.  Run control (t, dt,
// nstep, active) and CFL slots are per instance, on the device.
//
// Compiled three times, as hydro_kernels.hip: the FP64 unit
// (-ffp-contract=off), the mixed-precision unit (HYDRO_MIXED_UNIT,
// -ffp-contract=on) with only the FP32-storage sweeps and CFL kernels, and
// the fast unit (HYDRO_FAST_UNIT, -ffp-contract=on) with only the FP64
// fast-form sweeps.
#include "hydro_batch.hpp"
#include "hydro_sweeps.hpp"

namespace hydro {

// The next task index of a sweep launch for the whole workgroup: persistent
// workgroups pull tasks from one counter (zeroed before the launch), so a
// workgroup that drew cheap tasks takes more instead of idling while others
// finish.  Every index past the table ends the loop of the workgroup that
// drew it, so the launch drains.
__device__ __forceinline__ int next_task(uint64_t* qnext) {
  __shared__ int s_next;
  if (threadIdx.x == 0) s_next = (int) atomicAdd((unsigned long long*) qnext, 1ULL);
  __syncthreads();
  const int t = s_next;
  __syncthreads();  // every thread has read s_next before it is rewritten
  return t;
}

// x tasks of one phase (every task's kind is 1).  ST1: every listed
// instance has slope_type == 1 (the host splits each phase's lists), so the
// limiter's multiply/divide by slope_type is compiled out as in single grids.
// FD: the arithmetic form (persistent workgroups keep the SIMDs busy, so the
// FP64 batches take the range-gated divisions, kGated, for every instance).
// Built for 5 waves per SIMD, as the single-grid x sweep (the per-task state
// would otherwise take it to 98 VGPRs and 4 waves; one register spills).
template <typename S, bool ST1, int FD>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void k_batch_sweep_x(
    const InstDev<S>* __restrict__ inst, const BTask* __restrict__ tasks, int ntask, int phase,
    int xiter, uint64_t* qnext) {
  const int wpb = blockDim.x >> 6;
  const int wave = wave_id();
  for (;;) {
    const int t = next_task(qnext);
    if (t >= ntask) break;
    const BTask tk = tasks[t];
    const InstDev<S>& d = inst[tk.inst];
    if (!d.ctl->active) continue;  // workgroup-uniform
    const double dtdx = d.ctl->dt / d.dx;
    const TraceK<S> tr = make_tracek<S>((S) dtdx, d.ph.scheme);
    const GridView<S> in = d.buf[phase], out = d.buf[1 - phase];
    // the second sweep of an odd step is an x sweep; it produces the CFL
    // maxima the next (even) step's dt needs
    const bool cour = phase == 1;
    double cx = 0.0, cy = 0.0;
    // segments tk.a + it * wpb + wave, it = 0 .. xiter-1
    SegWalk w(tk.a + wave, d.nseg);
    for (int it = 0; it < xiter && w.row < d.g.ny; it++) {
      if (cour)
        x_segment<S, ST1, true, FD>(d.g, d.ph, tr, dtdx, in, out, w.row, w.seg, cx, cy);
      else
        x_segment<S, ST1, false, FD>(d.g, d.ph, tr, dtdx, in, out, w.row, w.seg, cx, cy);
      w.advance(wpb, d.nseg);
    }
    if (cour) commit_courant(cx, cy, d.cour, kInstSlots, t % kInstSlots);
  }
}

// y tasks of one phase (every task's kind is 2); y sweeps never end a step
// whose successor needs dt, so they fold no CFL maxima.
template <typename S, bool ST1, int FD>
__global__ __launch_bounds__(256) void k_batch_sweep_y(const InstDev<S>* __restrict__ inst,
                                                       const BTask* __restrict__ tasks,
                                                       int ntask, int phase, int yseg,
                                                       uint64_t* qnext) {
  const int wpb = blockDim.x >> 6;
  const int wave = wave_id();
  for (;;) {
    const int t = next_task(qnext);
    if (t >= ntask) break;
    const BTask tk = tasks[t];
    const InstDev<S>& d = inst[tk.inst];
    if (!d.ctl->active) continue;
    const double dtdx = d.ctl->dt / d.dx;
    const TraceK<S> tr = make_tracek<S>((S) dtdx, d.ph.scheme);
    const GridView<S> in = d.buf[phase], out = d.buf[1 - phase];
    const int col = tk.a * 64 + (threadIdx.x & 63);
    const bool colok = col < d.g.nx;
    const long i = 2 + (colok ? col : d.g.nx - 1);
    const int j0 = 2 + (tk.b * wpb + wave) * d.yseg;
    double cx = 0.0, cy = 0.0;
    if (j0 < d.g.ny + 2) {  // wave-uniform
      const int j1 = min(j0 + d.yseg, d.g.ny + 2);
      extern __shared__ __attribute__((aligned(16))) unsigned char ylds[];
      YStash<S>* ys = reinterpret_cast<YStash<S>*>(ylds) + wave;
      y_march<S, ST1, false, FD>(d.g, d.ph, tr, dtdx, in, out, colok, i, j0, j1, ys, cx, cy);
    }
  }
}

// Ghost fills (kind = direction) of (line, var) pairs [a, a + kBoundaryChunk)
// of one instance per task, from buffer `phase` into both buffers.
template <typename S>
__global__ __launch_bounds__(256) void k_batch_boundary(const InstDev<S>* inst,
                                                        const BTask* tasks, int ntask,
                                                        int phase) {
  for (int t = blockIdx.x; t < ntask; t += gridDim.x) {
    const BTask tk = tasks[t];
    const InstDev<S>& d = inst[tk.inst];
    if (!d.ctl->active) continue;
    const int idim = tk.kind;
    const int nl = (idim == 1) ? d.g.ny : d.g.nx;
    const GridView<S> src = d.buf[phase], dst = d.buf[1 - phase];
    const int kend = min(tk.a + kBoundaryChunk, 4 * nl);
    for (int k = tk.a + threadIdx.x; k < kend; k += blockDim.x) {
      const int line = k % nl, v = k / nl;
      boundary_line<S>(idim, d.g, d.bl, d.br, d.bd, d.bu, src.p[v], dst.p[v], v, line);
    }
  }
}

// CFL maxima of buffer 0 for the listed instances (admission / step 0).
template <typename S>
__global__ __launch_bounds__(256) void k_batch_courant(const InstDev<S>* inst, const int* ids,
                                                       int nids) {
  for (int t = blockIdx.x; t < nids * 64; t += gridDim.x) {
    const InstDev<S>& d = inst[ids[t / 64]];
    const int part = t % 64;
    const long n = (long) d.g.nx * d.g.ny;
    double cx = 0.0, cy = 0.0;
    for (long c = (long) part * blockDim.x + threadIdx.x; c < n; c += 64L * blockDim.x) {
      const long jj = c / d.g.nx, ii = c - jj * d.g.nx;
      const long o = (jj + 2) * d.g.pitch + ii + 2;
      S ax, ay;
      courant_cell<S>(d.buf[0].p[0][o], d.buf[0].p[1][o], d.buf[0].p[2][o], d.buf[0].p[3][o],
                      d.ph, ax, ay);
      cx = maxm(cx, (double) ax);
      cy = maxm(cy, (double) ay);
    }
    commit_courant(cx, cy, d.cour, kInstSlots, part % kInstSlots);
  }
}

// main.c:244-260 for every listed instance: one wave per instance.  A batch
// also stops an instance at its first dt that is not positive and finite
// (the reference would go on; this only differs for runs that have already
// broken down) and flags it, so one bad run cannot stall the others.
template <typename S>
__global__ __launch_bounds__(64) void k_batch_ctl_begin(const InstDev<S>* inst, const int* ids,
                                                        int nids, uint64_t* qctr) {
  // the step's sweep launches start their task queues from 0
  if (blockIdx.x == 0 && threadIdx.x < kQueues) qctr[threadIdx.x] = 0;
  for (int k = blockIdx.x; k < nids; k += gridDim.x) {
    const InstDev<S>& d = inst[ids[k]];
    RunCtl* c = d.ctl;
    const double t = c->t, tend = c->tend;
    const int nstep = c->nstep, nstepmax = c->nstepmax;
    bool act = (t < tend) && (nstep < nstepmax) && !c->fault;
    bool bad = false;
    if (act && (nstep % 2) == 0) {
      double dt = wave_reduce_dt(d.cour, kInstSlots, d.cf, d.dx, d.smallc);
      if (nstep == 0) dt = dt / 2.0;
      bad = !(dt > 0.0) || !__builtin_isfinite(dt);
      if (threadIdx.x == 0 && !bad) c->dt = dt;
    }
    if (threadIdx.x == 0) {
      if (bad) c->fault = 1;
      c->active = (act && !bad) ? 1 : 0;
    }
  }
}

// main.c:297-298
template <typename S>
__global__ void k_batch_ctl_end(const InstDev<S>* inst, const int* ids, int nids) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nids) return;
  RunCtl* c = inst[ids[k]].ctl;
  if (!c->active) return;
  c->nstep += 1;
  c->t += c->dt;
}

[[maybe_unused]] static inline unsigned cdivu(long a, long b) {
  return (unsigned) ((a + b - 1) / b);
}

#ifdef HYDRO_BASE_UNIT
__global__ void k_batch_set_ctl(RunCtl* c, double t, double dt, double tend, int nstep,
                                int nstepmax, int active) {
  if (threadIdx.x != 0) return;
  c->t = t;
  c->dt = dt;
  c->tend = tend;
  c->nstep = nstep;
  c->nstepmax = nstepmax;
  c->active = active;
  c->log = nullptr;
  c->log_cap = 0;
  c->log_base = 0;
  c->fault = 0;
}

__global__ void k_fill_u64(uint64_t* p, long n) {
  for (long k = (long) blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (long) gridDim.x * blockDim.x)
    p[k] = 0;
}

void launch_batch_set_ctl(RunCtl* c, double t, double dt, double tend, int nstep, int nstepmax,
                          int active, hipStream_t st) {
  hipLaunchKernelGGL(k_batch_set_ctl, dim3(1), dim3(64), 0, st, c, t, dt, tend, nstep, nstepmax,
                     active);
}

void launch_zero_u64(uint64_t* p, long n, hipStream_t st) {
  if (n <= 0) return;
  long nb = (n + 255) / 256;
  if (nb > 1024) nb = 1024;
  hipLaunchKernelGGL(k_fill_u64, dim3((unsigned) nb), dim3(256), 0, st, p, n);
}
#endif  // HYDRO_BASE_UNIT

// The arithmetic form of this unit's batched sweeps.
template <typename S>
constexpr int unit_form() {
#ifdef HYDRO_FAST_UNIT
  return kFast;
#else
  return sizeof(S) == 8 ? kGated : kIeee;
#endif
}

template <typename S, int FD>
static int occupancy_fd(int kind, int waves) {
  int nb = 0;
  const hipError_t e =
      kind == 1 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
                      &nb, k_batch_sweep_x<S, true, FD>, 64 * waves, 0)
                : hipOccupancyMaxActiveBlocksPerMultiprocessor(
                      &nb, k_batch_sweep_y<S, true, FD>, 64 * waves, waves * sizeof(YStash<S>));
  if (e != hipSuccess) return 2;
  return nb > 0 ? nb : 1;
}

template <typename S, bool ST1, int FD>
static void launch_sweep_part(int kind, const InstDev<S>* inst, const BTask* tasks, int ntask,
                              int phase, const BatchTune& tn, int nblocks, uint64_t* q,
                              hipStream_t st) {
  if (ntask <= 0) return;
  const int nb = ntask < nblocks ? ntask : nblocks;
  if (kind == 1)
    hipLaunchKernelGGL((k_batch_sweep_x<S, ST1, FD>), dim3(nb), dim3(64 * tn.waves), 0, st, inst,
                       tasks, ntask, phase, tn.xiter, q);
  else
    hipLaunchKernelGGL((k_batch_sweep_y<S, ST1, FD>), dim3(nb), dim3(64 * tn.waves),
                       tn.waves * sizeof(YStash<S>), st, inst, tasks, ntask, phase, tn.yseg, q);
}

template <typename S, int FD>
static void launch_batch_sweep_fd(int kind, const InstDev<S>* inst, const BTask* tasks,
                                  int ntask, int nst1, int phase, const BatchTune& tn,
                                  int nblocks, uint64_t* q, hipStream_t st) {
  launch_sweep_part<S, true, FD>(kind, inst, tasks, nst1, phase, tn, nblocks, q, st);
  launch_sweep_part<S, false, FD>(kind, inst, tasks + nst1, ntask - nst1, phase, tn, nblocks,
                                  q + 1, st);
}

#ifndef HYDRO_FAST_UNIT
template <typename S>
int batch_sweep_occupancy(int kind, int waves) {
  return occupancy_fd<S, unit_form<S>()>(kind, waves);
}

template <typename S>
void launch_batch_sweep(int kind, const InstDev<S>* inst, const BTask* tasks, int ntask,
                        int nst1, int phase, const BatchTune& tn, int nblocks, uint64_t* q,
                        hipStream_t st) {
  launch_batch_sweep_fd<S, unit_form<S>()>(kind, inst, tasks, ntask, nst1, phase, tn, nblocks, q,
                                           st);
}
#ifdef HYDRO_BASE_UNIT
template int batch_sweep_occupancy<double>(int, int);
#else
template int batch_sweep_occupancy<float>(int, int);
#endif
#else
int batch_sweep_occupancy_fast(int kind, int waves) {
  return occupancy_fd<double, kFast>(kind, waves);
}

void launch_batch_sweep_fast(int kind, const InstDev<double>* inst, const BTask* tasks,
                             int ntask, int nst1, int phase, const BatchTune& tn, int nblocks,
                             uint64_t* q, hipStream_t st) {
  launch_batch_sweep_fd<double, kFast>(kind, inst, tasks, ntask, nst1, phase, tn, nblocks, q, st);
}
#endif

template <typename S>
void launch_batch_boundary(const InstDev<S>* inst, const BTask* tasks, int ntask, int phase,
                           int nblocks, hipStream_t st) {
  if (ntask <= 0) return;
  const int nb = ntask < nblocks ? ntask : nblocks;
  hipLaunchKernelGGL(k_batch_boundary<S>, dim3(nb), dim3(256), 0, st, inst, tasks, ntask, phase);
}

template <typename S>
void launch_batch_courant(const InstDev<S>* inst, const int* ids, int nids, hipStream_t st) {
  if (nids <= 0) return;
  long nb = 64L * nids;
  if (nb > 4096) nb = 4096;
  hipLaunchKernelGGL(k_batch_courant<S>, dim3((unsigned) nb), dim3(256), 0, st, inst, ids, nids);
}

template <typename S>
void launch_batch_ctl(const InstDev<S>* inst, const int* ids, int nids, bool begin,
                      uint64_t* qctr, hipStream_t st) {
  if (nids <= 0) return;
  if (begin) {
    const int nb = nids < 1024 ? nids : 1024;
    hipLaunchKernelGGL(k_batch_ctl_begin<S>, dim3(nb), dim3(64), 0, st, inst, ids, nids, qctr);
  } else {
    hipLaunchKernelGGL(k_batch_ctl_end<S>, dim3(cdivu(nids, 256)), dim3(256), 0, st, inst, ids,
                       nids);
  }
}

#define INST_ARITH(S)                                                                            \
  template void launch_batch_sweep<S>(int, const InstDev<S>*, const BTask*, int, int, int,      \
                                      const BatchTune&, int, uint64_t*, hipStream_t);             \
  template void launch_batch_courant<S>(const InstDev<S>*, const int*, int, hipStream_t);
#define INST_MOVE(S)                                                                             \
  template void launch_batch_boundary<S>(const InstDev<S>*, const BTask*, int, int, int,        \
                                         hipStream_t);                                           \
  template void launch_batch_ctl<S>(const InstDev<S>*, const int*, int, bool, uint64_t*,      \
                                    hipStream_t);
#if defined(HYDRO_BASE_UNIT)
INST_ARITH(double)
INST_MOVE(double)
INST_MOVE(float)
#elif defined(HYDRO_MIXED_UNIT)
INST_ARITH(float)
#endif
#undef INST_ARITH
#undef INST_MOVE

}  // namespace hydro
