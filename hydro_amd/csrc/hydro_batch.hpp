// Internal interface of the batched kernels (hydro_batch.hip).
#pragma once
#include "hydro_kernels.hpp"
#include "hydro_numerics.hpp"

namespace hydro {

constexpr int kInstSlots = 64;        // CFL max slots per instance
constexpr int kBoundaryChunk = 1024;  // (line, var) ghost fills per boundary task
// Task-queue counters: [phase][x, y][slope_type == 1 variant, other]
constexpr int kQueues = 8;

// One instance as the batched kernels see it (device array, rebuilt by the
// host whenever the set of instances changes).
template <typename S>
struct InstDev {
  GridGeom g;
  Phys<S> ph;
  double dx, cf, smallc;  // dt = cf*dx / max(cx, cy, smallc)
  GridView<S> buf[2];     // state in buf[0] between steps
  RunCtl* ctl;
  uint64_t* cour;         // 2 * kInstSlots
  int bl, br, bd, bu;     // boundary types
  int nseg;               // 60-cell segments per row
  int yseg;               // rows each wave marches in this instance's y tasks
};

// One workgroup's work: kind 1 = x (a = first segment), 2 = y (a = 64-column
// block, b = row-segment group); for ghost fills kind is the direction.
struct BTask {
  int inst, kind, a, b;
};

struct BatchTune {
  int waves;  // waves per workgroup (1..4)
  int xiter;  // segments each wave walks in an x task
  int yseg;   // at most this many rows per wave in a y task (each instance
              // splits its rows evenly into row groups of waves x yseg)
};

// kind 1: x tasks, 2: y tasks (a phase launches both, on their own tables).
// The first nst1 tasks belong to instances with slope_type == 1.  q: the two
// task-queue counters of this list (zero at launch).
template <typename S>
void launch_batch_sweep(int kind, const InstDev<S>* inst, const BTask* tasks, int ntask,
                        int nst1, int phase, const BatchTune& tn, int nblocks, uint64_t* q,
                        hipStream_t st);
template <typename S>
void launch_batch_boundary(const InstDev<S>* inst, const BTask* tasks, int ntask, int phase,
                           int nblocks, hipStream_t st);
template <typename S>
void launch_batch_courant(const InstDev<S>* inst, const int* ids, int nids, hipStream_t st);
template <typename S>
void launch_batch_ctl(const InstDev<S>* inst, const int* ids, int nids, bool begin,
                      uint64_t* qctr, hipStream_t st);  // begin also zeroes qctr[0..kQueues)

void launch_batch_set_ctl(RunCtl* c, double t, double dt, double tend, int nstep, int nstepmax,
                          int active, hipStream_t st);
void launch_zero_u64(uint64_t* p, long n, hipStream_t st);
template <typename S>
int batch_sweep_occupancy(int kind, int waves);  // resident workgroups per CU
// The FP64 fast form's batched sweeps (hydro_batch.hip, fast unit): as
// launch_batch_sweep / batch_sweep_occupancy for double, with the kFast
// arithmetic.
void launch_batch_sweep_fast(int kind, const InstDev<double>* inst, const BTask* tasks,
                             int ntask, int nst1, int phase, const BatchTune& tn, int nblocks,
                             uint64_t* q, hipStream_t st);
int batch_sweep_occupancy_fast(int kind, int waves);

}  // namespace hydro
