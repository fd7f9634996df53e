// Internal (C++) interface between the C-ABI layer and the HIP kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// The kernel files (hydro_kernels.hip, hydro_batch.hip) are compiled once per
// arithmetic unit: HYDRO_MIXED_UNIT (FP32-storage sweeps), HYDRO_FAST_UNIT
// (FP64 fast-form sweeps) or neither (everything else, FP64 bitwise).
#if !defined(HYDRO_MIXED_UNIT) && !defined(HYDRO_FAST_UNIT)
#define HYDRO_BASE_UNIT 1
#endif

namespace hydro {

// Device-resident time-loop state (HydroC99_2DMpi/Src/main.c:244-298 moved
// onto the GPU so whole runs can be queued or graph-captured without host
// round trips).
struct RunCtl {
  double t, dt, tend;
  int nstep, nstepmax;
  int active;    // 1 while the current step runs (t < tend && nstep < nstepmax)
  int log_base;  // dt of step n goes to log[n - log_base] when in range
  int log_cap;
  int fault;     // batches: set when a computed dt was not positive and finite
  double* log;
};

// One grid as the kernels see it.  Element (i, j, v) of a state buffer lives
// at plane[v][j * pitch + i], i in [0, nx+4), j in [0, ny+4) (padded indices
// with two ghost layers, as in the reference).
template <typename S>
struct GridView {
  S* p[4];
};

struct GridGeom {
  int nx, ny;
  long pitch;
};

struct PhysArgs {
  double gamma, smallr, smallc, slope_type, courant_factor, dx;
  int niter, iorder, scheme;
  int bl, br, bd, bu;  // boundary types left/right/down/up
};

// Number of CFL max-slots (spread the atomics over many addresses).
constexpr int kCourSlots = 1024;
// x sweep: one 64-lane wave updates 60 cells of a row.
constexpr int kXSeg = 60;

struct SweepTune {
  int yseg;     // rows per y-sweep thread (column march length)
  int ywaves;   // waves per y-sweep workgroup (1..4)
  int xwaves;   // waves per x-sweep workgroup (1..4)
  int xiter;    // row segments each x-sweep wave walks
};

// ---- launchers (hydro_kernels.hip) -----------------------------------------
// dt handling: if ctl != nullptr the kernels read dt and the active flag from
// it; otherwise dt_val is used.  cour (2*kCourSlots uint64, zeroed) receives
// the CFL maxima of the written state when non-null.
template <typename S>
void launch_boundary(int idim, const GridGeom& g, const PhysArgs& ph, GridView<S> src,
                     GridView<S> dst2, const RunCtl* ctl, hipStream_t st);

template <typename S>
void launch_sweep(int idim, const GridGeom& g, const PhysArgs& ph, GridView<S> in,
                  GridView<S> out, double dt_val, const RunCtl* ctl, uint64_t* cour,
                  const SweepTune& tune, hipStream_t st);

// The FP64 fast form's sweeps (hydro_kernels.hip, fast unit): as
// launch_sweep, with the kFast arithmetic.
void launch_sweep_fast(int idim, const GridGeom& g, const PhysArgs& ph, GridView<double> in,
                       GridView<double> out, double dt_val, const RunCtl* ctl, uint64_t* cour,
                       const SweepTune& tune, hipStream_t st);

template <typename S>
void launch_courant(const GridGeom& g, const PhysArgs& ph, GridView<S> in, uint64_t* cour,
                    const RunCtl* ctl, hipStream_t st);

// dt_finalize reduces the CFL slots (clearing them) into
// dt = cfl*dx / max(cx, cy, smallc) at *dt_out.  ctl_begin opens a step of the
// device time loop: active = (t < tend && nstep < nstepmax); on even steps dt
// comes from the slots (halved at step 0).  ctl_end logs dt and advances
// nstep and t.
void launch_dt_finalize(uint64_t* cour, double cf, double dx, double smallc, double* dt_out,
                        hipStream_t st);
void launch_ctl_begin(RunCtl* ctl, uint64_t* cour, double cf, double dx, double smallc,
                      hipStream_t st);
void launch_ctl_end(RunCtl* ctl, hipStream_t st);
void launch_clear_slots(uint64_t* cour, hipStream_t st);

// Initial state (hydro_funcs.c:79-146): interior rho=1, u=v=0, E=1e-5, ghosts
// 0, then E=e_blast at padded cells (bx, by) for by in [by0, by1).  Fills
// both buffers when b is non-null.
template <typename S>
void launch_init(const GridGeom& g, GridView<S> a, GridView<S> b, int bx, int by0, int by1,
                 double e_blast, hipStream_t st);

void launch_copy_probe(const double* a, double* b, long n, hipStream_t st);
struct StreamProbe4 {
  const double* in[4];
  double* out[4];
};
void launch_stream_probe(const StreamProbe4& p, long n, hipStream_t st);

// Density image (hydro_image.hip): clears *key, reduces the maximum density
// of the padded plane `rho` into it, then writes w*h RGB pixels (3 bytes
// each, row cy = 0 first) sampled every `shrink` interior cells.
template <typename S>
void launch_image(const GridGeom& g, const S* rho, int shrink, int w, int h,
                  unsigned long long* key, unsigned char* rgb, hipStream_t st);

// Pitched device rows -> contiguous host rows through pinned staging buffers
// (hydro_stage.cpp); dst_esz == src_esz, or 8 with src_esz == 4 (widening).
int stage_d2h(void* dst, int dst_esz, const void* src, int src_esz, long src_pitch, long width,
              long rows, hipStream_t st);

}  // namespace hydro
