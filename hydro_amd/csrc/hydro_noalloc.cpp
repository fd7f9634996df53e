// Allocation-free CFL time step on caller-owned device memory -- the C ABI
// hydro_amd_compute_deltat_noalloc of include/hydro_amd.h.
//
// The kernels are those of hydro_amd_grid_compute_deltat (k_courant, then
// k_dt_finalize; hydro_kernels.hip).  Their scratch, the CFL max-slots, is a
// file-scope device array here instead of a grid's hipMalloc'ed buffer, so a
// call allocates nothing: it queues three kernels on the caller's stream and
// returns, which also makes it capturable into a HIP graph.
//
// The array exists once per GPU (each device holds its own copy until the
// process ends).  Host code never passes the array's name to a kernel: its
// device address is looked up once per device with hipGetSymbolAddress, with
// that device current, and that address is what the launches take.  All calls
// on one GPU share the array: calls that overlap on one GPU (different
// streams, or different host threads) race on it; calls queued in order on
// one stream do not.  Each call clears the slots first, with a fill kernel
// rather than a memset node.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/hydro_amd.h"
#include "hydro_kernels.hpp"

using namespace hydro;

extern "C" int hydro_amd_set_error(const char* fmt, ...);  // hydro_capi.cpp

namespace {
// 2 * kCourSlots 64-bit words (x and y maxima): 16 KiB per GPU.
constexpr int kScratchWords = 2 * kCourSlots;
static __device__ __attribute__((aligned(256))) uint64_t scratch[kScratchWords];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
uint64_t* g_scratch[kMaxDevices] = {};  // device address of `scratch`, per device
}  // namespace

// The current device's address of `scratch` (looked up on its first use).
static int scratch_of_current_device(uint64_t** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hydro_amd_set_error("hipGetDevice: %s", hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices)
    return hydro_amd_set_error("device %d: at most %d devices supported", dev, kMaxDevices);
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return hydro_amd_set_error("hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch[dev] = static_cast<uint64_t*>(p);
  }
  *out = g_scratch[dev];
  return 0;
}

extern "C" int hydro_amd_compute_deltat_noalloc(const hydro_amd_params* p, const double* d_uold,
                                                double* d_dt, void* stream) {
  if (!p || !d_uold || !d_dt) return hydro_amd_set_error("null argument");
  if (p->nx < 1 || p->ny < 1)
    return hydro_amd_set_error("nx, ny must be >= 1 (got %d, %d)", p->nx, p->ny);
  // the padded grid is addressed with 64-bit indices; the bound matches the
  // one on grids (hydro_amd_grid_create)
  if ((long) (p->nx + 4) * (p->ny + 4) > (1L << 40))
    return hydro_amd_set_error("grid too large: (nx+4)*(ny+4) must be <= 2^40");
  if (!(p->dx > 0)) return hydro_amd_set_error("dx must be > 0");
  if (!(p->gamma > 1)) return hydro_amd_set_error("gamma must be > 1");
  if (p->precision != HYDRO_AMD_FP64)
    return hydro_amd_set_error("compute_deltat_noalloc takes an FP64 state (precision %d)",
                               p->precision);
  uint64_t* slots = nullptr;
  if (scratch_of_current_device(&slots)) return -1;

  // the reference layout [4][ny+4][nx+4], x fastest
  GridGeom g;
  g.nx = p->nx;
  g.ny = p->ny;
  g.pitch = (long) p->nx + 4;
  const long plane = g.pitch * (p->ny + 4);
  GridView<double> v;
  for (int k = 0; k < 4; k++) v.p[k] = const_cast<double*>(d_uold) + k * plane;  // read only
  PhysArgs ph = {};
  ph.gamma = p->gamma;
  ph.smallr = p->smallr;
  ph.smallc = p->smallc;
  ph.slope_type = p->slope_type;
  ph.courant_factor = p->courant_factor;
  ph.dx = p->dx;
  ph.niter = p->niter_riemann;
  ph.iorder = p->iorder;
  ph.scheme = p->scheme;
  hipStream_t st = static_cast<hipStream_t>(stream);
  launch_clear_slots(slots, st);
  launch_courant<double>(g, ph, v, slots, nullptr, st);
  launch_dt_finalize(slots, p->courant_factor, p->dx, p->smallc, d_dt, st);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hydro_amd_set_error("launch: %s", hipGetErrorString(e));
  return 0;
}
