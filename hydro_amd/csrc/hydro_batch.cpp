// Batch of independent grids (instances) advanced per launch -- the C ABI
// hydro_amd_batch_* of include/hydro_amd.h.  Host bookkeeping only; the
// kernels are in hydro_batch.hip.
//
// Instances are admitted against an HBM budget; their state buffers come from
// the stream-ordered allocator (hipMallocAsync / hipFreeAsync on the batch's
// stream), so admitting or retiring an instance never synchronises the
// device.  The per-step work is two phase task tables (P and Q) built from
// each instance's step parity; two steps (P,Q then Q,P) are captured once as a
// HIP graph and replayed until the set of instances changes.
#include <hip/hip_runtime.h>

#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/hydro_amd.h"
#include "hydro_batch.hpp"

using namespace hydro;

extern "C" int hydro_amd_set_error(const char* fmt, ...);  // hydro_capi.cpp

#define BCHK(call)                                                                         \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess) return hydro_amd_set_error("%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

namespace {
constexpr long kOrigin = 30;  // same row origin as single grids (hydro_capi.cpp)
constexpr int kMaxInst = 8192;

struct Inst {
  bool live = false;
  hydro_amd_params p;
  GridGeom g;
  long plane = 0;
  void* mem[2] = {nullptr, nullptr};
  long long bytes = 0;
  bool need_cour = false;  // CFL maxima of the initial state still to compute
};
}  // namespace

struct hydro_amd_batch {
  int prec;
  size_t esz;
  long long budget, used = 0;
  hipStream_t st = nullptr;
  std::vector<Inst> inst;
  RunCtl* ctl = nullptr;         // kMaxInst run controls
  uint64_t* cour = nullptr;      // kMaxInst * 2 * kInstSlots
  void* idev = nullptr;          // InstDev<S>[kMaxInst]
  // task lists, in this order: Px Py (phase P: x of even-step instances, y of
  // odd ones), Qx Qy (phase Q: the other sweep of each), PB QB (ghost fills)
  BTask* tasks = nullptr;
  int task_cap = 0;
  int* ids = nullptr;            // live ids (device)
  int* cids = nullptr;           // ids needing a CFL pass (device)
  int n[6] = {0, 0, 0, 0, 0, 0}; // Px Py Qx Qy PB QB
  int off[6] = {0, 0, 0, 0, 0, 0};
  int n1[6] = {0, 0, 0, 0, 0, 0};  // leading tasks of slope_type == 1 instances
  uint64_t* qctr = nullptr;        // kQueues task-queue counters (device)
  int nids = 0;
  int maxid = -1;                // highest live id
  bool dirty = true;
  bool flip = false;             // next step runs Q before P
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  BatchTune tune = {2, 8, 32};  // set by precision at creation
  int nblk_x = 1024, nblk_y = 1024, cus = 256;
  std::vector<int> live_ids;
  RunCtl* hctl = nullptr;        // pinned host copy of the run controls
  void* hidev = nullptr;         // pinned host copy of the instance array
};

// FP64 storage (the FP64 and FP64 fast forms) rather than FP32.
static bool dbl(const hydro_amd_batch* b) { return b->prec != HYDRO_AMD_MIXED; }

static void size_grids(hydro_amd_batch* b) {
  const int w = b->tune.waves;
  if (b->prec == HYDRO_AMD_FP64_FAST) {
    b->nblk_x = b->cus * batch_sweep_occupancy_fast(1, w);
    b->nblk_y = b->cus * batch_sweep_occupancy_fast(2, w);
  } else if (b->prec == HYDRO_AMD_FP64) {
    b->nblk_x = b->cus * batch_sweep_occupancy<double>(1, w);
    b->nblk_y = b->cus * batch_sweep_occupancy<double>(2, w);
  } else {
    b->nblk_x = b->cus * batch_sweep_occupancy<float>(1, w);
    b->nblk_y = b->cus * batch_sweep_occupancy<float>(2, w);
  }
}

static void drop_graphs(hydro_amd_batch* b) {
  for (auto& g : b->gexec)
    if (g) {
      (void) hipGraphExecDestroy(g);
      g = nullptr;
    }
}

extern "C" void hydro_amd_batch_destroy(hydro_amd_batch* b) {
  if (!b) return;
  if (b->st) (void) hipStreamSynchronize(b->st);
  drop_graphs(b);
  for (auto& in : b->inst)
    for (auto& m : in.mem)
      if (m) (void) hipFree(m);
  for (void* p : {(void*) b->ctl, (void*) b->cour, b->idev, (void*) b->tasks, (void*) b->ids,
                  (void*) b->cids, (void*) b->qctr})
    if (p) (void) hipFree(p);
  if (b->hctl) (void) hipHostFree(b->hctl);
  if (b->hidev) (void) hipHostFree(b->hidev);
  if (b->st) (void) hipStreamDestroy(b->st);
  delete b;
}

extern "C" hydro_amd_batch* hydro_amd_batch_create(int precision, long long budget_bytes) {
  if (precision != HYDRO_AMD_FP64 && precision != HYDRO_AMD_MIXED &&
      precision != HYDRO_AMD_FP64_FAST) {
    hydro_amd_set_error("unknown precision %d", precision);
    return nullptr;
  }
  hydro_amd_batch* b = new hydro_amd_batch();
  b->prec = precision;
  b->esz = dbl(b) ? 8 : 4;
  // defaults (the serving loop runs with them): FP64 (2, 8, 32), what the
  // autotuner picks for the 512-instance mix; mixed (4, 32, 64), the fastest
  // measured on the serving mix (sides 128-4096: 48.4 solves/s against 45.0
  // for (4, 8, 16) at 1000 arrivals/s)
  b->tune = precision == HYDRO_AMD_MIXED ? BatchTune{4, 32, 64} : BatchTune{2, 8, 32};
  b->budget = budget_bytes > 0 ? budget_bytes : (1LL << 62);
  b->inst.resize(kMaxInst);
  const size_t isz = dbl(b) ? sizeof(InstDev<double>) : sizeof(InstDev<float>);
  bool ok = hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) == hipSuccess &&
            hipMalloc(&b->ctl, sizeof(RunCtl) * kMaxInst) == hipSuccess &&
            hipMalloc(&b->cour, sizeof(uint64_t) * 2 * kInstSlots * kMaxInst) == hipSuccess &&
            hipMalloc(&b->idev, isz * kMaxInst) == hipSuccess &&
            hipMalloc(&b->ids, sizeof(int) * kMaxInst) == hipSuccess &&
            hipMalloc(&b->cids, sizeof(int) * kMaxInst) == hipSuccess &&
            hipMalloc(&b->qctr, sizeof(uint64_t) * kQueues) == hipSuccess &&
            hipHostMalloc((void**) &b->hctl, sizeof(RunCtl) * kMaxInst) == hipSuccess &&
            hipHostMalloc(&b->hidev, isz * kMaxInst) == hipSuccess;
  if (!ok) {
    hydro_amd_set_error("batch allocation failed");
    hydro_amd_batch_destroy(b);
    return nullptr;
  }
  launch_zero_u64(b->cour, 2L * kInstSlots * kMaxInst, b->st);
  // keep freed instance memory in the pool instead of returning it
  hipMemPool_t pool;
  int dev = 0;
  if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
    uint64_t thr = ~0ULL;
    (void) hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) == hipSuccess) b->cus = prop.multiProcessorCount;
  size_grids(b);
  if (hipStreamSynchronize(b->st) != hipSuccess) {
    hydro_amd_set_error("batch setup failed");
    hydro_amd_batch_destroy(b);
    return nullptr;
  }
  return b;
}

static long long inst_bytes(const hydro_amd_batch* b, const GridGeom& g, long plane) {
  (void) g;
  return 2LL * (4 * plane + kOrigin + 64) * (long long) b->esz;
}

extern "C" long long hydro_amd_batch_bytes_needed(const hydro_amd_batch* b, int nx, int ny) {
  if (!b || nx < 1 || ny < 1) return -1;
  GridGeom g;
  g.nx = nx;
  g.ny = ny;
  g.pitch = ((long) nx + 4 + 31) / 32 * 32;
  return inst_bytes(b, g, g.pitch * (ny + 4));
}

extern "C" long long hydro_amd_batch_bytes_used(const hydro_amd_batch* b) {
  return b ? b->used : -1;
}

template <typename S>
static GridView<S> view_of(const hydro_amd_batch* b, const Inst& in, int k) {
  GridView<S> v;
  S* base = (S*) in.mem[k] + kOrigin;
  for (int q = 0; q < 4; q++) v.p[q] = base + q * in.plane;
  (void) b;
  return v;
}

extern "C" int hydro_amd_batch_add(hydro_amd_batch* b, const hydro_amd_params* p) {
  if (!b || !p) return hydro_amd_set_error("null argument");
  if (p->nx < 1 || p->ny < 1 || !(p->dx > 0) || !(p->gamma > 1) || p->scheme < 1 ||
      p->scheme > 3 || p->niter_riemann < 0)
    return hydro_amd_set_error("invalid instance parameters");
  if ((long) ((p->nx + kXSeg - 1) / kXSeg) * p->ny > (1L << 30))
    return hydro_amd_set_error("instance too large");
  if (p->precision != b->prec) return hydro_amd_set_error("instance precision differs from batch");
  int id = -1;
  for (int k = 0; k < kMaxInst; k++)
    if (!b->inst[k].live) {
      id = k;
      break;
    }
  if (id < 0) return hydro_amd_set_error("batch full (%d instances)", kMaxInst);
  Inst& in = b->inst[id];
  in.p = *p;
  in.g.nx = p->nx;
  in.g.ny = p->ny;
  in.g.pitch = ((long) p->nx + 4 + 31) / 32 * 32;
  in.plane = in.g.pitch * (p->ny + 4);
  in.bytes = inst_bytes(b, in.g, in.plane);
  if (b->used + in.bytes > b->budget)
    return hydro_amd_set_error("over budget: %lld + %lld > %lld bytes", b->used, in.bytes,
                               b->budget);
  const size_t sz = (size_t) in.bytes / 2;
  for (int k = 0; k < 2; k++) {
    if (hipMallocAsync(&in.mem[k], sz, b->st) != hipSuccess) {
      if (k == 1) (void) hipFreeAsync(in.mem[0], b->st);
      in.mem[0] = in.mem[1] = nullptr;
      return hydro_amd_set_error("device allocation of %zu bytes failed", sz);
    }
    BCHK(hipMemsetAsync(in.mem[k], 0, sz, b->st));
  }
  in.live = true;
  in.need_cour = true;
  b->used += in.bytes;
  launch_zero_u64(b->cour + (size_t) id * 2 * kInstSlots, 2 * kInstSlots, b->st);
  launch_batch_set_ctl(b->ctl + id, 0.0, 0.0, p->tend, 0, p->nstepmax, 0, b->st);
  BCHK(hipGetLastError());
  b->dirty = true;
  return id;
}

static Inst* live(hydro_amd_batch* b, int id) {
  if (!b || id < 0 || id >= kMaxInst || !b->inst[id].live) {
    hydro_amd_set_error("no live instance %d", id);
    return nullptr;
  }
  return &b->inst[id];
}

extern "C" int hydro_amd_batch_remove(hydro_amd_batch* b, int id) {
  Inst* in = live(b, id);
  if (!in) return -1;
  for (auto& m : in->mem) {
    if (m) BCHK(hipFreeAsync(m, b->st));
    m = nullptr;
  }
  in->live = false;
  b->used -= in->bytes;
  b->dirty = true;
  return 0;
}

static int init_inst(hydro_amd_batch* b, int id, int bx, int by0, int by1, double e) {
  Inst* in = live(b, id);
  if (!in) return -1;
  if (dbl(b))
    launch_init<double>(in->g, view_of<double>(b, *in, 0), view_of<double>(b, *in, 1), bx, by0,
                        by1, e, b->st);
  else
    launch_init<float>(in->g, view_of<float>(b, *in, 0), view_of<float>(b, *in, 1), bx, by0, by1,
                       e, b->st);
  launch_zero_u64(b->cour + (size_t) id * 2 * kInstSlots, 2 * kInstSlots, b->st);
  launch_batch_set_ctl(b->ctl + id, 0.0, 0.0, in->p.tend, 0, in->p.nstepmax, 0, b->st);
  BCHK(hipGetLastError());
  in->need_cour = true;
  b->dirty = true;  // parity back to even
  return 0;
}

extern "C" int hydro_amd_batch_init(hydro_amd_batch* b, int id, int testcase) {
  Inst* in = live(b, id);
  if (!in) return -1;
  const int nxt = in->p.nx + 4, nyt = in->p.ny + 4;
  const double eb = 1.0 / in->p.dx / in->p.dx;  // hydro_funcs.c:92
  switch (testcase) {
    case 0: return init_inst(b, id, nxt / 2, nyt / 2, nyt / 2 + 1, eb);
    case 1: return init_inst(b, id, 2, 2, 3, eb);
    case 2: return init_inst(b, id, 2, 2, nyt, eb);
    default: return hydro_amd_set_error("testcase %d not implemented", testcase);
  }
}

extern "C" int hydro_amd_batch_init_blast(hydro_amd_batch* b, int id, int bx, int by, double e) {
  Inst* in = live(b, id);
  if (!in) return -1;
  if (bx < 0 || by < 0 || bx >= in->p.nx || by >= in->p.ny)
    return hydro_amd_set_error("blast cell (%d, %d) outside %dx%d", bx, by, in->p.nx, in->p.ny);
  return init_inst(b, id, bx + 2, by + 2, by + 3, e);
}

extern "C" int hydro_amd_batch_download(hydro_amd_batch* b, int id, double* u) {
  Inst* in = live(b, id);
  if (!in || !u) return in ? hydro_amd_set_error("null argument") : -1;
  const long nxt = in->p.nx + 4, rows = 4L * (in->p.ny + 4);
  const void* src = (const char*) in->mem[0] + kOrigin * b->esz;
  return stage_d2h(u, 8, src, (int) b->esz, in->g.pitch, nxt, rows, b->st) ? -1 : 0;
}

// Where an instance's state lives: `rows` rows of `width` elements of `esz`
// bytes, `pitch` elements apart, starting at *src (layout [4][ny+4][nx+4]).
int hydro_amd_batch_raw_view(hydro_amd_batch* b, int id, const void** src, long* pitch,
                             long* width, long* rows, int* esz) {
  Inst* in = live(b, id);
  if (!in) return -1;
  *src = (const char*) in->mem[0] + kOrigin * b->esz;
  *pitch = in->g.pitch;
  *width = in->p.nx + 4;
  *rows = 4L * (in->p.ny + 4);
  *esz = (int) b->esz;
  return 0;
}

// Copies an instance's state as stored (element size 8 or 4), layout
// [4][ny+4][nx+4]; for the serving loop (no conversion on its critical path).
// The copy is queued on `st` (the batch's stream if null), so it can run
// beside kernels on the batch's stream that do not touch this instance.
int hydro_amd_batch_download_raw(hydro_amd_batch* b, int id, void* dst, int* elem_size,
                                 hipStream_t st) {
  Inst* in = live(b, id);
  if (!in || !dst) return in ? hydro_amd_set_error("null argument") : -1;
  const long nxt = in->p.nx + 4, rows = 4L * (in->p.ny + 4);
  const size_t w = b->esz;
  const void* src = (const char*) in->mem[0] + kOrigin * w;
  if (stage_d2h(dst, (int) w, src, (int) w, in->g.pitch, nxt, rows, st ? st : b->st)) return -1;
  if (elem_size) *elem_size = (int) w;
  return 0;
}

static int rebuild(hydro_amd_batch* b);

// Whether each listed live instance's interior state is finite: one CFL pass
// over it (k_batch_courant) into its slots, whose maxima are NaN or inf
// exactly when some cell's density, momenta or energy is (a NaN orders above
// +inf in the slots' integer maximum).  Meant for instances whose run has
// ended -- it overwrites their CFL slots -- such as the serving loop's
// finished requests.  Waits for the batch's stream.
int hydro_amd_batch_finite(hydro_amd_batch* b, const int* ids, int n, int* finite) {
  if (!b || n < 0 || (n > 0 && (!ids || !finite))) return hydro_amd_set_error("bad argument");
  if (n == 0) return 0;
  if (n > kMaxInst) return hydro_amd_set_error("too many instances");
  BCHK(hipStreamSynchronize(b->st));
  if (b->dirty && rebuild(b)) return -1;
  for (int k = 0; k < n; k++) {
    if (!live(b, ids[k])) return -1;
    launch_zero_u64(b->cour + (size_t) ids[k] * 2 * kInstSlots, 2 * kInstSlots, b->st);
  }
  BCHK(hipMemcpyAsync(b->cids, ids, sizeof(int) * n, hipMemcpyHostToDevice, b->st));
  if (dbl(b))
    launch_batch_courant<double>((const InstDev<double>*) b->idev, b->cids, n, b->st);
  else
    launch_batch_courant<float>((const InstDev<float>*) b->idev, b->cids, n, b->st);
  std::vector<uint64_t> slots((size_t) n * 2 * kInstSlots);
  for (int k = 0; k < n; k++)
    BCHK(hipMemcpyAsync(slots.data() + (size_t) k * 2 * kInstSlots,
                        b->cour + (size_t) ids[k] * 2 * kInstSlots,
                        sizeof(uint64_t) * 2 * kInstSlots, hipMemcpyDeviceToHost, b->st));
  BCHK(hipStreamSynchronize(b->st));  // (also: `ids` and `slots` outlive the copies)
  for (int k = 0; k < n; k++) {
    int ok = 1;
    for (int q = 0; q < 2 * kInstSlots; q++) {
      double v;
      memcpy(&v, &slots[(size_t) k * 2 * kInstSlots + q], sizeof v);
      if (!std::isfinite(v)) ok = 0;
    }
    finite[k] = ok;
  }
  return 0;
}

extern "C" int hydro_amd_batch_upload(hydro_amd_batch* b, int id, const double* u) {
  Inst* in = live(b, id);
  if (!in || !u) return in ? hydro_amd_set_error("null argument") : -1;
  const long nxt = in->p.nx + 4, rows = 4L * (in->p.ny + 4);
  std::vector<float> tmp;
  const void* srcp = u;
  size_t w = 8;
  if (!dbl(b)) {
    tmp.resize((size_t) nxt * rows);
    for (size_t k = 0; k < tmp.size(); k++) tmp[k] = (float) u[k];
    srcp = tmp.data();
    w = 4;
  }
  for (int k = 0; k < 2; k++) {
    void* dst = (char*) in->mem[k] + kOrigin * b->esz;
    BCHK(hipMemcpy2DAsync(dst, in->g.pitch * w, srcp, nxt * w, nxt * w, rows,
                          hipMemcpyHostToDevice, b->st));
  }
  BCHK(hipStreamSynchronize(b->st));
  launch_zero_u64(b->cour + (size_t) id * 2 * kInstSlots, 2 * kInstSlots, b->st);
  launch_batch_set_ctl(b->ctl + id, 0.0, 0.0, in->p.tend, 0, in->p.nstepmax, 0, b->st);
  in->need_cour = true;
  b->dirty = true;
  return 0;
}

// ---- task tables -------------------------------------------------------------------

// Rows per wave in an instance's y tasks: the instance's rows split evenly
// over as few row groups (of waves x tune.yseg rows) as cover them, so no
// group ends with a short segment that pays the full halo.
static int inst_yseg(const hydro_amd_batch* b, int ny) {
  const long W = b->tune.waves, per = W * b->tune.yseg;
  const long nrg = (ny + per - 1) / per;
  return (int) ((ny + W * nrg - 1) / (W * nrg));
}

template <typename S>
static InstDev<S> inst_dev(hydro_amd_batch* b, int id) {
  const Inst& in = b->inst[id];
  InstDev<S> d;
  d.g = in.g;
  d.ph = make_phys<S>(in.p.gamma, in.p.smallr, in.p.smallc, in.p.slope_type, in.p.niter_riemann,
                      in.p.iorder, in.p.scheme);
  d.dx = in.p.dx;
  d.cf = in.p.courant_factor;
  d.smallc = in.p.smallc;
  d.buf[0] = view_of<S>(b, in, 0);
  d.buf[1] = view_of<S>(b, in, 1);
  d.ctl = b->ctl + id;
  d.cour = b->cour + (size_t) id * 2 * kInstSlots;
  d.bl = in.p.boundary_left;
  d.br = in.p.boundary_right;
  d.bd = in.p.boundary_down;
  d.bu = in.p.boundary_up;
  d.nseg = (in.p.nx + kXSeg - 1) / kXSeg;
  d.yseg = inst_yseg(b, in.p.ny);
  return d;
}

static void tasks_of(const hydro_amd_batch* b, int id, int kind, std::vector<BTask>& out) {
  const Inst& in = b->inst[id];
  const int W = b->tune.waves;
  if (kind == 1) {
    const long nseg = (in.p.nx + kXSeg - 1) / kXSeg;
    const long tot = nseg * in.p.ny, per = (long) W * b->tune.xiter;
    for (long a = 0; a < tot; a += per) out.push_back(BTask{id, 1, (int) a, 0});
  } else {
    const int ncb = (in.p.nx + 63) / 64;
    const long rows = (long) W * inst_yseg(b, in.p.ny);
    const int nrg = (int) ((in.p.ny + rows - 1) / rows);
    for (int rg = 0; rg < nrg; rg++)
      for (int cb = 0; cb < ncb; cb++) out.push_back(BTask{id, 2, cb, rg});
  }
}

// Ghost fills of one instance in direction idim, kBoundaryChunk (line, var)
// pairs per task so a large instance spreads over many workgroups.
static void boundary_tasks_of(const hydro_amd_batch* b, int id, int idim,
                              std::vector<BTask>& out) {
  const Inst& in = b->inst[id];
  const int nl = idim == 1 ? in.p.ny : in.p.nx;
  for (int a = 0; a < 4 * nl; a += kBoundaryChunk) out.push_back(BTask{id, idim, a, 0});
}

// Reads the run controls of ids [0, maxid] into the pinned host copy.
static int read_ctl(hydro_amd_batch* b) {
  if (b->maxid < 0) return 0;
  BCHK(hipMemcpyAsync(b->hctl, b->ctl, sizeof(RunCtl) * (b->maxid + 1), hipMemcpyDeviceToHost,
                      b->st));
  BCHK(hipStreamSynchronize(b->st));
  return 0;
}

// Rebuilds the device instance array and the phase tables from the current
// live set and their step parities (the stream must be idle).
static int rebuild(hydro_amd_batch* b) {
  drop_graphs(b);
  b->live_ids.clear();
  b->maxid = -1;
  for (int k = 0; k < kMaxInst; k++)
    if (b->inst[k].live) {
      b->live_ids.push_back(k);
      b->maxid = k;
    }
  if (read_ctl(b)) return -1;
  // each sweep list: tasks of slope_type == 1 instances first (their own
  // kernel variant), then the others
  std::vector<BTask> L[6], G[6];
  for (int id : b->live_ids) {
    const bool even = (b->hctl[id].nstep % 2) == 0;
    std::vector<BTask>* S = b->inst[id].p.slope_type == 1.0 ? L : G;
    tasks_of(b, id, 1, even ? S[0] : S[2]);
    tasks_of(b, id, 2, even ? S[3] : S[1]);
    boundary_tasks_of(b, id, even ? 1 : 2, L[4]);
    boundary_tasks_of(b, id, even ? 2 : 1, L[5]);
  }
  int need = 0;
  for (int k = 0; k < 6; k++) {
    b->n1[k] = (int) L[k].size();
    L[k].insert(L[k].end(), G[k].begin(), G[k].end());
    b->off[k] = need;
    b->n[k] = (int) L[k].size();
    need += b->n[k];
  }
  b->nids = (int) b->live_ids.size();
  if (need > b->task_cap) {
    if (b->tasks) BCHK(hipFree(b->tasks));
    b->tasks = nullptr;
    b->task_cap = 0;
    const int cap = std::max(need, 1024) * 2;
    BCHK(hipMalloc(&b->tasks, sizeof(BTask) * cap));
    b->task_cap = cap;
  }
  std::vector<BTask> all;
  all.reserve(need);
  for (auto& l : L) all.insert(all.end(), l.begin(), l.end());
  if (!all.empty())
    BCHK(hipMemcpy(b->tasks, all.data(), sizeof(BTask) * all.size(), hipMemcpyHostToDevice));
  if (b->nids)
    BCHK(hipMemcpy(b->ids, b->live_ids.data(), sizeof(int) * b->nids, hipMemcpyHostToDevice));
  if (b->maxid >= 0) {
    if (dbl(b)) {
      auto* v = (InstDev<double>*) b->hidev;
      for (int id : b->live_ids) v[id] = inst_dev<double>(b, id);
      BCHK(hipMemcpy(b->idev, v, sizeof(v[0]) * (b->maxid + 1), hipMemcpyHostToDevice));
    } else {
      auto* v = (InstDev<float>*) b->hidev;
      for (int id : b->live_ids) v[id] = inst_dev<float>(b, id);
      BCHK(hipMemcpy(b->idev, v, sizeof(v[0]) * (b->maxid + 1), hipMemcpyHostToDevice));
    }
  }
  b->flip = false;
  b->dirty = false;
  return 0;
}

// The sweeps of task list k (kind 1: x, 2: y) in the batch's arithmetic form,
// reading buffer `phase`; q: the list's two task-queue counters.
template <typename S>
static void sweep_list(hydro_amd_batch* b, int kind, const InstDev<S>* I, int k, int phase,
                       int nblk, uint64_t* q) {
  if constexpr (sizeof(S) == 8) {
    if (b->prec == HYDRO_AMD_FP64_FAST) {
      launch_batch_sweep_fast(kind, I, b->tasks + b->off[k], b->n[k], b->n1[k], phase, b->tune,
                              nblk, q, b->st);
      return;
    }
  }
  launch_batch_sweep<S>(kind, I, b->tasks + b->off[k], b->n[k], b->n1[k], phase, b->tune, nblk,
                        q, b->st);
}

// One phase: its ghost fills, then its x tasks and its y tasks.
template <typename S>
static void enq_phase(hydro_amd_batch* b, int which, int phase) {
  const InstDev<S>* I = (const InstDev<S>*) b->idev;
  const int x = which == 0 ? 0 : 2, y = x + 1, bd = which == 0 ? 4 : 5;
  launch_batch_boundary<S>(I, b->tasks + b->off[bd], b->n[bd], phase, b->nblk_y, b->st);
  uint64_t* q = b->qctr + 4 * phase;  // x: q[0..1], y: q[2..3]
  sweep_list<S>(b, 1, I, x, phase, b->nblk_x, q);
  sweep_list<S>(b, 2, I, y, phase, b->nblk_y, q + 2);
}

template <typename S>
static void enq_step_t(hydro_amd_batch* b, bool flip) {
  const InstDev<S>* I = (const InstDev<S>*) b->idev;
  launch_batch_ctl<S>(I, b->ids, b->nids, true, b->qctr, b->st);
  enq_phase<S>(b, flip ? 1 : 0, 0);
  enq_phase<S>(b, flip ? 0 : 1, 1);
  launch_batch_ctl<S>(I, b->ids, b->nids, false, b->qctr, b->st);
}

static void enq_step(hydro_amd_batch* b, bool flip) {
  if (dbl(b)) enq_step_t<double>(b, flip);
  else enq_step_t<float>(b, flip);
}

static int pending_courant(hydro_amd_batch* b) {
  std::vector<int> need;
  for (int id : b->live_ids)
    if (b->inst[id].need_cour) {
      need.push_back(id);
      b->inst[id].need_cour = false;
    }
  if (need.empty()) return 0;
  BCHK(hipMemcpyAsync(b->cids, need.data(), sizeof(int) * need.size(), hipMemcpyHostToDevice,
                      b->st));
  if (dbl(b))
    launch_batch_courant<double>((const InstDev<double>*) b->idev, b->cids, (int) need.size(),
                                 b->st);
  else
    launch_batch_courant<float>((const InstDev<float>*) b->idev, b->cids, (int) need.size(),
                                b->st);
  // `need` must outlive the async copy
  BCHK(hipStreamSynchronize(b->st));
  return 0;
}

static int get_graph(hydro_amd_batch* b, int f) {
  if (b->gexec[f]) return 0;
  hipGraph_t graph;
  BCHK(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
  enq_step(b, f != 0);
  enq_step(b, f == 0);
  hipError_t e = hipStreamEndCapture(b->st, &graph);
  if (e != hipSuccess) return hydro_amd_set_error("graph capture: %s", hipGetErrorString(e));
  e = hipGraphInstantiate(&b->gexec[f], graph, nullptr, nullptr, 0);
  (void) hipGraphDestroy(graph);
  if (e != hipSuccess) {
    b->gexec[f] = nullptr;
    return hydro_amd_set_error("graph instantiate: %s", hipGetErrorString(e));
  }
  return 0;
}

extern "C" int hydro_amd_batch_run_async(hydro_amd_batch* b, int steps) {
  if (!b) return hydro_amd_set_error("null batch");
  if (steps < 0) return hydro_amd_set_error("steps < 0");
  if (b->dirty) {
    BCHK(hipStreamSynchronize(b->st));
    if (rebuild(b)) return -1;
  }
  if (b->nids == 0 || steps == 0) return 0;
  if (pending_courant(b)) return -1;
  int left = steps;
  if (left >= 2) {
    const int f = b->flip ? 1 : 0;
    if (get_graph(b, f)) return -1;
    for (; left >= 2; left -= 2) BCHK(hipGraphLaunch(b->gexec[f], b->st));
  }
  if (left == 1) {
    enq_step(b, b->flip);
    b->flip = !b->flip;
  }
  BCHK(hipGetLastError());
  return 0;
}

extern "C" int hydro_amd_batch_sync(hydro_amd_batch* b) {
  if (!b) return hydro_amd_set_error("null batch");
  BCHK(hipStreamSynchronize(b->st));
  return 0;
}

extern "C" int hydro_amd_batch_run(hydro_amd_batch* b, int steps) {
  if (hydro_amd_batch_run_async(b, steps)) return -1;
  return hydro_amd_batch_sync(b);
}

extern "C" int hydro_amd_batch_query(hydro_amd_batch* b, int* ids, double* t, int* nstep,
                                     double* dt, int* active, int cap) {
  if (!b) return hydro_amd_set_error("null batch");
  BCHK(hipStreamSynchronize(b->st));
  int top = -1;
  for (int k = 0; k < kMaxInst; k++)
    if (b->inst[k].live) top = k;
  b->maxid = top > b->maxid ? top : b->maxid;
  if (read_ctl(b)) return -1;
  int n = 0;
  for (int k = 0; k <= top; k++) {
    if (!b->inst[k].live) continue;
    if (n < cap) {
      const RunCtl& c = b->hctl[k];
      // a step opens only if t < tend && nstep < nstepmax (main.c:244);
      // -1: stopped at a dt that was not positive and finite
      const int more = c.fault ? -1 : ((c.t < c.tend) && (c.nstep < c.nstepmax));
      if (ids) ids[n] = k;
      if (t) t[n] = c.t;
      if (nstep) nstep[n] = c.nstep;
      if (dt) dt[n] = c.dt;
      if (active) active[n] = more;
    }
    n++;
  }
  return n;
}

extern "C" int hydro_amd_batch_set_tune(hydro_amd_batch* b, int waves, int xiter, int yseg) {
  if (!b) return hydro_amd_set_error("null batch");
  if (waves < 1 || waves > 4 || xiter < 1 || xiter > 256 || yseg < 1 || yseg > 1 << 16)
    return hydro_amd_set_error("tune out of range");
  BCHK(hipStreamSynchronize(b->st));
  b->tune = BatchTune{waves, xiter, yseg};
  size_grids(b);
  b->dirty = true;
  return 0;
}

// Times the phase sweeps of the current live set for one launch shape.  The
// run controls are saved, forced active for the timing (sweeps of phase 0
// read buffer 0 and write only buffer 1, the scratch), then restored: the
// state and the run are unchanged.
extern "C" int hydro_amd_batch_autotune(hydro_amd_batch* b, int reps, int* best) {
  if (!b) return hydro_amd_set_error("null batch");
  if (reps < 1) reps = 1;
  BCHK(hipStreamSynchronize(b->st));
  if (b->dirty && rebuild(b)) return -1;
  if (b->nids == 0) return hydro_amd_set_error("empty batch");
  const int nc = b->maxid + 1;
  std::vector<RunCtl> saved(nc);
  BCHK(hipMemcpy(saved.data(), b->ctl, sizeof(RunCtl) * nc, hipMemcpyDeviceToHost));
  std::vector<RunCtl> forced = saved;
  for (int id : b->live_ids) {
    forced[id].active = 1;
    forced[id].fault = 0;
    if (!(forced[id].dt > 0)) forced[id].dt = 1e-12;
  }
  BCHK(hipMemcpy(b->ctl, forced.data(), sizeof(RunCtl) * nc, hipMemcpyHostToDevice));
  const BatchTune keep = b->tune;
  BatchTune bt = keep;
  double bms = 1e30;
  hipEvent_t e0, e1;
  BCHK(hipEventCreate(&e0));
  BCHK(hipEventCreate(&e1));
  for (int w : {2, 4})
    for (int xi : {4, 8, 16, 32})
      for (int ys : {16, 32, 64, 128}) {
        if (hydro_amd_batch_set_tune(b, w, xi, ys) || rebuild(b)) break;
        for (int r = -1; r < reps; r++) {  // r = -1 warms up
          if (r == 0) (void) hipEventRecord(e0, b->st);
          // every x and y list at phase 0: reads buffer 0, writes buffer 1 only
          launch_zero_u64(b->qctr, kQueues, b->st);
          for (int k = 0; k < 4; k++) {
            const int kind = (k % 2 == 0) ? 1 : 2;
            const int nb = kind == 1 ? b->nblk_x : b->nblk_y;
            if (dbl(b))
              sweep_list<double>(b, kind, (const InstDev<double>*) b->idev, k, 0, nb,
                                 b->qctr + 2 * k);
            else
              sweep_list<float>(b, kind, (const InstDev<float>*) b->idev, k, 0, nb,
                                b->qctr + 2 * k);
          }
        }
        (void) hipEventRecord(e1, b->st);
        float ms = 1e30f;
        if (hipEventSynchronize(e1) == hipSuccess) (void) hipEventElapsedTime(&ms, e0, e1);
        if (ms < bms) {
          bms = ms;
          bt = b->tune;
        }
      }
  (void) hipEventDestroy(e0);
  (void) hipEventDestroy(e1);
  BCHK(hipMemcpy(b->ctl, saved.data(), sizeof(RunCtl) * nc, hipMemcpyHostToDevice));
  if (hydro_amd_batch_set_tune(b, bt.waves, bt.xiter, bt.yseg)) return -1;
  if (best) {
    best[0] = bt.waves;
    best[1] = bt.xiter;
    best[2] = bt.yseg;
  }
  return 0;
}
