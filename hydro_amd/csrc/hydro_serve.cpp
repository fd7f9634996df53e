// Serving loop over a batch -- the C ABI hydro_amd_server_* of
// include/hydro_amd.h.
//
// Continuous batching: requests (independent hydro runs) queue in arrival
// order; each pump admits queued requests while they fit the HBM budget,
// queues `chunk` steps of every running request as one batched device run
// (new arrivals join the runs already on the GPU), and while that runs copies
// out the results of the requests whose time loop ended (t >= tend or nstep
// >= nstepmax) in the previous pump -- on a second stream, so the copies
// overlap the kernels -- and returns their memory to the budget.
//
// Requests that can never run (parameters a batch rejects, a precision other
// than the server's, a state larger than the whole budget) fail at submission,
// so they never hold up the queue.  A request whose run stops being finite
// (its dt not a positive finite number, or a non-finite cell in its final
// state) is, on a mixed-precision server, run again from its initial state in
// FP64 -- on a second batch created on first use, within the same budget, at
// the head of the queue -- since the FP32 total energy of the mixed form can
// lose the pressure of hypersonic cells; if it breaks down in FP64 too (or on
// an FP64 server) it fails alone and the others go on.
#include <hip/hip_runtime.h>

#include <math.h>
#include <time.h>

#include <deque>
#include <map>
#include <memory>
#include <tuple>
#include <utility>
#include <vector>

#include "../../include/hydro_amd.h"

extern "C" int hydro_amd_set_error(const char* fmt, ...);  // hydro_capi.cpp
int hydro_amd_batch_download_raw(hydro_amd_batch* b, int id, void* dst, int* elem_size,
                                 hipStream_t st);  // hydro_batch.cpp
int hydro_amd_batch_raw_view(hydro_amd_batch* b, int id, const void** src, long* pitch,
                             long* width, long* rows, int* esz);
int hydro_amd_batch_finite(hydro_amd_batch* b, const int* ids, int n, int* finite);

namespace {
enum { QUEUED = 0, RUNNING = 1, DONE = 2, FAILED = 3 };

// Pinned host memory for results: first-fit blocks (256-byte aligned) with
// coalescing frees.  A result placed here is copied by DMA straight from the
// GPU, asynchronously, beside the kernels.
struct Arena {
  unsigned char* base = nullptr;
  size_t size = 0;
  std::map<size_t, size_t> free_;  // offset -> length
  bool alloc(size_t n, size_t* off) {
    n = (n + 255) & ~(size_t) 255;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < n) continue;
      *off = it->first;
      const size_t rest = it->second - n;
      const size_t at = it->first + n;
      free_.erase(it);
      if (rest) free_[at] = rest;
      return true;
    }
    return false;
  }
  void release(size_t off, size_t n) {
    n = (n + 255) & ~(size_t) 255;
    auto it = free_.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
};

struct Request {
  hydro_amd_params p;
  int bx, by;
  double e;
  double t_submit = 0, t_admit = 0, t_done = 0;
  int state = QUEUED;
  int inst = -1;
  int lane = 0;  // batch: 0 the server's, 1 the FP64 re-run batch
  std::vector<double> u0;  // initial state when given whole (else the blast)
  double t_sim = 0, dt = 0;
  int nstep = 0;
  std::unique_ptr<unsigned char[]> result;  // state as stored (8- or 4-byte elements)
  size_t result_bytes = 0;
  size_t arena_off = ~(size_t) 0;           // or here, in the server's pinned arena
  int esz = 0;
  const unsigned char* data(const Arena& a) const {
    return arena_off != ~(size_t) 0 ? a.base + arena_off : result.get();
  }
};

double now_s() {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}
}  // namespace

struct hydro_amd_server {
  // b[0]: the server's precision; b[1]: FP64 re-runs of mixed-precision runs
  // that broke down (created on first use)
  hydro_amd_batch* b[2] = {nullptr, nullptr};
  int prec = HYDRO_AMD_MIXED;
  long long budget = 0;           // over both batches
  hipStream_t copy_st = nullptr;  // result copies, beside the batches' kernels
  int chunk = 20;
  int keep = 0;
  std::vector<Request> req;
  std::deque<int> queue;
  std::map<int, int> running[2];   // per batch: instance id -> request id
  std::map<int, int> finished[2];  // ended on the device, result not yet copied out
  std::vector<std::tuple<int, int, int>> copying;  // (batch, instance, request): DMA queued
  Arena arena;
  int ndone = 0, nfailed = 0, nrerun = 0;
};

extern "C" double hydro_amd_now(void) { return now_s(); }

extern "C" hydro_amd_server* hydro_amd_server_create(int precision, long long budget_bytes,
                                                     int chunk_steps, int keep_results) {
  if (chunk_steps < 1) {
    hydro_amd_set_error("chunk_steps must be >= 1");
    return nullptr;
  }
  hydro_amd_batch* b = hydro_amd_batch_create(precision, budget_bytes);
  if (!b) return nullptr;
  hydro_amd_server* s = new hydro_amd_server();
  s->b[0] = b;
  s->prec = precision;
  s->budget = budget_bytes > 0 ? budget_bytes : (1LL << 62);
  if (hipStreamCreateWithFlags(&s->copy_st, hipStreamNonBlocking) != hipSuccess) {
    hydro_amd_set_error("stream creation failed");
    hydro_amd_batch_destroy(b);
    delete s;
    return nullptr;
  }
  s->chunk = chunk_steps;
  s->keep = keep_results;
  return s;
}

extern "C" void hydro_amd_server_destroy(hydro_amd_server* s) {
  if (!s) return;
  if (s->copy_st) (void) hipStreamSynchronize(s->copy_st);
  hydro_amd_batch_destroy(s->b[0]);
  hydro_amd_batch_destroy(s->b[1]);
  if (s->copy_st) (void) hipStreamDestroy(s->copy_st);
  if (s->arena.base) (void) hipHostFree(s->arena.base);
  delete s;
}

extern "C" int hydro_amd_server_reserve_results(hydro_amd_server* s, long long bytes) {
  if (!s || bytes < 0) return hydro_amd_set_error("bad argument");
  if (s->arena.base) return hydro_amd_set_error("result memory already reserved");
  if (bytes == 0) return 0;
  void* p = nullptr;
  if (hipHostMalloc(&p, (size_t) bytes, hipHostMallocDefault) != hipSuccess)
    return hydro_amd_set_error("pinned allocation of %lld bytes failed", bytes);
  s->arena.base = (unsigned char*) p;
  s->arena.size = (size_t) bytes & ~(size_t) 255;
  s->arena.free_[0] = s->arena.size;
  return 0;
}

extern "C" hydro_amd_batch* hydro_amd_server_batch(hydro_amd_server* s) {
  return s ? s->b[0] : nullptr;
}

static long long used_bytes(const hydro_amd_server* s) {
  long long u = hydro_amd_batch_bytes_used(s->b[0]);
  if (s->b[1]) u += hydro_amd_batch_bytes_used(s->b[1]);
  return u;
}

// Device bytes a run of p holds in the given precision (state buffers; the
// same for every batch of that precision).
static long long request_bytes(const hydro_amd_server* s, const hydro_amd_params& p, int prec) {
  const long long n = hydro_amd_batch_bytes_needed(s->b[0], p.nx, p.ny);
  if (n < 0) return -1;
  const int have = s->prec == HYDRO_AMD_MIXED ? 4 : 8, want = prec == HYDRO_AMD_MIXED ? 4 : 8;
  return n / have * want;
}

// Why a request can never run (nullptr if it can): what hydro_amd_batch_add
// rejects whatever the load, or a state larger than the whole budget.
static const char* never_runs(const hydro_amd_server* s, const hydro_amd_params& p) {
  if (p.precision != s->prec) return "precision differs from the server's";
  if (!(p.dx > 0) || !(p.gamma > 1) || p.scheme < 1 || p.scheme > 3 || p.niter_riemann < 0)
    return "invalid run parameters";
  if ((long) ((p.nx + 59) / 60) * p.ny > (1L << 30)) return "grid too large";
  if (request_bytes(s, p, s->prec) > s->budget) return "state larger than the whole budget";
  return nullptr;
}

static void fail_request(hydro_amd_server* s, Request& r) {
  r.state = FAILED;
  r.t_done = now_s();
  r.inst = -1;
  s->nfailed++;
}

extern "C" int hydro_amd_server_submit(hydro_amd_server* s, const hydro_amd_params* p, int bx,
                                       int by, double e) {
  if (!s || !p) return hydro_amd_set_error("null argument");
  if (p->nx < 1 || p->ny < 1 || bx < 0 || by < 0 || bx >= p->nx || by >= p->ny)
    return hydro_amd_set_error("invalid request geometry");
  Request r;
  r.p = *p;
  r.bx = bx;
  r.by = by;
  r.e = e;
  r.t_submit = now_s();
  s->req.push_back(std::move(r));
  const int id = (int) s->req.size() - 1;
  if (never_runs(s, *p)) fail_request(s, s->req[id]);  // at once: it would stall the queue
  else s->queue.push_back(id);
  return id;
}

extern "C" int hydro_amd_server_submit_state(hydro_amd_server* s, const hydro_amd_params* p,
                                             const double* uold) {
  if (!s || !p || !uold) return hydro_amd_set_error("null argument");
  if (p->nx < 1 || p->ny < 1) return hydro_amd_set_error("invalid request geometry");
  const int id = hydro_amd_server_submit(s, p, 0, 0, 0.0);
  if (id < 0) return -1;
  Request& r = s->req[id];
  if (r.state == QUEUED) r.u0.assign(uold, uold + (size_t) 4 * (p->nx + 4) * (p->ny + 4));
  return id;
}

// Queues the DMA of a finished request's result into the pinned arena, if it
// has room; returns 1 if queued (the request completes when the copy stream
// has drained), 0 if not.  The instance is inactive on the device (its run
// control ended), so kernels queued meanwhile on the batch's stream leave its
// state alone.
static int start_copy(hydro_amd_server* s, int lane, int inst, int rid) {
  Request& r = s->req[rid];
  const void* src;
  long pitch, width, rows;
  int esz;
  if (!s->arena.base ||
      hydro_amd_batch_raw_view(s->b[lane], inst, &src, &pitch, &width, &rows, &esz))
    return 0;
  const size_t bytes = (size_t) width * rows * esz;
  size_t off;
  if (!s->arena.alloc(bytes, &off)) return 0;
  if (hipMemcpy2DAsync(s->arena.base + off, (size_t) width * esz, src, (size_t) pitch * esz,
                       (size_t) width * esz, rows, hipMemcpyDeviceToHost, s->copy_st) !=
      hipSuccess) {
    s->arena.release(off, bytes);
    return hydro_amd_set_error("result copy failed");
  }
  r.arena_off = off;
  r.result_bytes = bytes;
  r.esz = esz;
  return 1;
}

// Completes a request: copies its result out (staged, on the copy stream)
// unless already in the arena, and frees its instance in the batch stream's
// order.
static int retire(hydro_amd_server* s, int lane, int inst, int rid, int state) {
  Request& r = s->req[rid];
  if (s->keep && state == DONE && r.arena_off == ~(size_t) 0) {
    const size_t n = (size_t) 4 * (r.p.nx + 4) * (r.p.ny + 4);
    r.result_bytes = n * (lane == 0 && s->prec == HYDRO_AMD_MIXED ? 4 : 8);
    r.result.reset(new unsigned char[r.result_bytes]);
    if (hydro_amd_batch_download_raw(s->b[lane], inst, r.result.get(), &r.esz, s->copy_st))
      return -1;
  }
  if (hydro_amd_batch_remove(s->b[lane], inst)) return -1;
  std::vector<double>().swap(r.u0);
  r.state = state;
  r.t_done = now_s();
  r.inst = -1;
  if (state == DONE) s->ndone++;
  else s->nfailed++;
  return 0;
}

// A mixed-precision run that broke down: drop its instance and queue it again
// at the head, from its initial state, for the FP64 batch (if its FP64 state
// fits the budget; otherwise it fails).
static int rerun_fp64(hydro_amd_server* s, int inst, int rid) {
  Request& r = s->req[rid];
  if (hydro_amd_batch_remove(s->b[0], inst)) return -1;
  r.inst = -1;
  hydro_amd_params p = r.p;
  p.precision = HYDRO_AMD_FP64;
  if (request_bytes(s, p, HYDRO_AMD_FP64) > s->budget) {
    fail_request(s, r);
    return 0;
  }
  r.lane = 1;
  r.state = QUEUED;
  r.t_sim = 0;
  r.nstep = 0;
  s->nrerun++;
  s->queue.push_front(rid);
  return 0;
}

static bool busy(const hydro_amd_server* s) {
  for (int l = 0; l < 2; l++)
    if (!s->running[l].empty() || !s->finished[l].empty()) return true;
  return !s->copying.empty();
}

// Admits queued requests in arrival order (re-runs first) while the head
// fits the budget.
static int admit(hydro_amd_server* s) {
  while (!s->queue.empty()) {
    const int rid = s->queue.front();
    Request& r = s->req[rid];
    hydro_amd_params p = r.p;
    if (r.lane == 1) {
      p.precision = HYDRO_AMD_FP64;
      if (!s->b[1] && !(s->b[1] = hydro_amd_batch_create(HYDRO_AMD_FP64, s->budget))) return -1;
    }
    const long long need = request_bytes(s, p, p.precision);
    if (used_bytes(s) + need > s->budget) break;  // fits once running requests retire
    const int inst = hydro_amd_batch_add(s->b[r.lane], &p);
    if (inst < 0) {
      if (!busy(s)) {  // nothing else holds memory: it will not fit later either
        s->queue.pop_front();
        fail_request(s, r);
        continue;
      }
      break;
    }
    if (r.u0.empty() ? hydro_amd_batch_init_blast(s->b[r.lane], inst, r.bx, r.by, r.e)
                     : hydro_amd_batch_upload(s->b[r.lane], inst, r.u0.data()))
      return -1;
    s->queue.pop_front();
    r.state = RUNNING;
    r.inst = inst;
    r.t_admit = now_s();
    s->running[r.lane][inst] = rid;
  }
  return 0;
}

// Reads the run state of a batch's running requests: ended runs move to
// `finished` (a mixed run whose final state is not finite is run again in
// FP64), runs that broke down are re-run (mixed) or fail.
static int poll(hydro_amd_server* s, int lane) {
  auto& run = s->running[lane];
  if (run.empty()) return 0;
  const int n = (int) run.size();
  std::vector<int> ids(n), nstep(n), more(n);
  std::vector<double> t(n), dt(n);
  const int got = hydro_amd_batch_query(s->b[lane], ids.data(), t.data(), nstep.data(),
                                        dt.data(), more.data(), n);
  if (got < 0) return -1;
  const bool can_rerun = lane == 0 && s->prec == HYDRO_AMD_MIXED;
  std::vector<int> ended;
  for (int k = 0; k < got && k < n; k++) {
    auto it = run.find(ids[k]);
    if (it == run.end()) continue;
    const int rid = it->second;
    Request& r = s->req[rid];
    r.t_sim = t[k];
    r.nstep = nstep[k];
    r.dt = dt[k];
    const bool sane = more[k] >= 0 && isfinite(t[k]);
    if (!sane) {
      run.erase(it);
      if (can_rerun) {
        if (rerun_fp64(s, ids[k], rid)) return -1;
      } else if (retire(s, lane, ids[k], rid, FAILED)) {
        return -1;
      }
    } else if (more[k] == 0) {
      ended.push_back(ids[k]);
    }
  }
  if (ended.empty()) return 0;
  // a breakdown in the last steps leaves a non-finite state without a bad dt
  std::vector<int> fin(ended.size(), 1);
  if (can_rerun && hydro_amd_batch_finite(s->b[lane], ended.data(), (int) ended.size(),
                                          fin.data()))
    return -1;
  for (size_t k = 0; k < ended.size(); k++) {
    const int rid = run[ended[k]];
    run.erase(ended[k]);
    if (fin[k]) {
      s->finished[lane][ended[k]] = rid;
    } else if (rerun_fp64(s, ended[k], rid)) {
      return -1;
    }
  }
  return 0;
}

// One scheduling round; returns the number of requests still queued,
// running or waiting for their result to be copied out (0 when the server is
// idle), or -1 on error.
extern "C" int hydro_amd_server_pump(hydro_amd_server* s) {
  if (!s) return hydro_amd_set_error("null server");
  if (admit(s)) return -1;
  // queue the next chunk, then copy out the runs that ended in the last one
  // while it executes
  for (int l = 0; l < 2; l++)
    if (!s->running[l].empty() && hydro_amd_batch_run_async(s->b[l], s->chunk)) return -1;
  bool freed = false;
  for (int l = 0; l < 2; l++) {
    for (const auto& f : s->finished[l]) {
      const int q = s->keep ? start_copy(s, l, f.first, f.second) : 0;
      if (q < 0) return -1;
      if (q) {
        s->copying.emplace_back(l, f.first, f.second);
      } else {
        if (retire(s, l, f.first, f.second, DONE)) return -1;
        freed = true;
      }
    }
    s->finished[l].clear();
  }
  // requests that fit in the memory just returned join from the next chunk
  if (freed && admit(s)) return -1;
  for (int l = 0; l < 2; l++)
    if (!s->running[l].empty() && hydro_amd_batch_sync(s->b[l])) return -1;
  if (!s->copying.empty()) {
    if (hipStreamSynchronize(s->copy_st) != hipSuccess)
      return hydro_amd_set_error("result copy failed");
    for (const auto& c : s->copying)
      if (retire(s, std::get<0>(c), std::get<1>(c), std::get<2>(c), DONE)) return -1;
    s->copying.clear();
  }
  for (int l = 0; l < 2; l++)
    if (poll(s, l)) return -1;
  size_t left = s->queue.size();
  for (int l = 0; l < 2; l++) left += s->running[l].size() + s->finished[l].size();
  return (int) left;
}

extern "C" int hydro_amd_server_counts(hydro_amd_server* s, int* queued, int* running, int* done,
                                       int* failed) {
  if (!s) return hydro_amd_set_error("null server");
  if (queued) *queued = (int) s->queue.size();
  if (running)
    *running = (int) (s->running[0].size() + s->finished[0].size() + s->running[1].size() +
                      s->finished[1].size());
  if (done) *done = s->ndone;
  if (failed) *failed = s->nfailed;
  return 0;
}

// times[3] = submit, admit, done (hydro_amd_now clock); returns the state
// (0 queued, 1 running, 2 done, 3 failed).
extern "C" int hydro_amd_server_info(hydro_amd_server* s, int rid, double* times, double* t_sim,
                                     int* nstep) {
  if (!s || rid < 0 || rid >= (int) s->req.size()) return hydro_amd_set_error("no request %d", rid);
  const Request& r = s->req[rid];
  if (times) {
    times[0] = r.t_submit;
    times[1] = r.t_admit;
    times[2] = r.t_done;
  }
  if (t_sim) *t_sim = r.t_sim;
  if (nstep) *nstep = r.nstep;
  return r.state;
}

extern "C" int hydro_amd_server_result(hydro_amd_server* s, int rid, double* uold) {
  if (!s || rid < 0 || rid >= (int) s->req.size()) return hydro_amd_set_error("no request %d", rid);
  if (!uold) return hydro_amd_set_error("null argument");
  Request& r = s->req[rid];
  const unsigned char* src = r.data(s->arena);
  if (r.state != DONE || !src) return hydro_amd_set_error("no result for %d", rid);
  const size_t n = r.result_bytes / r.esz;
  if (r.esz == 8) {
    const double* d = (const double*) src;
    for (size_t k = 0; k < n; k++) uold[k] = d[k];
  } else {
    const float* f = (const float*) src;
    for (size_t k = 0; k < n; k++) uold[k] = f[k];
  }
  return 0;
}

// Drops a kept result (frees host memory).
extern "C" int hydro_amd_server_release(hydro_amd_server* s, int rid) {
  if (!s || rid < 0 || rid >= (int) s->req.size()) return hydro_amd_set_error("no request %d", rid);
  Request& r = s->req[rid];
  if (r.arena_off != ~(size_t) 0) s->arena.release(r.arena_off, r.result_bytes);
  r.arena_off = ~(size_t) 0;
  r.result.reset();
  r.result_bytes = 0;
  return 0;
}

// The precision a request's result was computed in: the server's, or FP64
// for a mixed-precision run that was run again after breaking down.
extern "C" int hydro_amd_server_result_precision(hydro_amd_server* s, int rid) {
  if (!s || rid < 0 || rid >= (int) s->req.size()) return hydro_amd_set_error("no request %d", rid);
  return s->req[rid].lane == 1 ? HYDRO_AMD_FP64 : s->prec;
}

extern "C" int hydro_amd_server_reruns(hydro_amd_server* s) {
  return s ? s->nrerun : hydro_amd_set_error("null server");
}
