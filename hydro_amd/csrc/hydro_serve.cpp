// Serving loop over a batch -- the C ABI hydro_amd_server_* of
// include/hydro_amd.h.
//
// Continuous batching: requests (independent hydro runs) queue in arrival
// order; each pump admits queued requests while they fit the HBM budget,
// advances every running request by `chunk` steps in one batched device run
// (new arrivals join the runs already on the GPU), then retires the ones whose
// time loop has ended (t >= tend or nstep >= nstepmax): results are copied out
// and their memory returns to the budget.  A request whose dt stops being a
// positive finite number is failed and retired on its own; the others go on.
#include <hip/hip_runtime.h>

#include <math.h>
#include <time.h>

#include <deque>
#include <map>
#include <vector>

#include "../../include/hydro_amd.h"

int hydro_amd_set_error(const char* fmt, ...);  // hydro_capi.cpp
int hydro_amd_batch_download_raw(hydro_amd_batch* b, int id, void* dst, int* elem_size);

namespace {
enum { QUEUED = 0, RUNNING = 1, DONE = 2, FAILED = 3 };

struct Request {
  hydro_amd_params p;
  int bx, by;
  double e;
  double t_submit = 0, t_admit = 0, t_done = 0;
  int state = QUEUED;
  int inst = -1;
  double t_sim = 0, dt = 0;
  int nstep = 0;
  std::vector<unsigned char> result;  // state as stored (8- or 4-byte elements)
  int esz = 0;
};

double now_s() {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}
}  // namespace

struct hydro_amd_server {
  hydro_amd_batch* b = nullptr;
  int chunk = 20;
  int keep = 0;
  std::vector<Request> req;
  std::deque<int> queue;
  std::map<int, int> running;  // instance id -> request id
  int ndone = 0, nfailed = 0;
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
  s->b = b;
  s->chunk = chunk_steps;
  s->keep = keep_results;
  return s;
}

extern "C" void hydro_amd_server_destroy(hydro_amd_server* s) {
  if (!s) return;
  hydro_amd_batch_destroy(s->b);
  delete s;
}

extern "C" hydro_amd_batch* hydro_amd_server_batch(hydro_amd_server* s) { return s ? s->b : nullptr; }

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
  s->req.push_back(r);
  const int id = (int) s->req.size() - 1;
  s->queue.push_back(id);
  return id;
}

static int retire(hydro_amd_server* s, int inst, int rid, int state) {
  Request& r = s->req[rid];
  if (s->keep && state == DONE) {
    const size_t n = (size_t) 4 * (r.p.nx + 4) * (r.p.ny + 4);
    r.result.resize(n * (r.p.precision == HYDRO_AMD_FP64 ? 8 : 4));
    if (hydro_amd_batch_download_raw(s->b, inst, r.result.data(), &r.esz)) return -1;
  }
  if (hydro_amd_batch_remove(s->b, inst)) return -1;
  r.state = state;
  r.t_done = now_s();
  r.inst = -1;
  if (state == DONE) s->ndone++;
  else s->nfailed++;
  return 0;
}

// One scheduling round; returns the number of requests still queued or
// running (0 when the server is idle), or -1 on error.
extern "C" int hydro_amd_server_pump(hydro_amd_server* s) {
  if (!s) return hydro_amd_set_error("null server");
  // admit in arrival order while the head fits
  while (!s->queue.empty()) {
    const int rid = s->queue.front();
    Request& r = s->req[rid];
    const int inst = hydro_amd_batch_add(s->b, &r.p);
    if (inst < 0) {
      if (s->running.empty()) {  // can never fit: fail it instead of stalling
        s->queue.pop_front();
        r.state = FAILED;
        r.t_done = now_s();
        s->nfailed++;
        continue;
      }
      break;
    }
    if (hydro_amd_batch_init_blast(s->b, inst, r.bx, r.by, r.e)) return -1;
    s->queue.pop_front();
    r.state = RUNNING;
    r.inst = inst;
    r.t_admit = now_s();
    s->running[inst] = rid;
  }
  if (s->running.empty()) return (int) s->queue.size();
  if (hydro_amd_batch_run(s->b, s->chunk)) return -1;
  const int n = (int) s->running.size();
  std::vector<int> ids(n), nstep(n), more(n);
  std::vector<double> t(n), dt(n);
  const int got = hydro_amd_batch_query(s->b, ids.data(), t.data(), nstep.data(), dt.data(),
                                        more.data(), n);
  if (got < 0) return -1;
  for (int k = 0; k < got && k < n; k++) {
    auto it = s->running.find(ids[k]);
    if (it == s->running.end()) continue;
    const int rid = it->second;
    Request& r = s->req[rid];
    r.t_sim = t[k];
    r.nstep = nstep[k];
    r.dt = dt[k];
    const bool sane = more[k] >= 0 && isfinite(t[k]);
    if (!sane || more[k] == 0) {
      if (retire(s, ids[k], rid, sane ? DONE : FAILED)) return -1;
      s->running.erase(it);
    }
  }
  return (int) (s->queue.size() + s->running.size());
}

extern "C" int hydro_amd_server_counts(hydro_amd_server* s, int* queued, int* running, int* done,
                                       int* failed) {
  if (!s) return hydro_amd_set_error("null server");
  if (queued) *queued = (int) s->queue.size();
  if (running) *running = (int) s->running.size();
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
  Request& r = s->req[rid];
  if (r.state != DONE || r.result.empty()) return hydro_amd_set_error("no result for %d", rid);
  const size_t n = r.result.size() / r.esz;
  if (r.esz == 8) {
    const double* d = (const double*) r.result.data();
    for (size_t k = 0; k < n; k++) uold[k] = d[k];
  } else {
    const float* f = (const float*) r.result.data();
    for (size_t k = 0; k < n; k++) uold[k] = f[k];
  }
  return 0;
}

// Drops a kept result (frees host memory).
extern "C" int hydro_amd_server_release(hydro_amd_server* s, int rid) {
  if (!s || rid < 0 || rid >= (int) s->req.size()) return hydro_amd_set_error("no request %d", rid);
  std::vector<unsigned char>().swap(s->req[rid].result);
  return 0;
}
