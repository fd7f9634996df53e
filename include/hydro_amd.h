/*
 * hydro_amd.h -- handle-based C ABI of libhydro_amd beside the reference call
 * surface (hydro_surface.h).
 *
 * The reference has no API beyond its call surface; these entry points exist
 * for callers the surface cannot serve: several grids per process, a run of
 * many steps queued on the GPU without a host round trip per step, batched
 * grids advanced per launch, mixed precision and the serving loop.  Each one
 * states which reference function it generalises.
 *
 * Conventions: plain pointers and sizes only; int return 0 = success, < 0 =
 * error with a message from hydro_amd_last_error(); state arrays on the host
 * are the reference's [4][ny+4][nx+4] FP64 layout (slot 0 rho, 1 rho*u,
 * 2 rho*v, 3 E, x fastest, two ghost layers).  Calls on one handle must not
 * overlap; different handles may be used from different threads.
 */
#ifndef HYDRO_AMD_H
#define HYDRO_AMD_H

#include "hydro_surface.h"

#ifdef __cplusplus
extern "C" {
#endif

#define HYDRO_AMD_FP64 0   /* FP64 storage and arithmetic (bitwise the reference's) */
#define HYDRO_AMD_FP64_FAST 2 /* FP64 storage and arithmetic, not bit for bit: fused
                              multiply-adds, quotients and square roots from the
                              hardware estimates refined to about an ulp, and the
                              Newton iteration with one division per step.  Agrees
                              with FP64 to a per-field relative L1 of 1e-12 on
                              finite flows (smallr, smallc > 0); the fastest FP64
                              form. */
#define HYDRO_AMD_MIXED 1  /* FP32 storage / slopes / traces / fluxes; FP64 Newton,
                              update accumulation and CFL reduction.  The total
                              energy is stored in FP32, so a cell whose internal
                              energy is below ~1e-5 of its kinetic energy
                              (Mach ~ 1e2-1e3) loses its pressure to cancellation
                              and the run can go non-finite: use FP64 for such
                              flows. */

/* Run parameters: the physics/scheme fields of hydroparam_t
 * (HydroC99_2DMpi/Src/parametres.h:55-95) plus precision. */
typedef struct {
  int nx, ny;
  double dx, gamma, courant_factor, smallc, smallr, slope_type;
  int niter_riemann, iorder, scheme;
  int boundary_left, boundary_right, boundary_down, boundary_up;
  double tend;
  int nstepmax;
  int precision;
} hydro_amd_params;

typedef struct hydro_amd_grid hydro_amd_grid;

const char *hydro_amd_last_error(void);
/* Defaults of HydroC99_2DMpi/Src/parametres.c:64-109 with precision FP64. */
void hydro_amd_params_default(hydro_amd_params *p);
void hydro_amd_params_from_surface(const hydroparam_t *H, hydro_amd_params *p);

/* hydro_init + allocate_work_space + compute_deltat_init_mem for one grid
 * (HydroC99_2DMpi/Src/hydro_funcs.c:51-76, :165-293; compute_deltat.c:147):
 * GPU state, ping-pong buffer, CFL slots and run control.  State starts 0. */
hydro_amd_grid *hydro_amd_grid_create(const hydro_amd_params *p);
void hydro_amd_grid_destroy(hydro_amd_grid *g);
/* Initial conditions of hydro_funcs.c:79-146 (testcase 0, 1, 2) built on the
 * GPU; init_blast puts energy e at interior cell (bx, by) over the same
 * background (rho 1, u = v = 0, E = 1e-5).  Both reset t, nstep and dt. */
int hydro_amd_grid_init(hydro_amd_grid *g, int testcase);
int hydro_amd_grid_init_blast(hydro_amd_grid *g, int bx, int by, double e);
/* Whole-array copies in the reference layout (FP32 storage converts). */
int hydro_amd_grid_upload(hydro_amd_grid *g, const double *uold);
int hydro_amd_grid_download(hydro_amd_grid *g, double *uold);
/* Density image of the current state, HydroC99_2Domp/Src/Image.c:111-166 for
 * one rank: shrink = the smallest integer with nx/shrink and ny/shrink both
 * <= shrink_size; *width = nx/shrink, *height = ny/shrink; pixel (cx, cy) is
 * interior cell (cx*shrink, cy*shrink), rows cy = 0 first, 3 equal bytes
 * (uint8_t) fabs(255 * rho / max), max over every padded cell.  With rgb NULL
 * only the size is returned; otherwise cap must hold 3*w*h bytes. */
int hydro_amd_grid_image(hydro_amd_grid *g, int shrink_size, unsigned char *rgb, long long cap,
                         int *width, int *height);
/* hydro_godunov (hydro_godunov.c:63), make_boundary (make_boundary.c:146)
 * and compute_deltat (compute_deltat.c:166) on one grid. */
int hydro_amd_grid_godunov(hydro_amd_grid *g, int idim, double dt);
int hydro_amd_grid_make_boundary(hydro_amd_grid *g, int idim);
int hydro_amd_grid_compute_deltat(hydro_amd_grid *g, double *dt);
/* The time loop of HydroC99_2DMpi/Src/main.c:244-298 run on the GPU: up to
 * max_steps steps while t < tend and nstep < nstepmax, dt recomputed on even
 * steps and halved at step 0, sweep order alternating.  Steps are queued as a
 * replayed HIP graph, 256 at a time, and the host stops queueing when the
 * device loop has ended; dt_log (host, may be NULL, max_steps entries)
 * receives the dt of each step taken.  Returns the number of steps taken. */
int hydro_amd_grid_run(hydro_amd_grid *g, int max_steps, double *dt_log);
/* Enqueue min(max_steps, nstepmax - nstep) steps of the same loop without
 * waiting (steps past the end of the loop are no-ops); pair with
 * hydro_amd_grid_sync. */
int hydro_amd_grid_run_async(hydro_amd_grid *g, int max_steps);
int hydro_amd_grid_sync(hydro_amd_grid *g);
int hydro_amd_grid_state(hydro_amd_grid *g, double *t, int *nstep, double *dt);
int hydro_amd_grid_set_state(hydro_amd_grid *g, double t, int nstep, double dt);
int hydro_amd_grid_set_limits(hydro_amd_grid *g, double tend, int nstepmax);
/* Sweep launch shape (results do not depend on it): rows each y-sweep lane
 * marches, waves per y-sweep workgroup (1..4), waves per x-sweep workgroup
 * (1..4) and row pieces each x-sweep workgroup walks.  With 2 or 4 x waves
 * the workgroup shares one row piece of 64*xwaves cells through LDS; with 1
 * or 3 each wave takes its own 64-cell piece and exchanges neighbours by DPP
 * lane rotations.  A new grid starts with (64, 1, 4, 8) in FP64 and
 * (64, 4, 4, 16) in mixed precision. */
int hydro_amd_grid_set_tune(hydro_amd_grid *g, int yseg, int ywaves, int xwaves, int xiter);
/* Autotuner: times the candidate launch shapes of both sweeps on this grid
 * (reps launches each, HIP events; the sweeps read the state and write only
 * the scratch buffer, so the state is unchanged), keeps the fastest and
 * caches it per (nx, ny, precision) for the process.  best[4] (may be NULL)
 * receives yseg, ywaves, xwaves, xiter. */
int hydro_amd_grid_autotune(hydro_amd_grid *g, int reps, int *best);
/* Kernel timing with HIP events recorded on the grid's own stream around
 * every x-sweep, y-sweep, ghost-fill and CFL launch (graph replay is off
 * while enabled).  set_profile resets the sums; profile returns summed
 * milliseconds and launch counts in the order x sweep, y sweep, boundary,
 * CFL (arrays of 4). */
int hydro_amd_grid_set_profile(hydro_amd_grid *g, int on);
int hydro_amd_grid_profile(hydro_amd_grid *g, double *ms, long long *n);
/* Device bytes held by the grid. */
long long hydro_amd_grid_bytes(const hydro_amd_grid *g);

/* ---- batches: many independent grids advanced per launch ----------------
 * Each instance is a separate run of the reference time loop (own size,
 * gamma, blast, dt, t, nstep, tend/nstepmax), bit-identical to running it
 * alone with hydro_amd_grid_run.  One persistent workgroup grid per phase
 * walks a task table mixing the instances' x and y sweeps; two steps are
 * replayed as a HIP graph.  Instances are admitted against an HBM budget
 * (state buffers only) with the stream-ordered allocator, so admission and
 * retirement never synchronise the device.  All instances share one
 * precision. */
typedef struct hydro_amd_batch hydro_amd_batch;
hydro_amd_batch *hydro_amd_batch_create(int precision, long long budget_bytes);
void hydro_amd_batch_destroy(hydro_amd_batch *b);
/* Admit an instance (state zero, t = nstep = 0); returns its id >= 0, or -1
 * (e.g. over budget) with the reason in hydro_amd_last_error(). */
int hydro_amd_batch_add(hydro_amd_batch *b, const hydro_amd_params *p);
int hydro_amd_batch_remove(hydro_amd_batch *b, int id);
long long hydro_amd_batch_bytes_needed(const hydro_amd_batch *b, int nx, int ny);
long long hydro_amd_batch_bytes_used(const hydro_amd_batch *b);
/* hydro_amd_grid_init / init_blast / upload / download for one instance. */
int hydro_amd_batch_init(hydro_amd_batch *b, int id, int testcase);
int hydro_amd_batch_init_blast(hydro_amd_batch *b, int id, int bx, int by, double e);
int hydro_amd_batch_upload(hydro_amd_batch *b, int id, const double *uold);
int hydro_amd_batch_download(hydro_amd_batch *b, int id, double *uold);
/* Advance every live instance by up to `steps` steps of its own time loop
 * (instances that reach tend / nstepmax stop; the rest go on). */
int hydro_amd_batch_run(hydro_amd_batch *b, int steps);
int hydro_amd_batch_run_async(hydro_amd_batch *b, int steps);
int hydro_amd_batch_sync(hydro_amd_batch *b);
/* Waits, then reports the live instances: ids, t, nstep, dt and `active`:
 * 1 if another step would open (t < tend && nstep < nstepmax), 0 if the run
 * has ended, -1 if it was stopped at a dt that was not positive and finite
 * (the batch stops such a run; the reference would go on, so this differs
 * only for runs that have already broken down).  Fills up to cap entries;
 * returns the number of live instances. */
int hydro_amd_batch_query(hydro_amd_batch *b, int *ids, double *t, int *nstep, double *dt,
                          int *active, int cap);
/* Task shape (waves per workgroup 1..4, segments per wave in x tasks, rows per
 * wave in y tasks) and its autotuner (times the shapes on the live set; state
 * unchanged).  best[3] receives waves, xiter, yseg. */
int hydro_amd_batch_set_tune(hydro_amd_batch *b, int waves, int xiter, int yseg);
int hydro_amd_batch_autotune(hydro_amd_batch *b, int reps, int *best);

/* ---- serving: continuous batching of independent runs --------------------
 * Requests queue in arrival order.  Each pump admits queued requests while
 * they fit the HBM budget (FIFO), advances every running request by
 * chunk_steps steps in one batched device run -- requests admitted now join
 * those already on the GPU -- and retires the runs whose time loop has ended
 * (t >= tend or nstep >= nstepmax), copying results out when keep_results and
 * returning their memory.  A request that can never run (invalid parameters,
 * a precision other than the server's, a state larger than the whole budget)
 * is failed at submission, so it never holds up the queue.  A run whose dt
 * stops being positive and finite (or whose final state is not finite) is,
 * on a mixed-precision server, run again from its initial state in FP64 at
 * the head of the queue; otherwise, or if it breaks down in FP64 too, it is
 * failed alone.
 * Timestamps use hydro_amd_now() (CLOCK_MONOTONIC seconds). */
typedef struct hydro_amd_server hydro_amd_server;
double hydro_amd_now(void);
hydro_amd_server *hydro_amd_server_create(int precision, long long budget_bytes, int chunk_steps,
                                          int keep_results);
void hydro_amd_server_destroy(hydro_amd_server *s);
hydro_amd_batch *hydro_amd_server_batch(hydro_amd_server *s);
/* Queue a run of p (its tend / nstepmax end it) from the blast initial state
 * (energy e at interior cell bx, by); returns the request id. */
int hydro_amd_server_submit(hydro_amd_server *s, const hydro_amd_params *p, int bx, int by,
                            double e);
/* Queue a run of p from a whole initial state (host, the reference layout
 * [4][ny+4][nx+4] FP64; copied at submission, uploaded at admission). */
int hydro_amd_server_submit_state(hydro_amd_server *s, const hydro_amd_params *p,
                                  const double *uold);
/* One round; returns requests still queued or running (0 = idle), -1 error. */
int hydro_amd_server_pump(hydro_amd_server *s);
int hydro_amd_server_counts(hydro_amd_server *s, int *queued, int *running, int *done,
                            int *failed);
/* times[3] = submitted, admitted, finished; returns the state: 0 queued,
 * 1 running, 2 done, 3 failed (or -1 for an unknown id). */
int hydro_amd_server_info(hydro_amd_server *s, int req, double *times, double *t_sim, int *nstep);
int hydro_amd_server_result(hydro_amd_server *s, int req, double *uold);
int hydro_amd_server_release(hydro_amd_server *s, int req);
/* The precision a finished request's result was computed in: the server's,
 * or HYDRO_AMD_FP64 for a mixed-precision run that broke down (its dt not a
 * positive finite number, or a non-finite cell at its end) and was run again
 * from its initial state in FP64.  reruns: how many such re-runs there were. */
int hydro_amd_server_result_precision(hydro_amd_server *s, int req);
int hydro_amd_server_reruns(hydro_amd_server *s);
/* Pin `bytes` of host memory for kept results (once per server): a result
 * that fits is copied by DMA straight into it while the next chunk runs;
 * others are copied through staging buffers.  release() returns the space. */
int hydro_amd_server_reserve_results(hydro_amd_server *s, long long bytes);

/* Measurement support (no counterpart in the reference): runs a plain copy
 * kernel (k_copy_probe) over `bytes` bytes `reps` times and returns GB/s
 * counting bytes read + written; used as the achievable-bandwidth reference
 * and to calibrate HBM byte counters. */
double hydro_amd_copy_probe(long long bytes, int reps);

/* Measurement support (no counterpart in the reference): a copy in the
 * sweeps' access shape -- four arrays of `cells` doubles read, four written,
 * one cell per thread, non-temporal loads and stores (k_stream_probe) -- run
 * `reps` times; returns GB/s counting 64 bytes per cell, the bandwidth a
 * sweep could reach with no arithmetic.  Allocates 64 * cells bytes. */
double hydro_amd_stream_probe(long long cells, int reps);

/* Allocation-free compute_deltat.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same result as
 * hydro_amd_grid_compute_deltat (CFL max over the interior, then
 * dt = courant_factor*dx / max(cx, cy, smallc), compute_deltat.c:166-200) for
 * an FP64 state the caller keeps on the device in the reference layout
 * [4][ny+4][nx+4] (d_uold); dt is written to the device double d_dt.  The
 * kernels are queued on `stream` (a hipStream_t; NULL = the null stream) and
 * the call returns without waiting.  Scratch: a 16 KiB file-scope device
 * array per GPU (the CFL max-slots), cleared at the start of every call; no
 * hipMalloc or hipFree, any grid size up to (nx+4)*(ny+4) = 2^40.  Calls that
 * overlap on one GPU (other streams, other threads) race on that array. */
int hydro_amd_compute_deltat_noalloc(const hydro_amd_params *p, const double *d_uold,
                                     double *d_dt, void *stream);

/* Host mirror of the reference surface (hydro_surface.h): refresh Hv->uold
 * from the GPU, or push host edits of Hv->uold to the GPU. */
int hydro_amd_sync_host(hydrovar_t *Hv);
int hydro_amd_sync_device(hydrovar_t *Hv);
/* The grid behind a surface uold (NULL if Hv was not made by hydro_init). */
hydro_amd_grid *hydro_amd_surface_grid(hydrovar_t *Hv);
/* FP64 binary dump of the interior (the VTK dump of vtkfile.c is Float32):
 * "HYDROF64", int32 nx, ny, nvar, nstep, double t, double 0, then
 * double[nvar][ny][nx].  Refreshes the host mirror first. */
int hydro_amd_dump(const char *path, const hydroparam_t H, hydrovar_t *Hv);

/* Checkpoint / restart.  HydroC99_2DMpi has none; this keeps what the C++
 * variant's Continue.dump keeps for one domain (HydroCplusMPI/
 * Domain_Prot.cpp:98-181: the state, t, dt, the step count) in the dump's
 * layout, with dt in the second double.  An existing file is first renamed
 * to path.prev (Domain_Prot.cpp:268-282).  Ghost cells are not stored: every
 * sweep refills the ghosts it reads. */
int hydro_amd_checkpoint(const char *path, const hydroparam_t H, hydrovar_t *Hv, double dt);
/* Loads a checkpoint into a state built by hydro_init from the same deck
 * (nx, ny must match): the interior of Hv->uold, H->t, H->nstep and *dt (the
 * dt an odd step reuses); refreshes the GPU copy.  0, or -1 with
 * hydro_amd_last_error() set. */
int hydro_amd_restore(const char *path, hydroparam_t *H, hydrovar_t *Hv, double *dt);

/* Image cadence keys of the last deck process_args() read (HydroC99_2Domp/
 * Src/parametres.c:69-70, 165-166, 202-203: `nimage` every n steps, or
 * `dtimage` in simulated time; both 0 by default).  hydroparam_t of
 * HydroC99_2DMpi has no fields for them, so they are kept beside it. */
void hydro_amd_image_cadence(int *nimage, double *dtimage);

#ifdef __cplusplus
}
#endif
#endif
