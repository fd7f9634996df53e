/*
 * Test-only harness around the reference HydroC99_2DMpi sources.
 *
 * Built by oracle/Makefile into oracle/_ref/libhydroref.so together with the
 * reference's own Src/*.c (read in place, never copied).  Test infrastructure:
 * tests and the golden-vector script call it through ctypes to (1) run the
 * reference's stage functions on chosen slabs and (2) drive the reference's
 * time loop with full-precision access to uold and the dt sequence, which the
 * reference executable only prints to ~6 digits.
 *
 * ref_steps() restates the loop of HydroC99_2DMpi/Src/main.c:244-298 (dt on
 * even steps, halved at step 0, hydro_godunov(1|2) alternating, t += dt); the
 * unchanged reference executable (oracle/_ref/hydro_d) is used by the tests to
 * cross-check it against the printed step lines.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "parametres.h"
#include "hydro_funcs.h"
#include "compute_deltat.h"
#include "hydro_godunov.h"
#include "make_boundary.h"
#include "vtkfile.h"

/* main.c:69 defines this global; every stage adds its timings to it. */
double functim[TIM_END];

static hydroparam_t H;
static hydrovar_t Hv;
static hydrovarwork_t Hvw_deltat, Hvw_godunov;
static hydrowork_t Hw_deltat, Hw_godunov;
static int opened = 0;

/* dims: nx, ny, nxt, nyt, nxyt, nxystep, nstepmax, testcase */
int ref_open(const char *nml, int *dims) {
  char *argv[3];
  if (opened) return -1;
  argv[0] = "hydro";
  argv[1] = "-i";
  argv[2] = (char *) nml;
  memset(functim, 0, sizeof(functim));
  process_args(3, argv, &H);
  hydro_init(&H, &Hv);
  allocate_work_space(H.nxyt, H, &Hw_godunov, &Hvw_godunov);
  compute_deltat_init_mem(H, &Hw_deltat, &Hvw_deltat);
  opened = 1;
  if (dims) {
    dims[0] = H.nx; dims[1] = H.ny; dims[2] = H.nxt; dims[3] = H.nyt;
    dims[4] = H.nxyt; dims[5] = H.nxystep; dims[6] = H.nstepmax; dims[7] = H.testCase;
  }
  return 0;
}

void ref_close(void) {
  if (!opened) return;
  deallocate_work_space(H.nxyt, H, &Hw_godunov, &Hvw_godunov);
  compute_deltat_clean_mem(H, &Hw_deltat, &Hvw_deltat);
  hydro_finish(H, &Hv);
  opened = 0;
}

/* reals: t, tend, dx, gamma, courant_factor, smallc, smallr, slope_type */
void ref_get_reals(double *r) {
  r[0] = H.t; r[1] = H.tend; r[2] = H.dx; r[3] = H.gamma;
  r[4] = H.courant_factor; r[5] = H.smallc; r[6] = H.smallr; r[7] = H.slope_type;
}

/* ints: nstep, nstepmax, niter_riemann, iorder, scheme, bl, br, bd, bu */
void ref_get_ints(int *v) {
  v[0] = H.nstep; v[1] = H.nstepmax; v[2] = H.niter_riemann; v[3] = H.iorder;
  v[4] = H.scheme; v[5] = H.boundary_left; v[6] = H.boundary_right;
  v[7] = H.boundary_down; v[8] = H.boundary_up;
}

/* gamma is hard-wired in the reference's default_values(); let tests vary it. */
void ref_set_gamma(double g) { H.gamma = g; }

/* uold crosses as FP64 whatever real_t is: the build with -DPREC_SP (the
 * reference's all-FP32 mode, parametres.h:41-45) converts on the way. */
long ref_uold_len(void) { return (long) H.nvar * H.nxt * H.nyt; }
void ref_get_uold(double *dst) {
  for (long k = 0; k < ref_uold_len(); k++) dst[k] = Hv.uold[k];
}
void ref_set_uold(const double *src) {
  for (long k = 0; k < ref_uold_len(); k++) Hv.uold[k] = (real_t) src[k];
}

void ref_compute_deltat(double *dt) {
  real_t d = 0;
  compute_deltat(&d, H, &Hw_deltat, &Hv, &Hvw_deltat);
  *dt = d;
}
void ref_godunov(int idim, double dt) {
  hydro_godunov(idim, (real_t) dt, H, &Hv, &Hw_godunov, &Hvw_godunov);
}
void ref_make_boundary(int idim) { make_boundary(idim, H, &Hv); }
void ref_vtkfile(int step) { vtkfile(step, H, &Hv); }

/* Runs up to k steps of the reference time loop; returns the steps taken.
 * dt_log[s] / t_log[s] receive dt used by, and t after, each step. */
int ref_steps(int k, double *dt_log, double *t_log) {
  static real_t dt = 0;  /* main.c:144 */
  int done = 0;
  while ((H.t < H.tend) && (H.nstep < H.nstepmax) && done < k) {
    if ((H.nstep % 2) == 0) {
      dt = 0;
      compute_deltat(&dt, H, &Hw_deltat, &Hv, &Hvw_deltat);
      if (H.nstep == 0) dt = dt / 2.0;
    }
    if ((H.nstep % 2) == 0)
      hydro_godunov(1, dt, H, &Hv, &Hw_godunov, &Hvw_godunov);
    else
      hydro_godunov(2, dt, H, &Hv, &Hw_godunov, &Hvw_godunov);
    H.nstep++;
    H.t += dt;
    if (dt_log) dt_log[done] = dt;
    if (t_log) t_log[done] = H.t;
    done++;
  }
  return done;
}

#ifndef PREC_SP
/* Calls the reference riemann() on one line of nface faces (SoA [4][nl]
 * arrays), allocating the hydrowork_t scratch it expects. */
#include "riemann.h"
void ref_riemann_line(int nface, int nl, double smallr, double smallc, double gamma, int niter,
                      double *ql, double *qr, double *qg, int *sgnm) {
  hydrowork_t w;
  memset(&w, 0, sizeof(w));
  double *blk = (double *) calloc((size_t) nface * 10, sizeof(double));
  w.pstar = blk; w.rl = blk + nface; w.ul = blk + 2 * nface; w.pl = blk + 3 * nface;
  w.rr = blk + 4 * nface; w.ur = blk + 5 * nface; w.pr = blk + 6 * nface;
  w.cl = blk + 7 * nface; w.cr = blk + 8 * nface;
  w.goon = (int *) calloc((size_t) nface, sizeof(int));
  riemann(nface, smallr, smallc, gamma, niter, 4, nl, 1, 1, (void *) ql, (void *) qr, (void *) qg,
          (void *) sgnm, &w);
  free(w.goon);
  free(blk);
}
#endif

/* The reference's process_args alone (no allocation): reals as ref_get_reals,
 * ints: nx, ny, nxystep, nstepmax, testcase, niter_riemann, iorder, scheme,
 * boundary left, right, down, up, noutput, nvar. */
void ref_parse(const char *nml, double *reals, int *ints) {
  hydroparam_t P;
  char *argv[3];
  argv[0] = "hydro";
  argv[1] = "-i";
  argv[2] = (char *) nml;
  process_args(3, argv, &P);
  reals[0] = P.t; reals[1] = P.tend; reals[2] = P.dx; reals[3] = P.gamma;
  reals[4] = P.courant_factor; reals[5] = P.smallc; reals[6] = P.smallr; reals[7] = P.slope_type;
  reals[8] = P.dtoutput;
  ints[0] = P.nx; ints[1] = P.ny; ints[2] = P.nxystep; ints[3] = P.nstepmax; ints[4] = P.testCase;
  ints[5] = P.niter_riemann; ints[6] = P.iorder; ints[7] = P.scheme; ints[8] = P.boundary_left;
  ints[9] = P.boundary_right; ints[10] = P.boundary_down; ints[11] = P.boundary_up;
  ints[12] = P.noutput; ints[13] = P.nvar;
}
