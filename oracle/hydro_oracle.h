/*
 * hydro_oracle.h -- CPU restatement of HydroC99's split Godunov/MUSCL step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the oracle: tests, the smoke check and the
 * CPU-baseline leg of bench.py use it as the reference check.  The product path (the
 * HIP library) never links or calls it.
 *
 * Layout follows the reference so arrays compare element for element:
 * uold is [4][nyt][nxt] doubles, nxt = nx + 4, nyt = ny + 4 (two ghost layers),
 * index (i, j, v) -> i + nxt * (j + nyt * v)   (HydroC99_2DMpi/Src/parametres.h:107).
 * Variable slots: 0 = rho, 1 = rho*u, 2 = rho*v, 3 = E.
 *
 * Line ("pencil") arrays used by the stage functions are SoA [4][n]: the
 * reference's [nvar][Hstep][Hnxyt] work arrays with Hstep = 1 and Hnxyt = n.
 */
#ifndef HYDRO_ORACLE_H
#define HYDRO_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int nx, ny;                 /* interior cells */
  double dx;
  double gamma;
  double courant_factor;
  double smallc, smallr;
  double slope_type;
  int niter_riemann;
  int iorder;
  int scheme;                 /* 1 muscl, 2 plmde, 3 collela */
  int boundary_left, boundary_right, boundary_down, boundary_up;
} ho_params;

/* defaults of HydroC99_2DMpi/Src/parametres.c:64-109 */
void ho_default_params(ho_params *p);

/* initial state of HydroC99_2DMpi/Src/hydro_funcs.c:79-146 (single rank) */
void ho_init(const ho_params *p, int testcase, double *uold);
/* generalised blast used by the batched workloads: background rho=1, u=v=0,
 * E=1e-5 and E=e_blast at interior cell (bx, by) (0-based interior index) */
void ho_init_blast(const ho_params *p, int bx, int by, double e_blast, double *uold);

/* stage functions over one line of n cells (SoA [4][n]) */
void ho_constoprim(int n, const double *u, double *q, double *e, double smallr);
void ho_eos(int n, const double *e, double *q, double *c, double smallc, double gamma);
void ho_slope(int n, const double *q, double *dq, double slope_type);
void ho_trace(int n, const double *q, const double *dq, const double *c, double dtdx,
              int scheme, double *qxm, double *qxp);
void ho_qleftright(int nface, int n, const double *qxm, const double *qxp, double *ql, double *qr);
void ho_riemann(int nface, int nl, const double *ql, const double *qr, double *qg, int *sgnm,
                double smallr, double smallc, double gamma, int niter);
void ho_cmpflx(int nface, int nl, const double *qg, double *flux, double gamma);

/* whole-grid entry points (the HydroC99_2DMpi call surface, restated) */
void ho_make_boundary(int idim, const ho_params *p, double *uold);
double ho_compute_deltat(const ho_params *p, const double *uold);
void ho_godunov(int idim_start, double dt, const ho_params *p, double *uold);

/* time loop of HydroC99_2DMpi/Src/main.c:244-298.  Runs while t < tend and
 * nstep < nstepmax and at most max_steps steps; *t, *nstep and *dt carry the
 * loop state between calls.  dt_log (may be NULL) gets the dt of each step. */
int ho_run(const ho_params *p, double *uold, double tend, int nstepmax, int max_steps,
           double *t, int *nstep, double *dt, double *dt_log);

/* density image of HydroC99_2Domp/Src/Image.c:111-166 (one rank): returns
 * the shrink factor, sets *w, *h; fills rgb (3*w*h bytes) unless it is NULL */
int ho_image(int nx, int ny, const double *uold, int shrink_size, unsigned char *rgb, int *w,
             int *h);

#ifdef __cplusplus
}
#endif
#endif
