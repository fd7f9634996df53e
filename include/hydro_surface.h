/*
 * hydro_surface.h -- the HydroC99_2DMpi call surface, served by libhydro_amd.
 *
 * A program written against the reference's headers (parametres.h,
 * hydro_funcs.h, compute_deltat.h, hydro_godunov.h, make_boundary.h,
 * vtkfile.h) links against libhydro_amd.so unchanged: the structs below have
 * the reference's field order and types (real_t = double) and every function
 * has the reference's name and signature.  Behind them the state lives in GPU
 * memory and the hot path runs as HIP kernels on gfx950.
 *
 * Semantics that differ from the CPU reference, by design:
 *  - Hv->uold is a host mirror.  hydro_godunov / make_boundary update the GPU
 *    copy only; vtkfile() and hydro_amd_sync_host() refresh the mirror.  A
 *    caller that edits Hv->uold on the host must call hydro_amd_sync_device()
 *    (declared in hydro_amd.h) before the next step.  The reference driver
 *    (main.c) never touches uold between steps, so it needs neither.
 *  - hydro_init() builds the initial state on the host only; the GPU state
 *    (state + ping-pong buffer + CFL slots) is created from it on the first
 *    GPU-side call (allocate_work_space, compute_deltat, hydro_godunov, ...).
 *    hydrowork_t / hydrovarwork_t are opaque: the GPU path keeps no host work
 *    arrays, their pointers are left NULL.
 *  - Failures print a message and abort(), as the reference's asserts do.
 *  - Not reentrant, one thread at a time (as the reference).
 */
#ifndef HYDRO_SURFACE_H
#define HYDRO_SURFACE_H

#ifdef __cplusplus
extern "C" {
#endif

/* HydroC99_2DMpi/Src/parametres.h:41-45 */
typedef double real_t;

/* HydroC99_2DMpi/Src/parametres.h:47-53 */
typedef enum {
  XMIN_BOX, XMAX_BOX,
  YMIN_BOX, YMAX_BOX,
  UP_BOX, DOWN_BOX,
  LEFT_BOX, RIGHT_BOX,
  MAX_BOX
} Box_t;

/* HydroC99_2DMpi/Src/parametres.h:55-95 (field order and types preserved) */
typedef struct _hydroparam {
  int prt;
  real_t t, tend;
  int nstep, nstepmax;
  int noutput;
  real_t dtoutput;
  int imin, imax, jmin, jmax, nx, ny, nxt, nyt, nxyt, nxystep;
  int nproc, mype;
  int globnx, globny;
  int box[MAX_BOX];
  int nvar;
  real_t dx;
  real_t gamma;
  real_t courant_factor;
  real_t smallc, smallr;
  int niter_riemann;
  int iorder;
  real_t slope_type;
  int scheme;
  int boundary_right, boundary_left, boundary_down, boundary_up;
  int testCase;
} hydroparam_t;

#define HSCHEME_MUSCL 1
#define HSCHEME_PLMDE 2
#define HSCHEME_COLLELA 3

/* HydroC99_2DMpi/Src/parametres.h:102-104: uold is [nvar][nyt][nxt] */
typedef struct _hydrovar {
  real_t *uold;
} hydrovar_t;

/* HydroC99_2DMpi/Src/parametres.h:112-115 (opaque here) */
typedef struct _hydrovarwork {
  real_t *u, *q, *qxm, *qxp, *dq;
  real_t *qleft, *qright, *qgdnv, *flux;
} hydrovarwork_t;

/* HydroC99_2DMpi/Src/parametres.h:122-136 (opaque here) */
typedef struct _hydrowork {
  real_t *c;
  real_t *e;
  real_t *tmpm1, *tmpm2;
  real_t *rl, *ul, *pl, *cl, *wl;
  real_t *rr, *ur, *pr, *cr, *wr;
  real_t *ro, *uo, *po, *co, *wo;
  real_t *rstar, *ustar, *pstar, *cstar;
  real_t *spin, *spout, *ushock;
  int *sgnm;
  int *goon;
  real_t *frac, *scr, *delp, *pold;
  int *ind, *ind2;
} hydrowork_t;

/* HydroC99_2DMpi/Src/parametres.h:188-204 */
typedef enum {
  TIM_GATCON, TIM_CONPRI, TIM_EOS, TIM_SLOPE, TIM_TRACE, TIM_QLEFTR,
  TIM_RIEMAN, TIM_CMPFLX, TIM_UPDCON, TIM_COMPDT, TIM_MAKBOU, TIM_ALLRED,
  TIM_END
} Timers_t;
extern double functim[TIM_END];

/* HydroC99_2DMpi/Src/perfcnt.h:57-59 (kept at zero: no per-stage FLOP
 * counting on the GPU) */
extern long flopsAri, flopsSqr, flopsMin, flopsTra;
extern double MflopsSUM;
extern long nbFLOPS;

/* Parses `-i deck.nml` (and -v) into H: HydroC99_2DMpi/Src/parametres.c:262-452.
 * The .nml keys are the reference's (parametres.c:162-257) plus `gamma`
 * (default 1.4, which the reference hard-wires in default_values()).
 * Unknown keys are ignored, as in the reference. */
void process_args(int argc, char **argv, hydroparam_t *H);

/* HydroC99_2DMpi/Src/hydro_funcs.h:41: grid dims, host uold, initial
 * condition of hydro_funcs.c:79-146 (testcases 0 centre, 1 corner, 2 Sod
 * column; others abort as in the reference). */
void hydro_init(hydroparam_t *H, hydrovar_t *Hv);
/* HydroC99_2DMpi/Src/hydro_funcs.h:42 */
void hydro_finish(const hydroparam_t H, hydrovar_t *Hv);
/* HydroC99_2DMpi/Src/hydro_funcs.h:44-45 */
void allocate_work_space(int ngrid, const hydroparam_t H, hydrowork_t *Hw, hydrovarwork_t *Hvw);
void deallocate_work_space(int ngrid, const hydroparam_t H, hydrowork_t *Hw, hydrovarwork_t *Hvw);
/* HydroC99_2DMpi/Src/compute_deltat.h:41-44 */
void compute_deltat_init_mem(const hydroparam_t H, hydrowork_t *Hw, hydrovarwork_t *Hvw);
void compute_deltat_clean_mem(const hydroparam_t H, hydrowork_t *Hw, hydrovarwork_t *Hvw);
void compute_deltat(real_t *dt, const hydroparam_t H, hydrowork_t *Hw, hydrovar_t *Hv,
                    hydrovarwork_t *Hvw);
/* HydroC99_2DMpi/Src/hydro_godunov.h:40 -- both sweeps, idim first */
void hydro_godunov(int idim, real_t dt, const hydroparam_t H, hydrovar_t *Hv, hydrowork_t *Hw,
                   hydrovarwork_t *Hvw);
/* HydroC99_2DMpi/Src/make_boundary.h:40 */
void make_boundary(int idim, const hydroparam_t H, hydrovar_t *Hv);
/* HydroC99_2DMpi/Src/vtkfile.h:39 -- Dep/NNNN/NN/Hydro_RRRRR_SSSS.vtr + .pvtr
 * + Hydro.pvd, base64 Float32 cell data, the reference's format */
void vtkfile(int step, const hydroparam_t H, hydrovar_t *Hv);
/* HydroC99_2Domp/Src/Image.h:30 (the OpenMP variant's image output, its
 * default PPM build): writes IMAGE_nnnnnn.ppm (prefixed by $HYDROC_IMG_PREFIX
 * when set), "P6\n<w> <h>\n255\n", w*h grey RGB pixels (uint8_t)
 * fabs(255 * rho / max rho), then "#EOF\n"; the image is at most 1000
 * pixels a side (Image.c:124-126, shrinkSize 1000 of parametres.c:68).
 * Computed on the GPU from the device state; aborts when the file cannot be
 * opened, as Image.c:298-310. */
void DumpImage(int n, hydroparam_t *H, hydrovar_t *Hv);

#ifdef __cplusplus
}
#endif
#endif
