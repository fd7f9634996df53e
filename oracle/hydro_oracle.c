/*
 * hydro_oracle.c -- CPU restatement of HydroC99's split Godunov/MUSCL step.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle): used by tests/, the smoke check and
 * bench.py's CPU-baseline leg as the reference check; never linked into the product.
 *
 * Every function restates one piece of HydroC/HydroC99_2DMpi/Src and cites the
 * lines it follows.  The operation order of each expression is the reference's
 * (C evaluates a*b*c as (a*b)*c, etc.) and the file is built with
 * -ffp-contract=off, so results are bit-identical to the reference on the same
 * inputs.  The structure is not the reference's: it works one line (row or
 * column) at a time through small per-thread buffers instead of slabs of
 * whole-stage arrays.
 *
 * Parity pins (tests/test_oracle.py): every stage and the whole trajectory are
 * compared bitwise with the compiled reference (oracle/_ref) and with the
 * committed golden vectors, plus closed-form / invariant checks that do not
 * depend on the reference (conservation, symmetry, special cases).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "hydro_oracle.h"

#define NV 4
#define ID 0
#define IU 1
#define IV 2
#define IP 3

/* HydroC99_2DMpi/Src/utils.h:41-47 */
#define SQ(x) ((x) * (x))
#define MAXM(x, y) ((x) > (y) ? (x) : (y))

static const double HALF = 1.0 / 2.0;

void ho_default_params(ho_params *p) {
  /* HydroC99_2DMpi/Src/parametres.c:64-109 */
  memset(p, 0, sizeof(*p));
  p->nx = 20;
  p->ny = 20;
  p->dx = 1.0;
  p->gamma = 1.4;
  p->courant_factor = 1.0 / 2.0;
  p->smallc = 1e-10;
  p->smallr = 1e-10;
  p->niter_riemann = 10;
  p->iorder = 2;
  p->slope_type = 1.;
  p->scheme = 1;
  p->boundary_left = p->boundary_right = p->boundary_down = p->boundary_up = 1;
}

static inline long idx(const ho_params *p, int i, int j, int v) {
  long nxt = p->nx + 4, nyt = p->ny + 4;
  return i + nxt * (j + nyt * (long) v);
}

void ho_init(const ho_params *p, int testcase, double *uold) {
  /* HydroC99_2DMpi/Src/hydro_funcs.c:79-128, nproc == 1 branches; the ghost
   * layers start at zero (utils.c:93-110 callocs). */
  int nxt = p->nx + 4, nyt = p->ny + 4;
  memset(uold, 0, sizeof(double) * NV * (size_t) nxt * nyt);
  for (int j = 2; j < nyt - 2; j++)
    for (int i = 2; i < nxt - 2; i++) {
      uold[idx(p, i, j, ID)] = 1.0;
      uold[idx(p, i, j, IU)] = 0.0;
      uold[idx(p, i, j, IV)] = 0.0;
      uold[idx(p, i, j, IP)] = 1e-5;
    }
  double eb = 1.0 / p->dx / p->dx;
  if (testcase == 0) uold[idx(p, nxt / 2, nyt / 2, IP)] = eb;
  if (testcase == 1) uold[idx(p, 2, 2, IP)] = eb;
  if (testcase == 2)
    for (int j = 2; j < nyt; j++) uold[idx(p, 2, j, IP)] = eb;
}

void ho_init_blast(const ho_params *p, int bx, int by, double e_blast, double *uold) {
  ho_init(p, -1, uold);
  uold[idx(p, bx + 2, by + 2, IP)] = e_blast;
}

/* ---- stage functions over one line --------------------------------------- */

void ho_constoprim(int n, const double *u, double *q, double *e, double smallr) {
  /* HydroC99_2DMpi/Src/constoprim.c:62-78 */
  for (int i = 0; i < n; i++) {
    double r = MAXM(u[ID * n + i], smallr);
    double vx = u[IU * n + i] / r;
    double vy = u[IV * n + i] / r;
    double eken = HALF * (SQ(vx) + SQ(vy));
    double ei = u[IP * n + i] / r - eken;
    q[ID * n + i] = r;
    q[IU * n + i] = vx;
    q[IV * n + i] = vy;
    q[IP * n + i] = ei;
    e[i] = ei;
  }
}

void ho_eos(int n, const double *e, double *q, double *c, double smallc, double gamma) {
  /* HydroC99_2DMpi/Src/equation_of_state.c:64-81 */
  double smallp = SQ(smallc) / gamma;
  for (int k = 0; k < n; k++) {
    double rho = q[ID * n + k];
    double p = (gamma - 1.0) * rho * e[k];
    p = MAXM(p, rho * smallp);
    q[IP * n + k] = p;
    c[k] = sqrt(gamma * p / rho);
  }
}

void ho_slope(int n, const double *q, double *dq, double st) {
  /* HydroC99_2DMpi/Src/slope.c:71-90 (the non-OLDSTYLE limiter) */
  for (int v = 0; v < NV; v++) {
    const double *a = q + v * n;
    for (int i = 1; i < n - 1; i++) {
      double dl = st * (a[i] - a[i - 1]);
      double dr = st * (a[i + 1] - a[i]);
      double dc = HALF * (dl + dr) / st;
      double sg = (dc > 0) ? 1.0 : -1.0;
      int ll = (dl * dr) <= 0.0;
      double t1 = fmin(fabs(dl), fabs(dr));
      dq[v * n + i] = sg * fmin((1 - ll) * t1, fabs(dc));
    }
  }
}

void ho_trace(int n, const double *q, const double *dq, const double *c, double dtdx,
              int scheme, double *qxm, double *qxp) {
  /* HydroC99_2DMpi/Src/trace.c:69-147 */
  double zerol = 0.0, zeror = 0.0, project = 0.0;
  if (scheme == 1) { zerol = -100.0 / dtdx; zeror = 100.0 / dtdx; project = 1.0; }
  if (scheme == 2) { zerol = 0.0; zeror = 0.0; project = 1.0; }
  if (scheme == 3) { zerol = 0.0; zeror = 0.0; project = 0.0; }
  for (int i = 1; i < n - 1; i++) {
    double cc = c[i];
    double csq = SQ(cc);
    double r = q[ID * n + i], u = q[IU * n + i], v = q[IV * n + i], p = q[IP * n + i];
    double dr = dq[ID * n + i], du = dq[IU * n + i], dv = dq[IV * n + i], dp = dq[IP * n + i];
    double rOcc = r / cc;
    double OrOcc = cc / r;
    double dprcc = dp / (r * cc);
    double am = HALF * (dprcc - du) * rOcc;
    double ap = HALF * (dprcc + du) * rOcc;
    double a0r = dr - dp / csq;
    double a0v = dv;
    double upcc = u + cc, umcc = u - cc;
    double upccx = upcc * dtdx, umccx = umcc * dtdx, ux = u * dtdx;
    /* plus side (qxp) */
    double spm = (umcc >= zeror) ? project : umccx + 1.0;
    double spp = (upcc >= zeror) ? project : upccx + 1.0;
    double sp0 = (u >= zeror) ? project : ux + 1.0;
    double apr = -HALF * spp * ap;
    double amr = -HALF * spm * am;
    double azr = -HALF * sp0 * a0r;
    double azv = -HALF * sp0 * a0v;
    qxp[ID * n + i] = r + (apr + amr + azr);
    qxp[IU * n + i] = u + (apr - amr) * OrOcc;
    qxp[IV * n + i] = v + (azv);
    qxp[IP * n + i] = p + (apr + amr) * csq;
    /* minus side (qxm) */
    spm = (umcc <= zerol) ? -project : umccx - 1.0;
    spp = (upcc <= zerol) ? -project : upccx - 1.0;
    sp0 = (u <= zerol) ? -project : ux - 1.0;
    double apl = -HALF * spp * ap;
    double aml = -HALF * spm * am;
    double azl = -HALF * sp0 * a0r;
    double azvl = -HALF * sp0 * a0v;
    qxm[ID * n + i] = r + (apl + aml + azl);
    qxm[IU * n + i] = u + (apl - aml) * OrOcc;
    qxm[IV * n + i] = v + (azvl);
    qxm[IP * n + i] = p + (apl + aml) * csq;
  }
}

void ho_qleftright(int nface, int n, const double *qxm, const double *qxp, double *ql, double *qr) {
  /* HydroC99_2DMpi/Src/qleftright.c:63-78: face k takes qxm of cell k+1 and
   * qxp of cell k+2 */
  for (int v = 0; v < NV; v++)
    for (int k = 0; k < nface; k++) {
      ql[v * n + k] = qxm[v * n + k + 1];
      qr[v * n + k] = qxp[v * n + k + 2];
    }
}

void ho_riemann(int nface, int nl, const double *ql, const double *qr, double *qg, int *sgnm,
                double smallr, double smallc, double gamma, int niter) {
  /* HydroC99_2DMpi/Src/riemann.c:91-287, one face at a time.  The reference
   * runs the Newton loop over all faces with a per-face `goon` flag; stopping
   * one face early is the same as its flag going false. */
  const double smallp = SQ(smallc) / gamma;
  const double gamma6 = (gamma + 1.0) / (2.0 * gamma);
  const double smallpp = smallr * smallp;
  for (int i = 0; i < nface; i++) {
    double rl = fmax(ql[ID * nl + i], smallr);
    double ul = ql[IU * nl + i];
    double pl = fmax(ql[IP * nl + i], rl * smallp);
    double rr = fmax(qr[ID * nl + i], smallr);
    double ur = qr[IU * nl + i];
    double pr = fmax(qr[IP * nl + i], rr * smallp);
    double cl = gamma * pl * rl;
    double cr = gamma * pr * rr;
    double wl = sqrt(cl);
    double wr = sqrt(cr);
    double pst = fmax(((wr * pl + wl * pr) + wl * wr * (ul - ur)) / (wl + wr), 0.0);
    for (int it = 0; it < niter; it++) {
      /* riemann.c:176-201 */
      double wwl = sqrt(cl * (1.0 + gamma6 * (pst - pl) / pl));
      double wwr = sqrt(cr * (1.0 + gamma6 * (pst - pr) / pr));
      double swwl = SQ(wwl);
      double qlw = 2.0 * wwl * swwl / (swwl + cl);
      double qrw = 2.0 * wwr * SQ(wwr) / (SQ(wwr) + cr);
      double usl = ul - (pst - pl) / wwl;
      double usr = ur + (pst - pr) / wwr;
      double t = (qrw * qlw / (qrw + qlw) * (usl - usr));
      double d = (t > (-pst)) ? t : (-pst);
      pst += d;
      double rel = d / (pst + smallpp);
      double arel = (rel > 0) ? rel : -rel;
      if (!(arel > 1e-6)) break;
    }
    /* riemann.c:211-276 */
    wr = sqrt(cr * (1.0 + gamma6 * (pst - pr) / pr));
    wl = sqrt(cl * (1.0 + gamma6 * (pst - pl) / pl));
    double ustar = HALF * (ul + (pl - pst) / wl + ur - (pr - pst) / wr);
    int left = ustar > 0;
    int sg = left ? 1 : -1;
    double ro = left ? rl : rr;
    double uo = left ? ul : ur;
    double po = left ? pl : pr;
    double wo = left ? wl : wr;
    double co = sqrt(fabs(gamma * po / ro));
    co = fmax(smallc, co);
    double rstar = ro / (1.0 + ro * (po - pst) / SQ(wo));
    rstar = fmax(rstar, smallr);
    double cstar = sqrt(fabs(gamma * pst / rstar));
    cstar = fmax(smallc, cstar);
    double spout = co - sg * uo;
    double spin = cstar - sg * ustar;
    double ushock = wo / ro - sg * uo;
    if (pst >= po) {
      spin = ushock;
      spout = ushock;
    }
    double scr = fmax(spout - spin, smallc + fabs(spout + spin));
    double frac = (1.0 + (spout + spin) / scr) * HALF;
    frac = fmax(0.0, fmin(1.0, frac));
    double gr, gu, gp;
    if (spout < 0.0) {
      gr = ro; gu = uo; gp = po;
    } else if (spin > 0.0) {
      gr = rstar; gu = ustar; gp = pst;
    } else {
      gr = frac * rstar + (1.0 - frac) * ro;
      gu = frac * ustar + (1.0 - frac) * uo;
      gp = frac * pst + (1.0 - frac) * po;
    }
    qg[ID * nl + i] = gr;
    qg[IU * nl + i] = gu;
    qg[IP * nl + i] = gp;
    qg[IV * nl + i] = left ? ql[IV * nl + i] : qr[IV * nl + i];
    if (sgnm) sgnm[i] = sg;
  }
}

void ho_cmpflx(int nface, int nl, const double *qg, double *flux, double gamma) {
  /* HydroC99_2DMpi/Src/cmpflx.c:62-91 */
  double entho = 1.0 / (gamma - 1.0);
  for (int i = 0; i < nface; i++) {
    double r = qg[ID * nl + i], u = qg[IU * nl + i], p = qg[IP * nl + i], v = qg[IV * nl + i];
    double m = r * u;
    flux[ID * nl + i] = m;
    flux[IU * nl + i] = m * u + p;
    flux[IV * nl + i] = m * v;
    double ekin = HALF * r * (SQ(u) + SQ(v));
    double etot = p * entho + ekin;
    flux[IP * nl + i] = u * (etot + p);
  }
}

/* ---- whole grid ------------------------------------------------------------ */

void ho_make_boundary(int idim, const ho_params *p, double *uold) {
  /* HydroC99_2DMpi/Src/make_boundary.c:237-289 (x) and 381-433 (y), with the
   * reference's source-index choices for types 1 (reflect), 2 (absorb) and
   * any other value (periodic); ghosts are filled only along interior
   * rows/columns. */
  const int nx = p->nx, ny = p->ny;
  if (idim == 1) {
    if (p->boundary_left > 0)
      for (int v = 0; v < NV; v++)
        for (int i = 0; i < 2; i++) {
          double sign = 1.0;
          int i0;
          if (p->boundary_left == 1) { i0 = 3 - i; if (v == IU) sign = -1.0; }
          else if (p->boundary_left == 2) i0 = 2;
          else i0 = nx + i;
          for (int j = 2; j < ny + 2; j++) uold[idx(p, i, j, v)] = uold[idx(p, i0, j, v)] * sign;
        }
    if (p->boundary_right > 0)
      for (int v = 0; v < NV; v++)
        for (int i = nx + 2; i < nx + 4; i++) {
          double sign = 1.0;
          int i0;
          if (p->boundary_right == 1) { i0 = 2 * nx + 3 - i; if (v == IU) sign = -1.0; }
          else if (p->boundary_right == 2) i0 = nx + 2;
          else i0 = i - nx;
          for (int j = 2; j < ny + 2; j++) uold[idx(p, i, j, v)] = uold[idx(p, i0, j, v)] * sign;
        }
  } else {
    if (p->boundary_down > 0)
      for (int v = 0; v < NV; v++)
        for (int j = 0; j < 2; j++) {
          double sign = 1.0;
          int j0;
          if (p->boundary_down == 1) { j0 = 3 - j; if (v == IV) sign = -1.0; }
          else if (p->boundary_down == 2) j0 = 4;
          else j0 = ny + j;
          for (int i = 2; i < nx + 2; i++) uold[idx(p, i, j, v)] = uold[idx(p, i, j0, v)] * sign;
        }
    if (p->boundary_up > 0)
      for (int v = 0; v < NV; v++)
        for (int j = ny + 2; j < ny + 4; j++) {
          double sign = 1.0;
          int j0;
          if (p->boundary_up == 1) { j0 = 2 * ny + 3 - j; if (v == IV) sign = -1.0; }
          else if (p->boundary_up == 2) j0 = ny + 1;
          else j0 = j - ny;
          for (int i = 2; i < nx + 2; i++) uold[idx(p, i, j, v)] = uold[idx(p, i, j0, v)] * sign;
        }
  }
}

double ho_compute_deltat(const ho_params *p, const double *uold) {
  /* HydroC99_2DMpi/Src/compute_deltat.c:55-95 (primitives of interior cells),
   * equation_of_state.c:64-81, courantOnXY :100-145 and the dt formula :197.
   * For finite Courant numbers the maxima are exact, so the per-row partial
   * maxima combine in any order.  Once the state holds a NaN, MAXM (x > y ?
   * x : y, the reference's MAX) drops the running maximum at the NaN, and
   * this result -- like the reference's -- depends on the reduction order,
   * i.e. on the OpenMP thread count; parity is checked on finite runs. */
  const int nx = p->nx, ny = p->ny;
  const double smallp = SQ(p->smallc) / p->gamma;
  double cx = 0.0, cy = 0.0;
#pragma omp parallel for reduction(max : cx, cy) schedule(static)
  for (int j = 2; j < ny + 2; j++) {
    for (int i = 2; i < nx + 2; i++) {
      double r = MAXM(uold[idx(p, i, j, ID)], p->smallr);
      double vx = uold[idx(p, i, j, IU)] / r;
      double vy = uold[idx(p, i, j, IV)] / r;
      double eken = HALF * (SQ(vx) + SQ(vy));
      double e = uold[idx(p, i, j, IP)] / r - eken;
      double pr = (p->gamma - 1.0) * r * e;
      pr = MAXM(pr, r * smallp);
      double c = sqrt(p->gamma * pr / r);
      double ax = c + fabs(vx), ay = c + fabs(vy);
      cx = MAXM(cx, ax);
      cy = MAXM(cy, ay);
    }
  }
  return p->courant_factor * p->dx / MAXM(cx, MAXM(cy, p->smallc));
}

/* One sweep over every interior line of direction idim
 * (HydroC99_2DMpi/Src/hydro_godunov.c:101-237 restated per line). */
static void sweep(int idim, double dtdx, const ho_params *p, double *uold) {
  const int nx = p->nx, ny = p->ny, nxt = nx + 4, nyt = ny + 4;
  const int n = (idim == 1) ? nxt : nyt;          /* cells on a line */
  const int nface = (idim == 1) ? nx + 1 : ny + 1; /* Hndim_1 */
  const int nlines = (idim == 1) ? ny : nx;
  /* gather/update slots: the y sweep swaps IU and IV (conservar.c:106-109,
   * 177-180) so the normal momentum always sits in slot IU */
  const int vmap[NV] = {ID, idim == 1 ? IU : IV, idim == 1 ? IV : IU, IP};
#pragma omp parallel
  {
    double *buf = (double *) malloc(sizeof(double) * (size_t) n * (NV * 9 + 2));
    int *sg = (int *) malloc(sizeof(int) * (size_t) n);
    double *u = buf, *q = u + NV * n, *dq = q + NV * n, *qxm = dq + NV * n, *qxp = qxm + NV * n;
    double *ql = qxp + NV * n, *qr = ql + NV * n, *qg = qr + NV * n, *fl = qg + NV * n;
    double *e = fl + NV * n, *c = e + n;
#pragma omp for schedule(static)
    for (int line = 0; line < nlines; line++) {
      const int l = line + 2;
      for (int v = 0; v < NV; v++)
        for (int k = 0; k < n; k++)
          u[v * n + k] = (idim == 1) ? uold[idx(p, k, l, vmap[v])] : uold[idx(p, l, k, vmap[v])];
      ho_constoprim(n, u, q, e, p->smallr);
      ho_eos(n, e, q, c, p->smallc, p->gamma);
      if (p->iorder != 1) ho_slope(n, q, dq, p->slope_type);
      else memset(dq, 0, sizeof(double) * NV * n);
      ho_trace(n, q, dq, c, dtdx, p->scheme, qxm, qxp);
      ho_qleftright(nface, n, qxm, qxp, ql, qr);
      ho_riemann(nface, n, ql, qr, qg, sg, p->smallr, p->smallc, p->gamma, p->niter_riemann);
      ho_cmpflx(nface, n, qg, fl, p->gamma);
      /* conservar.c:147-197: U_i += (F_{i-2} - F_{i-1}) * dtdx */
      for (int v = 0; v < NV; v++)
        for (int k = 2; k < n - 2; k++) {
          double nu = u[v * n + k] + (fl[v * n + k - 2] - fl[v * n + k - 1]) * dtdx;
          if (idim == 1) uold[idx(p, k, l, vmap[v])] = nu;
          else uold[idx(p, l, k, vmap[v])] = nu;
        }
    }
    free(sg);
    free(buf);
  }
}

void ho_godunov(int idim_start, double dt, const ho_params *p, double *uold) {
  /* HydroC99_2DMpi/Src/hydro_godunov.c:98-110: both directions, boundary
   * refilled before each */
  for (int k = 0; k < 2; k++) {
    int idim = (idim_start - 1 + k) % 2 + 1;
    double dtdx = dt / p->dx;
    ho_make_boundary(idim, p, uold);
    sweep(idim, dtdx, p, uold);
  }
}

int ho_run(const ho_params *p, double *uold, double tend, int nstepmax, int max_steps,
           double *t, int *nstep, double *dt, double *dt_log) {
  /* HydroC99_2DMpi/Src/main.c:244-298 */
  int done = 0;
  while ((*t < tend) && (*nstep < nstepmax) && done < max_steps) {
    if ((*nstep % 2) == 0) {
      *dt = ho_compute_deltat(p, uold);
      if (*nstep == 0) *dt = *dt / 2.0;
    }
    ho_godunov((*nstep % 2) == 0 ? 1 : 2, *dt, p, uold);
    if (dt_log) dt_log[done] = *dt;
    (*nstep)++;
    *t += *dt;
    done++;
  }
  return done;
}

int ho_image(int nx, int ny, const double *uold, int shrink_size, unsigned char *rgb, int *w,
             int *h) {
  /* HydroC99_2Domp/Src/Image.c:111-166 for one rank: image size (:124-131),
   * the maximum density over every padded cell scanned from cell (0, 0) with
   * `v > max` (getMaxVarValues, :71-110), then grey pixels of interior cells
   * (2 + cx*shrink, 2 + cy*shrink).  The byte is the reference's own C
   * expression, so this compiles to the same conversion as the reference's
   * gcc build (cvttsd2si to int32, low byte). */
  const int nxt = nx + 4, nyt = ny + 4;
  int shrink = 1;
  while ((nx / shrink) > shrink_size || (ny / shrink) > shrink_size) shrink++;
  *w = nx / shrink;
  *h = ny / shrink;
  if (!rgb) return shrink;
  double mx = uold[0];
  for (int y = 0; y < nyt; y++)
    for (int x = 0; x < nxt; x++) {
      double v = uold[(long) y * nxt + x];
      if (v > mx) mx = v;
    }
  unsigned char *ptr = rgb;
  for (int cy = 0; cy < *h; cy++)
    for (int cx = 0; cx < *w; cx++) {
      int x = 2 + cx * shrink, y = 2 + cy * shrink;
      unsigned char vd = (unsigned char) fabs(255 * uold[(long) y * nxt + x] / mx);
      *(ptr++) = vd;
      *(ptr++) = vd;
      *(ptr++) = vd;
    }
  return shrink;
}
