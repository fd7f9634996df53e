/*
 * hydro -- the HydroC99 driver over libhydro_amd.
 *
 *   hydro -i deck.nml [-v] [-dtlog FILE] [-dump FILE] [-native [-chunk K]]
 *         [-restart FILE] [-checkpoint FILE]
 *
 * Default mode calls the reference call surface step by step exactly as
 * HydroC99_2DMpi/Src/main.c:244-364 does (dt on even steps, halved at step 0,
 * hydro_godunov(1) on even steps and (2) on odd ones, t += dt, VTK output on
 * noutput / dtoutput, the "--> step=" line with MC/s, the average MC/s over
 * steps > 5).  Density images follow HydroC99_2Domp/Src/main.c:212-221,
 * 462-479: with `nimage` or `dtimage` in the deck, IMAGE_000000.ppm of the
 * initial state, then one image every nimage steps (or each time t passes
 * the next multiple of dtimage), marked " <nnnn>" on the step line after
 * the VTK marker.  -native runs the same loop on the GPU (hydro_amd_grid_run) in
 * chunks of K steps, with one host round trip per chunk.
 *
 * -dtlog writes "step t dt" per step with %.17g; -dump writes the final state
 * as an FP64 binary dump (hydro_amd_dump).
 *
 * Checkpoint / restart and the time guard follow the C++ variant
 * (HydroCplusMPI/TimeLimit.cpp:22-55, Domain.cpp:103-122,
 * Domain_Compute.cpp:531-569): the allotted time comes from HYDROC_MAXTIME
 * (else BRIDGE_MPRUN_MAXTIME, BRIDGE_MSUB_MAXTIME), less the time already
 * spent since HYDROC_START_TIME (epoch seconds) if set; the guard is 900 s,
 * 600 s below an hour, 300 s below 30 min, 20 s below a minute; a run with
 * at most 1 s allotted aborts.  After each step (each chunk with -native) the
 * loop stops once the remaining time is below the guard.  Unlike the C++
 * variant, no limit applies when none of these variables is set (the C99
 * driver has none).  -checkpoint writes the state, t, nstep and dt at the end
 * of the loop, however it ended; -restart continues from such a file.
 */
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "../../include/hydro_amd.h"

static double now(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* Seconds allotted to this run (< 0: no limit). */
static double allotted_time(void) {
  const char *p = getenv("HYDROC_MAXTIME");
  if (!p) p = getenv("BRIDGE_MPRUN_MAXTIME");
  if (!p) p = getenv("BRIDGE_MSUB_MAXTIME");
  if (!p) return -1.0;
  double t = strtod(p, NULL);
  const char *s = getenv("HYDROC_START_TIME");
  if (s) t -= (double) ((long) time(NULL) - strtol(s, NULL, 10));
  return t;
}

int main(int argc, char **argv) {
  hydroparam_t H;
  hydrovar_t Hv;
  hydrowork_t Hw_godunov, Hw_deltat;
  hydrovarwork_t Hvw_godunov, Hvw_deltat;
  const char *dtlog_path = NULL, *dump_path = NULL, *restart_path = NULL, *ckpt_path = NULL;
  int native = 0, chunk = 100;
  char **pass = (char **) calloc((size_t) argc + 1, sizeof(char *));
  int npass = 0;
  pass[npass++] = argv[0];
  for (int n = 1; n < argc; n++) {
    if (strcmp(argv[n], "-dtlog") == 0 && n + 1 < argc) dtlog_path = argv[++n];
    else if (strcmp(argv[n], "-dump") == 0 && n + 1 < argc) dump_path = argv[++n];
    else if (strcmp(argv[n], "-native") == 0) native = 1;
    else if (strcmp(argv[n], "-chunk") == 0 && n + 1 < argc) chunk = atoi(argv[++n]);
    else if (strcmp(argv[n], "-restart") == 0 && n + 1 < argc) restart_path = argv[++n];
    else if (strcmp(argv[n], "-checkpoint") == 0 && n + 1 < argc) ckpt_path = argv[++n];
    else pass[npass++] = argv[n];
  }
  if (chunk < 1) chunk = 1;
  memset(functim, 0, sizeof(functim));
  process_args(npass, pass, &H);
  free(pass);
  const double t_begin = now();
  const double allotted = allotted_time();
  double guard = 0;
  if (allotted >= 0) {
    if (allotted <= 1) {
      fprintf(stderr, "Hydro: allocated time too short %g s\n", allotted);
      abort();
    }
    guard = allotted < 60 ? 20 : allotted < 1800 ? 300 : allotted < 3600 ? 600 : 900;
    fprintf(stdout, "Hydro: allocated time %g s, time guard %g s\n", allotted, guard);
  }
  hydro_init(&H, &Hv);
  double dt = 0;
  if (restart_path) {
    if (hydro_amd_restore(restart_path, &H, &Hv, &dt) != 0) {
      fprintf(stderr, "hydro: %s\n", hydro_amd_last_error());
      return 1;
    }
    fprintf(stdout, "Hydro restarts at step %d, t = %.17g\n", H.nstep, H.t);
  }
  fprintf(stdout, "Hydro starts in double precision.\n");
  fprintf(stdout, "Hydro: libhydro_amd on gfx950 (%s loop)\n", native ? "device" : "host");

  FILE *dtlog = dtlog_path ? fopen(dtlog_path, "w") : NULL;
  int nvtk = 0, time_output = 0;
  double next_output_time = 0;
  if (H.dtoutput > 0) {
    time_output = 1;
    next_output_time += H.dtoutput;
    while (next_output_time <= H.t) next_output_time += H.dtoutput;  /* after a restart */
  }
  if (H.dtoutput > 0 || H.noutput > 0) vtkfile(++nvtk, H, &Hv);
  int nimage = 0, nimg = 0, time_image = 0;
  double dtimage = 0, next_image_time = 0;
  hydro_amd_image_cadence(&nimage, &dtimage);
  if (dtimage > 0) {
    time_image = 1;
    next_image_time += dtimage;
    while (next_image_time <= H.t) next_image_time += dtimage;  /* after a restart */
  }
  if (dtimage > 0 || nimage > 0) DumpImage(nimg++, &H, &Hv);
  fprintf(stdout, "Hydro starts main loop.\n");
  allocate_work_space(H.nxyt, H, &Hw_godunov, &Hvw_godunov);
  compute_deltat_init_mem(H, &Hw_deltat, &Hvw_deltat);

  double mcsavg = 0, mcsmin = FLT_MAX, mcsmax = 0, mcssig = 0;
  long nmcs = 0;
  const double t_start = now();
  hydro_amd_grid *g = native ? hydro_amd_surface_grid(&Hv) : NULL;
  if (g && hydro_amd_grid_set_state(g, H.t, H.nstep, dt) != 0) {
    fprintf(stderr, "hydro: %s\n", hydro_amd_last_error());
    return 1;
  }
  while ((H.t < H.tend) && (H.nstep < H.nstepmax)) {
    double t0 = now();
    int steps = 1;
    double dts[1] = {0};
    double *chunk_dt = NULL;
    if (!native) {
      if ((H.nstep % 2) == 0) {
        dt = 0;
        compute_deltat(&dt, H, &Hw_deltat, &Hv, &Hvw_deltat);
        if (H.nstep == 0) {
          dt = dt / 2.0;
          fprintf(stdout, "Hydro computes initial deltat: %le\n", dt);
        }
      }
      hydro_godunov((H.nstep % 2) == 0 ? 1 : 2, dt, H, &Hv, &Hw_godunov, &Hvw_godunov);
      dts[0] = dt;
      chunk_dt = dts;
    } else {
      int want = chunk;
      if (H.noutput > 0 && !time_output) want = H.noutput - (H.nstep % H.noutput);
      if (nimage > 0 && !time_image && nimage - (H.nstep % nimage) < want)
        want = nimage - (H.nstep % nimage);
      if (time_output || time_image) want = 1;  /* output times are only known step by step */
      if (want > H.nstepmax - H.nstep) want = H.nstepmax - H.nstep;
      chunk_dt = (double *) malloc(sizeof(double) * (size_t) want);
      hydro_amd_grid_set_limits(g, H.tend, H.nstepmax);
      steps = hydro_amd_grid_run(g, want, chunk_dt);
      if (steps < 0) {
        fprintf(stderr, "hydro: %s\n", hydro_amd_last_error());
        return 1;
      }
      if (steps == 0) {
        free(chunk_dt);
        break;
      }
    }
    double t1 = now();
    double mcs = (double) H.globnx * H.globny * steps / (t1 - t0) / 1e6;
    for (int s = 0; s < steps; s++) {
      dt = chunk_dt[s];
      H.nstep++;
      H.t += dt;
      if (dtlog) fprintf(dtlog, "%d %.17g %.17g\n", H.nstep, H.t, dt);
      char out[64] = "";
      if (time_output == 0 && H.noutput > 0) {
        if ((H.nstep % H.noutput) == 0) {
          vtkfile(++nvtk, H, &Hv);
          snprintf(out, sizeof(out), " [%04d]", nvtk);
        }
      } else if (time_output == 1 && H.t >= next_output_time) {
        vtkfile(++nvtk, H, &Hv);
        next_output_time += H.dtoutput;
        snprintf(out, sizeof(out), " [%04d]", nvtk);
      }
      if (time_image == 0 && nimage > 0) {
        if ((H.nstep % nimage) == 0) {
          DumpImage(nimg, &H, &Hv);
          snprintf(out + strlen(out), sizeof(out) - strlen(out), " <%04d>", nimg++);
        }
      } else if (time_image == 1 && H.t >= next_image_time) {
        DumpImage(nimg, &H, &Hv);
        next_image_time += dtimage;
        snprintf(out + strlen(out), sizeof(out) - strlen(out), " <%04d>", nimg++);
      }
      fprintf(stdout, "--> step=%4d, %12.5e, %10.5e %.3lf MC/s%s\n", H.nstep, H.t, dt, mcs, out);
      if (H.nstep > 5) {
        mcsavg += mcs;
        nmcs++;
        if (mcsmin > mcs) mcsmin = mcs;
        if (mcsmax < mcs) mcsmax = mcs;
        mcssig += mcs * mcs;
      }
    }
    if (native) free(chunk_dt);
    fflush(stdout);
    if (allotted >= 0 && (H.t < H.tend) && (H.nstep < H.nstepmax)) {
      const double remain = allotted - (now() - t_begin);
      if (remain < guard) {
        fprintf(stderr, " Hydro stops by time limit %g < %g\n", remain, guard);
        break;
      }
    }
  }
  const double elaps = now() - t_start;
  if (dtlog) fclose(dtlog);
  if (ckpt_path) {
    if (hydro_amd_checkpoint(ckpt_path, H, &Hv, dt) != 0) {
      fprintf(stderr, "hydro: %s\n", hydro_amd_last_error());
      return 1;
    }
    fprintf(stdout, "Hydro: checkpoint at step %d written to %s\n", H.nstep, ckpt_path);
  }
  if (dump_path && hydro_amd_dump(dump_path, H, &Hv) != 0)
    fprintf(stderr, "hydro: cannot write %s\n", dump_path);
  deallocate_work_space(H.nxyt, H, &Hw_godunov, &Hvw_godunov);
  compute_deltat_clean_mem(H, &Hw_deltat, &Hvw_deltat);
  hydro_finish(H, &Hv);
  fprintf(stdout, "Hydro ends in %.3lfs.\n", elaps);
  if (nmcs > 0) {
    mcsavg /= nmcs;
    mcssig = sqrt(fabs(mcssig / nmcs - mcsavg * mcsavg));
    fprintf(stdout, "Average MC/s: %lf min %lf, max %lf sig %lf\n", mcsavg, mcsmin, mcsmax,
            mcssig);
  }
  return 0;
}
