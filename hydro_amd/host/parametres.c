/*
 * process_args: command line and .nml deck -> hydroparam_t.
 *
 * Behaviour follows HydroC99_2DMpi/Src/parametres.c:
 *  - defaults of default_values() (:64-109);
 *  - options --help, -v (prt++), -i deck, -c (FTI, unsupported) (:278-304);
 *  - deck lines split at the first '=', the key cut at its first blank,
 *    blank/tab prefixes skipped, unknown keys and '&SECTION' / '/' lines
 *    ignored (:112-138, :159-258); int keys read with %d, real keys with %lf,
 *    scheme one of muscl/plmde/collela;
 *  - single-rank geometry (:372-377) and the nxystep default/clamp (:430-436);
 *  - the run summary box (:438-451).
 * One key is added: `gamma` (the reference hard-wires 1.4).  The image
 * cadence keys of HydroC99_2Domp (`nimage` %d, `dtimage` %lf,
 * HydroC99_2Domp/Src/parametres.c:165-166, 202-203) are read too and kept
 * outside hydroparam_t, which has no fields for them.
 */
#include <limits.h>
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/hydro_amd.h"
#include "../../include/hydro_surface.h"

static void usage(void) {
  fprintf(stderr, "options of hydro\n--help\n-i input\n-v :: to increase verbosity\n");
  exit(1);
}

static void set_defaults(hydroparam_t *H) {
  memset(H, 0, sizeof(*H));
  H->prt = 0;
  H->nx = 20;
  H->ny = 20;
  H->globnx = H->nx;
  H->globny = H->ny;
  H->nproc = 1;
  H->mype = 0;
  for (int k = 0; k < MAX_BOX; k++) H->box[k] = -1;
  H->nxystep = -1;
  H->nvar = 4;
  H->dx = 1.0;
  H->t = 0.0;
  H->nstep = 0;
  H->tend = 0.0;
  H->gamma = 1.4;
  H->courant_factor = 0.5;
  H->smallc = 1e-10;
  H->smallr = 1e-10;
  H->niter_riemann = 10;
  H->iorder = 2;
  H->slope_type = 1.;
  H->scheme = HSCHEME_MUSCL;
  H->boundary_right = 1;
  H->boundary_left = 1;
  H->boundary_up = 1;
  H->boundary_down = 1;
  H->noutput = 0;
  H->nstepmax = INT_MAX;
  H->dtoutput = 0.0;
  H->testCase = 0;
}

/* Splits "  key = value\n" in place; *val is NULL when the line has no '='. */
static void split_line(char *line, char **key, char **val) {
  char *nl = strchr(line, '\n');
  if (nl) *nl = 0;
  char *k = line;
  while (*k == ' ' || *k == '\t') k++;
  char *eq = strchr(line, '=');
  if (eq) {
    *eq = 0;
    *val = eq + 1;
  } else {
    *val = NULL;
  }
  size_t n = strcspn(k, " \t");
  k[n] = 0;
  *key = k;
}

struct int_key {
  const char *name;
  size_t off;
};
struct real_key {
  const char *name;
  size_t off;
};

static const struct int_key int_keys[] = {
    {"nstepmax", offsetof(hydroparam_t, nstepmax)},
    {"prt", offsetof(hydroparam_t, prt)},
    {"nx", offsetof(hydroparam_t, nx)},
    {"ny", offsetof(hydroparam_t, ny)},
    {"nxystep", offsetof(hydroparam_t, nxystep)},
    {"boundary_left", offsetof(hydroparam_t, boundary_left)},
    {"boundary_right", offsetof(hydroparam_t, boundary_right)},
    {"boundary_up", offsetof(hydroparam_t, boundary_up)},
    {"boundary_down", offsetof(hydroparam_t, boundary_down)},
    {"niter_riemann", offsetof(hydroparam_t, niter_riemann)},
    {"noutput", offsetof(hydroparam_t, noutput)},
    {"iorder", offsetof(hydroparam_t, iorder)},
    {"testcase", offsetof(hydroparam_t, testCase)},
};
static const struct real_key real_keys[] = {
    {"slope_type", offsetof(hydroparam_t, slope_type)},
    {"tend", offsetof(hydroparam_t, tend)},
    {"dx", offsetof(hydroparam_t, dx)},
    {"courant_factor", offsetof(hydroparam_t, courant_factor)},
    {"smallr", offsetof(hydroparam_t, smallr)},
    {"smallc", offsetof(hydroparam_t, smallc)},
    {"dtoutput", offsetof(hydroparam_t, dtoutput)},
    {"gamma", offsetof(hydroparam_t, gamma)},
};

/* nimage / dtimage of the last deck read (0 / 0.0: no images). */
static int g_nimage = 0;
static double g_dtimage = 0.0;

void hydro_amd_image_cadence(int *nimage, double *dtimage) {
  if (nimage) *nimage = g_nimage;
  if (dtimage) *dtimage = g_dtimage;
}

static void read_deck(const char *path, hydroparam_t *H) {
  FILE *fd = fopen(path, "r");
  char line[1024];
  if (!fd) {
    fprintf(stderr, "can't read input file\n");
    exit(1);
  }
  while (fgets(line, sizeof(line), fd) == line) {
    char *key, *val;
    int done = 0;
    split_line(line, &key, &val);
    if (!val) continue;
    for (size_t k = 0; !done && k < sizeof(int_keys) / sizeof(int_keys[0]); k++)
      if (strcmp(key, int_keys[k].name) == 0) {
        sscanf(val, "%d", (int *) ((char *) H + int_keys[k].off));
        done = 1;
      }
    for (size_t k = 0; !done && k < sizeof(real_keys) / sizeof(real_keys[0]); k++)
      if (strcmp(key, real_keys[k].name) == 0) {
        sscanf(val, "%lf", (double *) ((char *) H + real_keys[k].off));
        done = 1;
      }
    if (!done && strcmp(key, "nimage") == 0) {
      sscanf(val, "%d", &g_nimage);
      done = 1;
    }
    if (!done && strcmp(key, "dtimage") == 0) {
      sscanf(val, "%lf", &g_dtimage);
      done = 1;
    }
    if (!done && strcmp(key, "scheme") == 0) {
      if (strcmp(val, "muscl") == 0) H->scheme = HSCHEME_MUSCL;
      else if (strcmp(val, "plmde") == 0) H->scheme = HSCHEME_PLMDE;
      else if (strcmp(val, "collela") == 0) H->scheme = HSCHEME_COLLELA;
      else {
        fprintf(stderr, "Scheme name <%s> is unknown, should be one of [muscl,plmde,collela]\n",
                val);
        exit(1);
      }
    }
  }
  fclose(fd);
}

void process_args(int argc, char **argv, hydroparam_t *H) {
  const char *deck = NULL;
  set_defaults(H);
  for (int n = 1; n < argc; n++) {
    if (strcmp(argv[n], "--help") == 0) {
      usage();
    } else if (strcmp(argv[n], "-v") == 0) {
      H->prt++;
    } else if (strcmp(argv[n], "-i") == 0 && n + 1 < argc) {
      deck = argv[++n];
    } else if (strcmp(argv[n], "-c") == 0) {
      n++;
      fprintf(stderr, "FTI is not available\n");
    } else {
      fprintf(stderr, "Key %s is unkown\n", argv[n]);
    }
  }
  if (!deck) {
    fprintf(stderr, "Option -i is missing\n");
    exit(1);
  }
  g_nimage = 0;
  g_dtimage = 0.0;
  read_deck(deck, H);

  H->globnx = H->nx;
  H->globny = H->ny;
  H->box[XMIN_BOX] = 0;
  H->box[XMAX_BOX] = H->nx;
  H->box[YMIN_BOX] = 0;
  H->box[YMAX_BOX] = H->ny;
  if (H->nxystep == -1) {
    H->nxystep = (H->nx < H->ny) ? H->nx : H->ny;
  } else {
    if (H->nxystep > H->nx) H->nxystep = H->nx;
    if (H->nxystep > H->ny) H->nxystep = H->ny;
  }
  if (H->mype == 0) {
    printf("+-------------------+\n");
    printf("|GlobNx=%-7d     |\n", H->globnx);
    printf("|GlobNy=%-7d     |\n", H->globny);
    printf("|nx=%-7d         |\n", H->nx);
    printf("|ny=%-7d         |\n", H->ny);
    printf("|nxystep=%-7d    |\n", H->nxystep);
    printf("|tend=%-10.3f    |\n", H->tend);
    printf("|nstepmax=%-7d   |\n", H->nstepmax);
    printf("|noutput=%-7d    |\n", H->noutput);
    printf("|dtoutput=%-10.3f|\n", H->dtoutput);
    printf("+-------------------+\n");
  }
}
