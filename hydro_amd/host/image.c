/*
 * DumpImage(): the density image of HydroC99_2Domp (Src/Image.c:385-396 with
 * the PPM writer of :291-347; the reference's default build, PNG=N in its
 * Makefile, writes PPM).  For one rank the reference fills grey pixels
 * (uint8_t) fabs(255 * rho / max rho) sampled every `shrink` interior cells
 * (:111-166) and writes
 *     "P6\n<w> <h>\n255\n", w*h RGB triples, "#EOF\n"
 * to IMAGE_%06d.ppm, prefixed by $HYDROC_IMG_PREFIX when that is set.
 * The pixels come from hydro_amd_grid_image(), which computes them on the
 * GPU from the resident state.
 */
#include <stdio.h>
#include <stdlib.h>

#include "../../include/hydro_amd.h"

#define IMAGE_SHRINK_SIZE 1000 /* HydroC99_2Domp/Src/parametres.c:68 */

void DumpImage(int n, hydroparam_t *H, hydrovar_t *Hv) {
  (void) H;
  hydro_amd_grid *g = hydro_amd_surface_grid(Hv);
  int w = 0, h = 0;
  if (!g || hydro_amd_grid_image(g, IMAGE_SHRINK_SIZE, NULL, 0, &w, &h) != 0) {
    fprintf(stderr, "DumpImage: %s\n", g ? hydro_amd_last_error() : "uold was not created by hydro_init");
    abort();
  }
  const size_t bytes = (size_t) 3 * (size_t) w * (size_t) h;
  unsigned char *rgb = (unsigned char *) malloc(bytes ? bytes : 1);
  if (!rgb || hydro_amd_grid_image(g, IMAGE_SHRINK_SIZE, rgb, (long long) bytes, &w, &h) != 0) {
    fprintf(stderr, "DumpImage: %s\n", rgb ? hydro_amd_last_error() : "out of memory");
    abort();
  }
  char name[256], path[1024];
  snprintf(name, sizeof(name), "IMAGE_%06d.ppm", n);
  const char *prefix = getenv("HYDROC_IMG_PREFIX");
  snprintf(path, sizeof(path), "%s%s", prefix ? prefix : "", name);
  FILE *f = fopen(path, "w");
  if (!f) {
    fprintf(stderr, "File %s could not be opened for writing\n", path);
    abort();
  }
  fprintf(f, "P6\n%d %d\n255\n", w, h);
  if (bytes) fwrite(rgb, 1, bytes, f);
  fprintf(f, "#EOF\n");
  fclose(f);
  free(rgb);
}
