/*
 * Field dumps.
 *
 * vtkfile(): the reference's VTK output, byte for byte the format of
 * HydroC99_2DMpi/Src/vtkfile.c:199-400 for a single rank:
 *   Dep/NNNN/NN/Hydro_00000_SSSS.vtr   RectilinearGrid piece, ascii Float32
 *                                      coordinates, base64 Float32 cell data
 *                                      (a base64 4-byte length header, then
 *                                      the base64 payload, :282-299)
 *   Dep/NNNN/NN/Hydro_SSSS.pvtr        the parallel wrapper (:343-393)
 *   Hydro.pvd                          the time collection (:153-177)
 * where NNNNNN is the 6-digit step number split 4 + 2 (:179-197).
 *
 * hydro_amd_dump(): an FP64 binary dump (the VTK output is Float32 and too
 * coarse to check parity at 1e-12).  Layout, little endian:
 *   char magic[8] = "HYDROF64"; int32 nx, ny, nvar, nstep; double t, 0;
 *   double interior[nvar][ny][nx]
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/types.h>

#include "../../include/hydro_amd.h"

int hydro_amd_set_error(const char *fmt, ...); /* hydro_capi.cpp */

static char six_bits(unsigned b) {
  if (b < 26) return (char) ('A' + b);
  if (b < 52) return (char) ('a' + b - 26);
  if (b < 62) return (char) ('0' + b - 52);
  return b == 62 ? '+' : '/';
}

/* Standard base64 with '=' padding; caller frees. */
static char *base64(const unsigned char *d, size_t n) {
  size_t blocks = (n + 2) / 3;
  char *s = (char *) malloc(blocks * 4 + 1);
  if (!s) {
    fprintf(stderr, "vtkfile: out of memory\n");
    abort();
  }
  for (size_t b = 0; b < blocks; b++) {
    size_t k = b * 3;
    unsigned c0 = d[k];
    unsigned c1 = k + 1 < n ? d[k + 1] : 0;
    unsigned c2 = k + 2 < n ? d[k + 2] : 0;
    s[b * 4 + 0] = six_bits(c0 >> 2);
    s[b * 4 + 1] = six_bits(((c0 & 3) << 4) | (c1 >> 4));
    s[b * 4 + 2] = k + 1 < n ? six_bits(((c1 & 15) << 2) | (c2 >> 6)) : '=';
    s[b * 4 + 3] = k + 2 < n ? six_bits(c2 & 63) : '=';
  }
  s[blocks * 4] = 0;
  return s;
}

static void step_dir(char *dir, size_t cap, int step) {
  char tmp[16];
  snprintf(tmp, sizeof(tmp), "%06d", step);
  snprintf(dir, cap, "Dep");
  mkdir(dir, 0777);
  snprintf(dir, cap, "Dep/%.4s", tmp);
  mkdir(dir, 0777);
  snprintf(dir, cap, "Dep/%.4s/%.2s", tmp, tmp + 4);
  mkdir(dir, 0777);
}

static void write_pvd(int nout) {
  FILE *f = fopen("Hydro.pvd", "w");
  if (!f) return;
  fprintf(f, "<?xml version=\"1.0\"?>\n");
  fprintf(f, " <VTKFile type=\"Collection\" version=\"0.1\" byte_order=\"LittleEndian\">\n");
  fprintf(f, "  <Collection>\n");
  for (int i = 1; i <= nout; i++) {
    char tmp[16];
    snprintf(tmp, sizeof(tmp), "%06d", i);
    fprintf(f, "  <DataSet timestep=\"%d\" part=\"0\" file=\"Dep/%.4s/%.2s/Hydro_%04d.pvtr\"  "
               "name=\"Asmb:FRAME\"/>\n", i, tmp, tmp + 4, i);
  }
  fprintf(f, " </Collection>\n");
  fprintf(f, "</VTKFile>\n");
  fclose(f);
}

static const char *var_name[4] = {"varID", "varIU", "varIV", "varIP"};

void vtkfile(int step, const hydroparam_t H, hydrovar_t *Hv) {
  char dir[256], name[512];
  /* a grid created by hydro_init lives on the GPU; refresh the host copy */
  (void) hydro_amd_sync_host(Hv);
  step_dir(dir, sizeof(dir), step);
  snprintf(name, sizeof(name), "%s/Hydro_%05d_%04d.vtr", dir, H.mype, step);
  FILE *f = fopen(name, "w");
  if (!f) {
    fprintf(stderr, "Ouverture du fichier %s impossible\n", name);
    exit(1);
  }
  const int x0 = H.box[XMIN_BOX], x1 = H.box[XMAX_BOX];
  const int y0 = H.box[YMIN_BOX], y1 = H.box[YMAX_BOX];
  fprintf(f, "<?xml version=\"1.0\"?>\n");
  fprintf(f, "<VTKFile type=\"RectilinearGrid\" byte_order=\"LittleEndian\">\n");
  fprintf(f, " <RectilinearGrid WholeExtent=\" %d %d %d %d %d %d\">\n", x0, x1, y0, y1, 0, 1);
  fprintf(f, "  <Piece Extent=\" %d %d %d %d %d %d\" GhostLevel=\"0\">\n", x0, x1, y0, y1, 0, 1);
  fprintf(f, "   <Coordinates>\n");
  fprintf(f, "    <DataArray type=\"Float32\" format=\"ascii\" NumberOfComponents=\"1\">\n");
  for (int i = x0; i <= x1; i++) fprintf(f, "%f ", i * H.dx);
  fprintf(f, "\n");
  fprintf(f, "    </DataArray>\n");
  fprintf(f, "    <DataArray type=\"Float32\" format=\"ascii\" NumberOfComponents=\"1\">\n");
  for (int j = y0; j <= y1; j++) fprintf(f, "%f ", j * H.dx);
  fprintf(f, "\n");
  fprintf(f, "    </DataArray>\n");
  fprintf(f, "    <DataArray type=\"Float32\" format=\"ascii\" NumberOfComponents=\"1\">\n");
  fprintf(f, "%f %f\n", 0., 1. * H.dx);
  fprintf(f, "    </DataArray>\n");
  fprintf(f, "   </Coordinates>\n");
  fprintf(f, "   <CellData Scalars=\" varID varIU varIV varIP\">\n");
  const size_t nxt = (size_t) H.nxt, nyt = (size_t) H.nyt;
  float *buf = (float *) malloc(sizeof(float) * (size_t) H.nx * (size_t) H.ny + 16);
  if (!buf) {
    fprintf(stderr, "vtkfile: out of memory\n");
    abort();
  }
  for (int v = 0; v < 4; v++) {
    fprintf(f, "    <DataArray Name=\"%s\" type=\"Float32\" format=\"binary\" "
               "encoding=\"base64\" NumberOfComponents=\"1\">\n", var_name[v]);
    size_t p = 0;
    for (size_t j = 2; j < nyt - 2; j++)
      for (size_t i = 2; i < nxt - 2; i++) buf[p++] = (float) Hv->uold[i + nxt * (j + nyt * v)];
    /* the header is the payload byte count as a little-endian int32 */
    uint32_t nbytes = (uint32_t) (p * sizeof(float));
    char *h64 = base64((const unsigned char *) &nbytes, sizeof(nbytes));
    fwrite(h64, 1, strlen(h64), f);
    free(h64);
    char *d64 = base64((const unsigned char *) buf, p * sizeof(float));
    fwrite(d64, 1, strlen(d64), f);
    free(d64);
    fprintf(f, "    </DataArray>\n");
  }
  free(buf);
  fprintf(f, "   </CellData>\n");
  fprintf(f, "  </Piece>\n");
  fprintf(f, " </RectilinearGrid>\n");
  fprintf(f, "</VTKFile>\n");
  fclose(f);

  if (H.mype == 0) {
    snprintf(name, sizeof(name), "%s/Hydro_%04d.pvtr", dir, step);
    FILE *vf = fopen(name, "w");
    if (!vf) {
      fprintf(stderr, "Ouverture du fichier %s impossible\n", name);
      exit(1);
    }
    fprintf(vf, "<?xml version=\"1.0\"?>\n");
    fprintf(vf, "<VTKFile type=\"PRectilinearGrid\" byte_order=\"LittleEndian\">\n");
    fprintf(vf, "<PRectilinearGrid WholeExtent=\"0 %d 0 %d 0 %d\"  GhostLevel=\"0\" >\n",
            H.globnx, H.globny, 1);
    fprintf(vf, " <PCellData>\n");
    for (int v = 0; v < 4; v++)
      fprintf(vf, "  <PDataArray Name=\"%s\" type=\"Float32\" format=\"binary\" "
                  "encoding=\"base64\" NumberOfComponents=\"1\"/>\n", var_name[v]);
    fprintf(vf, " </PCellData>\n");
    fprintf(vf, " <PCoordinates>\n");
    for (int k = 0; k < 3; k++)
      fprintf(vf, "  <PDataArray type=\"Float32\" format=\"ascii\" NumberOfComponents=\"1\"/>\n");
    fprintf(vf, " </PCoordinates>\n");
    /* one rank: the piece is the whole domain */
    fprintf(vf, " <Piece Extent=\"%d %d %d %d %d %d\" Source=\"Hydro_%05d_%04d.vtr\"/>\n", 0,
            H.globnx, 0, H.globny, 0, 1, 0, step);
    fprintf(vf, "</PRectilinearGrid>\n");
    fprintf(vf, "</VTKFile>\n");
    fclose(vf);
    write_pvd(step);
  }
}

static int write_state(const char *path, const hydroparam_t H, hydrovar_t *Hv, double dt) {
  if (hydro_amd_sync_host(Hv) != 0) return -1;
  FILE *f = fopen(path, "wb");
  if (!f) return -1;
  const int32_t hdr[4] = {H.nx, H.ny, 4, H.nstep};
  const double td[2] = {H.t, dt};
  int ok = fwrite("HYDROF64", 1, 8, f) == 8 && fwrite(hdr, sizeof(int32_t), 4, f) == 4 &&
           fwrite(td, sizeof(double), 2, f) == 2;
  const size_t nxt = (size_t) H.nxt, nyt = (size_t) H.nyt;
  for (int v = 0; v < 4 && ok; v++)
    for (size_t j = 2; j < nyt - 2 && ok; j++)
      ok = fwrite(Hv->uold + 2 + nxt * (j + nyt * v), sizeof(double), (size_t) H.nx, f) ==
           (size_t) H.nx;
  return (fclose(f) == 0 && ok) ? 0 : -1;
}

int hydro_amd_dump(const char *path, const hydroparam_t H, hydrovar_t *Hv) {
  return write_state(path, H, Hv, 0.0);
}

int hydro_amd_checkpoint(const char *path, const hydroparam_t H, hydrovar_t *Hv, double dt) {
  char prev[4096];
  if (snprintf(prev, sizeof(prev), "%s.prev", path) < (int) sizeof(prev)) {
    struct stat st;
    if (stat(path, &st) == 0) (void) rename(path, prev);
  }
  if (write_state(path, H, Hv, dt) != 0) {
    hydro_amd_set_error("cannot write checkpoint %s", path);
    return -1;
  }
  return 0;
}

int hydro_amd_restore(const char *path, hydroparam_t *H, hydrovar_t *Hv, double *dt) {
  FILE *f = fopen(path, "rb");
  if (!f) {
    hydro_amd_set_error("cannot open checkpoint %s", path);
    return -1;
  }
  char magic[8];
  int32_t hdr[4];
  double td[2];
  int ok = fread(magic, 1, 8, f) == 8 && memcmp(magic, "HYDROF64", 8) == 0 &&
           fread(hdr, sizeof(int32_t), 4, f) == 4 && fread(td, sizeof(double), 2, f) == 2;
  if (!ok || hdr[0] != H->nx || hdr[1] != H->ny || hdr[2] != 4 || hdr[3] < 0) {
    fclose(f);
    hydro_amd_set_error("checkpoint %s does not match this %dx%d deck", path, H->nx, H->ny);
    return -1;
  }
  const size_t nxt = (size_t) H->nxt, nyt = (size_t) H->nyt;
  for (int v = 0; v < 4 && ok; v++)
    for (size_t j = 2; j < nyt - 2 && ok; j++)
      ok = fread(Hv->uold + 2 + nxt * (j + nyt * v), sizeof(double), (size_t) H->nx, f) ==
           (size_t) H->nx;
  fclose(f);
  if (!ok) {
    hydro_amd_set_error("checkpoint %s is truncated", path);
    return -1;
  }
  H->nstep = hdr[3];
  H->t = td[0];
  if (dt) *dt = td[1];
  return hydro_amd_sync_device(Hv);
}
