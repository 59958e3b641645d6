/*
 * gp_ranks.c -- the MPI driver's per-rank VTK snapshots and the reconstruct tool's joined snapshot.
 *
 * Decomposition (GPm/solverloop/mpiinfo_xy.h:37-56, 323-326): the halo-inclusive grid, rows_x = MESH_X + 6 by
 * rows_y = MESH_Y + 6, is cut into workers_x x workers_y blocks. Rank r sits at (r / workers_y, r % workers_y);
 * along each axis a block has floor(rows / workers) rows, one more for the first rows % workers blocks, and its
 * offset is the sum of the blocks before it. The blocks tile the grid: the boundary ranks hold the global halo.
 *
 * Per-rank file <dir>/Processor_<r>/<out>_<t>.vtk (GPm/solverloop/file_writer.h:52-75, 339-533). The fields are
 * the phases, Mu_<c>, Composition_<c> and, for non-isothermal runs, the temperature, each over the rank's block in
 * x-major order.
 *   ASCII:  rows_X, rows_Y, offset_X, offset_Y, one per line; then one line "<global index> <value>" per cell,
 *           global index = (x + offset_X) * rows_y + (y + offset_Y), as "%ld %le\n" for the phases and Mu and
 *           "%ld %le \n" for Composition and the temperature; an empty line ends every field but the temperature.
 *   Binary: the four counts as native longs, then the values as native doubles, nothing between the fields.
 *
 * Joined file <dir>/<out>_<t>.vtk (reconstruct.c:88-275, writers :277-533): every field's interior cells in the
 * serial layout. ASCII: header DIMENSIONS MESH_Y 1 MESH_X and SPACING 1 1 1, values "%le\n" ("%le \n" for the
 * temperature). Binary: header DIMENSIONS MESH_Y MESH_X 1 and SPACING DELTA_X DELTA_Y DELTA_Z, big-endian values.
 * Both end every field but the temperature with a newline. The per-rank files are placed by their own counts and
 * offsets (ASCII: by each line's global index), so the join needs only the number of ranks.
 */
#include "gp_host.h"
#include "gp_snapshot.h"

#include <errno.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

int gp_rank_block(const gp_params *p, int workers_x, int workers_y, int rank, int64_t rows[2], int64_t offset[2]) {
  if (!p || workers_x < 1 || workers_y < 1 || rank < 0 || rank >= workers_x * workers_y) return GP_E_ARG;
  const int64_t n[2] = {gp_rows_x(p), gp_rows_y(p)};
  const int64_t w[2] = {workers_x, workers_y};
  const int64_t at[2] = {rank / workers_y, rank % workers_y};
  for (int d = 0; d < 2; d++) {
    const int64_t ave = n[d] / w[d], extra = n[d] % w[d];
    if (ave < 1) return GP_E_ARG;
    if (rows) rows[d] = ave + (at[d] < extra ? 1 : 0);
    if (offset) offset[d] = at[d] * ave + (at[d] < extra ? at[d] : extra);
  }
  return GP_OK;
}

static void rank_path(char *buf, size_t len, const char *dir, int rank, const char *out, int64_t t) {
  snprintf(buf, len, "%s/Processor_%d/%s_%ld.vtk", dir, rank, out, (long)t);
}

int gp_write_rank_vtk(const gp_input *in, const char *dir, const char *out, int64_t t, int workers_x, int workers_y,
                      const double *phi, const double *mu, const double *T_rows) {
  const gp_params *p = &in->p;
  const int N = p->nphases, NC = p->ncomp - 1;
  const int thermal = p->tempgrady != 0;
  if (thermal && !T_rows) return GP_E_ARG;
  if (gp_rank_block(p, workers_x, workers_y, 0, NULL, NULL)) return GP_E_ARG;
  const int64_t ry = gp_rows_y(p);
  gp_field f[GP_SNAPSHOT_MAXF];
  int nf = 0, rc = gp_snapshot_fields(in, phi, mu, T_rows, f, &nf);
  if (rc) return rc;
  char path[4096];
  for (int r = 0; r < workers_x * workers_y && rc == GP_OK; r++) {
    int64_t rows[2], off[2];
    gp_rank_block(p, workers_x, workers_y, r, rows, off);
    snprintf(path, sizeof path, "%s/Processor_%d", dir, r);
    if (mkdir(path, 0777) != 0 && errno != EEXIST) {
      rc = GP_E_IO;
      break;
    }
    rank_path(path, sizeof path, dir, r, out, t);
    FILE *fp = fopen(path, in->ascii ? "w" : "wb");
    if (!fp) {
      rc = GP_E_IO;
      break;
    }
    if (in->ascii) {
      fprintf(fp, "%ld\n%ld\n%ld\n%ld\n", (long)rows[0], (long)rows[1], (long)off[0], (long)off[1]);
    } else {
      const int64_t hdr[4] = {rows[0], rows[1], off[0], off[1]};
      fwrite(hdr, sizeof(int64_t), 4, fp);
    }
    for (int k = 0; k < nf; k++) {
      const int temp = thermal && k == nf - 1;
      const char *fmt = k < N + NC ? "%ld %le\n" : "%ld %le \n";
      for (int64_t x = 0; x < rows[0]; x++) {
        const int64_t g0 = (x + off[0]) * ry + off[1];
        if (in->ascii) {
          for (int64_t y = 0; y < rows[1]; y++) fprintf(fp, fmt, (long)(g0 + y), f[k].data[g0 + y]);
        } else {
          fwrite(f[k].data + g0, sizeof(double), (size_t)rows[1], fp);
        }
      }
      if (in->ascii && !temp) fprintf(fp, "\n");
    }
    if (ferror(fp)) rc = GP_E_IO;
    if (fclose(fp) != 0) rc = GP_E_IO;
  }
  gp_snapshot_free(f, nf);
  return rc;
}

/* Reads the first nread fields of one per-rank file into dst[k] (halo-inclusive arrays, rows_x * rows_y) and marks
 * its block in covered; a cell already covered, a block outside the grid or a line whose global index is not the
 * cell's is an error. */
static int read_rank_file(const gp_input *in, const char *path, int nread, double *const *dst, unsigned char *covered) {
  const gp_params *p = &in->p;
  const int64_t rx = gp_rows_x(p), ry = gp_rows_y(p);
  FILE *fp = fopen(path, in->ascii ? "r" : "rb");
  if (!fp) return GP_E_IO;
  int64_t hdr[4] = {0, 0, 0, 0};
  int ok = 1;
  if (in->ascii) {
    for (int i = 0; i < 4 && ok; i++) {
      long v;
      ok = fscanf(fp, "%ld", &v) == 1;
      hdr[i] = v;
    }
  } else {
    ok = fread(hdr, sizeof(int64_t), 4, fp) == 4;
  }
  const int64_t rows[2] = {hdr[0], hdr[1]}, off[2] = {hdr[2], hdr[3]};
  if (ok) ok = rows[0] >= 1 && rows[1] >= 1 && off[0] >= 0 && off[1] >= 0 && off[0] + rows[0] <= rx &&
               off[1] + rows[1] <= ry;
  for (int64_t x = 0; x < rows[0] && ok; x++)
    for (int64_t y = 0; y < rows[1] && ok; y++) {
      const int64_t g = (x + off[0]) * ry + (y + off[1]);
      if (covered[g]) ok = 0;
      covered[g] = 1;
    }
  for (int k = 0; k < nread && ok; k++)
    for (int64_t x = 0; x < rows[0] && ok; x++) {
      const int64_t g0 = (x + off[0]) * ry + off[1];
      if (in->ascii) {
        for (int64_t y = 0; y < rows[1] && ok; y++) {
          long gi;
          double v;
          ok = fscanf(fp, "%ld %le", &gi, &v) == 2 && gi == g0 + y;
          if (ok) dst[k][g0 + y] = v;
        }
      } else {
        ok = fread(dst[k] + g0, sizeof(double), (size_t)rows[1], fp) == (size_t)rows[1];
      }
    }
  fclose(fp);
  return ok ? GP_OK : GP_E_IO;
}

/* Reads <dir>/Processor_<r>/<out>_<t>.vtk for r < nranks; every cell of the grid must come from exactly one file. */
static int read_ranks(const gp_input *in, const char *dir, const char *out, int64_t t, int nranks, int nread,
                      double *const *dst) {
  const gp_params *p = &in->p;
  const int64_t n = gp_rows_x(p) * gp_rows_y(p);
  if (nranks < 1) return GP_E_ARG;
  unsigned char *covered = (unsigned char *)calloc((size_t)n, 1);
  if (!covered) return GP_E_OOM;
  char path[4096];
  int rc = GP_OK;
  for (int r = 0; r < nranks && rc == GP_OK; r++) {
    rank_path(path, sizeof path, dir, r, out, t);
    rc = read_rank_file(in, path, nread, dst, covered);
  }
  for (int64_t i = 0; i < n && rc == GP_OK; i++)
    if (!covered[i]) rc = GP_E_IO;
  free(covered);
  return rc;
}

int gp_read_rank_vtk(const gp_input *in, const char *dir, const char *out, int64_t t, int nranks, double *phi,
                     double *mu) {
  const gp_params *p = &in->p;
  const int N = p->nphases, NC = p->ncomp - 1;
  const int64_t n = gp_rows_x(p) * gp_rows_y(p);
  double *dst[GP_MAXP + GP_MAXNC];
  for (int a = 0; a < N; a++) dst[a] = phi + (int64_t)a * n;
  for (int c = 0; c < NC; c++) dst[N + c] = mu + (int64_t)c * n;
  return read_ranks(in, dir, out, t, nranks, N + NC, dst);
}

static void put_be(FILE *fp, double v) {
  unsigned char b[8], o[8];
  memcpy(b, &v, 8);
  for (int i = 0; i < 8; i++) o[i] = b[7 - i];
  fwrite(o, 1, 8, fp);
}

int gp_reconstruct(const gp_input *in, const char *dir, const char *out, int nranks, int64_t t) {
  const gp_params *p = &in->p;
  const int N = p->nphases, NC = p->ncomp - 1;
  const int thermal = p->tempgrady != 0;
  const int nf = N + 2 * NC + thermal;
  const int64_t ry = gp_rows_y(p), n = gp_rows_x(p) * ry;
  double *buf = (double *)malloc(sizeof(double) * (size_t)nf * (size_t)n);
  if (!buf) return GP_E_OOM;
  double *dst[GP_SNAPSHOT_MAXF];
  for (int k = 0; k < nf; k++) dst[k] = buf + (int64_t)k * n;
  int rc = read_ranks(in, dir, out, t, nranks, nf, dst);
  if (rc) {
    free(buf);
    return rc;
  }
  char path[4096];
  snprintf(path, sizeof path, "%s/%s_%ld.vtk", dir, out, (long)t);
  FILE *fp = fopen(path, in->ascii ? "w" : "wb");
  if (!fp) {
    free(buf);
    return GP_E_IO;
  }
  fprintf(fp, "# vtk DataFile Version 3.0\n");
  fprintf(fp, "Microsim_fields\n");
  fprintf(fp, in->ascii ? "ASCII\n" : "BINARY\n");
  fprintf(fp, "DATASET STRUCTURED_POINTS\n");
  if (in->ascii) {
    fprintf(fp, "DIMENSIONS %ld %ld %ld\n", (long)p->mesh_y, 1L, (long)p->mesh_x);
    fprintf(fp, "ORIGIN 0 0 0\n");
    fprintf(fp, "SPACING %le %le %le\n", 1.0, 1.0, 1.0);
  } else {
    fprintf(fp, "DIMENSIONS %ld %ld %ld\n", (long)p->mesh_y, (long)p->mesh_x, 1L);
    fprintf(fp, "ORIGIN 0 0 0\n");
    fprintf(fp, "SPACING %le %le %le\n", p->dx, p->dy, in->dz);
  }
  fprintf(fp, "POINT_DATA %ld\n", (long)p->mesh_x * (long)p->mesh_y);
  for (int k = 0; k < nf; k++) {
    const int temp = thermal && k == nf - 1;
    if (k < N) fprintf(fp, "SCALARS %s double 1\n", in->phases[k]);
    else if (k < N + NC) fprintf(fp, "SCALARS Mu_%s double 1\n", in->components[k - N]);
    else if (!temp) fprintf(fp, "SCALARS Composition_%s double 1\n", in->components[k - N - NC]);
    else fprintf(fp, "SCALARS Temperature double 1\n");
    fprintf(fp, "LOOKUP_TABLE default\n");
    for (int64_t x = 3; x <= p->mesh_x + 2; x++)
      for (int64_t y = 3; y <= p->mesh_y + 2; y++) {
        const double v = dst[k][x * ry + y];
        if (in->ascii) fprintf(fp, temp ? "%le \n" : "%le\n", v);
        else put_be(fp, v);
      }
    if (!temp) fprintf(fp, "\n");
  }
  if (ferror(fp)) rc = GP_E_IO;
  if (fclose(fp) != 0) rc = GP_E_IO;
  free(buf);
  return rc;
}
