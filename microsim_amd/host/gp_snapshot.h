/*
 * gp_snapshot.h -- internal to libgp_host: the fields of the MPI driver's snapshots (its HDF5 files and its
 * per-rank VTK files), in the order it writes them (GPm/solverloop/file_writer.h:339-533, 534-633): the phases,
 * Mu_<component>, Composition_<component> and, for non-isothermal runs, the temperature, each over the
 * halo-inclusive grid [rows_x][rows_y].
 */
#ifndef GP_SNAPSHOT_H_
#define GP_SNAPSHOT_H_

#include "gp_host.h"

#define GP_SNAPSHOT_MAXF (GP_MAXP + 2 * GP_MAXNC + 1)

typedef struct {
  char name[GP_NAMELEN + 16];
  const double *data; /* rows_x * rows_y values */
  double *tmp;        /* owned by the field list (freed by gp_snapshot_free), or NULL */
} gp_field;

/* Fills f[0 .. *nf). The composition is sum_b c_b(mu) h_b(phi) of each cell, at the row's temperature (T_rows)
 * when non-isothermal; the temperature field is named "T" (the HDF5 dataset name). */
__attribute__((visibility("hidden"))) int gp_snapshot_fields(const gp_input *in, const double *phi,
                                                             const double *mu, const double *T_rows, gp_field *f,
                                                             int *nf);
__attribute__((visibility("hidden"))) void gp_snapshot_free(gp_field *f, int nf);

#endif
