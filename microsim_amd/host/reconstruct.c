/*
 * reconstruct -- joins the MPI layout's per-rank snapshots into serial-layout VTK files, with the reference tool's
 * command line (GPm/reconstruct.c:88-130):
 *
 *   ./reconstruct Input.in Output numworkers t_start t_end
 *
 * For t = t_start, t_start + SAVET, ... <= t_end it reads DATA/Processor_<r>/Output_<t>.vtk for r < numworkers and
 * writes DATA/Output_<t>.vtk (gp_reconstruct). The file format (ASCII or binary) is Input.in's WRITEFORMAT.
 */
#include <stdio.h>
#include <stdlib.h>

#include "gp_host.h"

int main(int argc, char *argv[]) {
  if (argc < 6) {
    fprintf(stderr, "usage: %s Input.in Output numworkers t_start t_end\n", argv[0]);
    return 2;
  }
  static gp_input in;
  char err[512];
  if (gp_read_input(argv[1], &in, 1, err, sizeof err)) {
    fprintf(stderr, "reconstruct: Input: %s\n", err);
    return 1;
  }
  const int nranks = atoi(argv[3]);
  const long t0 = atol(argv[4]), t1 = atol(argv[5]);
  if (in.savet <= 0) {
    fprintf(stderr, "reconstruct: SAVET must be positive\n");
    return 1;
  }
  for (long t = t0; t <= t1; t += in.savet) {
    int rc = gp_reconstruct(&in, "DATA", argv[2], nranks, t);
    if (rc) {
      fprintf(stderr, "reconstruct: DATA/Processor_*/%s_%ld.vtk: error %d\n", argv[2], t, rc);
      return 1;
    }
  }
  return 0;
}
