/*
 * write_xdmf -- the XDMF descriptions of the MPI layout's HDF5 snapshots, with the reference tool's command line
 * (GPm/write_xdmf.c:20-40):
 *
 *   ./write_xdmf Input.in Output t_start t_end
 *
 * For t = t_start, t_start + SAVET, ... <= t_end it writes DATA/Output_<t>.xmf describing Output_<t>.h5
 * (gp_write_xdmf). The driver writes the same file next to every snapshot it writes with WRITEHDF5 = 1.
 */
#include <stdio.h>
#include <stdlib.h>

#include "gp_host.h"

int main(int argc, char *argv[]) {
  if (argc < 5) {
    fprintf(stderr, "usage: %s Input.in Output t_start t_end\n", argv[0]);
    return 2;
  }
  static gp_input in;
  char err[512];
  if (gp_read_input(argv[1], &in, 1, err, sizeof err)) {
    fprintf(stderr, "write_xdmf: Input: %s\n", err);
    return 1;
  }
  const long t0 = atol(argv[3]), t1 = atol(argv[4]);
  if (in.savet <= 0) {
    fprintf(stderr, "write_xdmf: SAVET must be positive\n");
    return 1;
  }
  char xmf[1200], h5[1100];
  for (long t = t0; t <= t1; t += in.savet) {
    snprintf(xmf, sizeof xmf, "DATA/%s_%ld.xmf", argv[2], t);
    snprintf(h5, sizeof h5, "%s_%ld.h5", argv[2], t);
    if (gp_write_xdmf(&in, xmf, h5)) {
      fprintf(stderr, "write_xdmf: cannot write %s\n", xmf);
      return 1;
    }
  }
  return 0;
}
