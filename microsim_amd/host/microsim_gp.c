/*
 * microsim_gp -- drop-in driver for the serial grand-potential solver with the solver loop on the GPU.
 *
 *   ./microsim_gp Input.in Filling.in Output            (serial form, GPs/microsim_gp.c)
 *   ./microsim_gp Input.in Filling.in Output wx wy      (MPI form, GPm/microsim_gp.c: one GPU holds the whole
 *                                                        domain; wx x wy sets the layout of the per-rank files)
 *
 * Same sequence as the reference main (GPs = Grand_potential_Finite_difference_2D_serial, GPs/microsim_gp.c:41-151):
 * mkdir DATA, read Input.in (+ <stem>.out / <stem>.bd echoes), fill the domain or read the restart snapshot,
 * apply the boundary conditions, write t = 0, NSMOOTH-1 smoothing steps, NTIMESTEPS steps with a snapshot every
 * SAVET (plus the DATA/shift.dat line) and the progress block every TRACK_PROGRESS. The loop body
 * (GPs/microsim_gp.c:103-106) is gp_step; steps between two output events are issued as one batch.
 *
 * The MPI driver's snapshots: WRITEHDF5 = 1 selects its HDF5 files (GPm/microsim_gp.c:202-210, 368-377) with their
 * XDMF descriptions; in the MPI form without WRITEHDF5 each snapshot is its per-rank VTK files
 * DATA/Processor_<r>/<out>_<t>.vtk for the wx x wy decomposition (GPm/microsim_gp.c:113-118, 368-376), which the
 * reconstruct tool joins. With either, the file at STARTTIME is written before the smoothing steps and again after
 * them (GPm/microsim_gp.c:248-285), and a restart reads that driver's files (:201-210).
 *
 * Non-isothermal runs (ISOTHERMAL = 0 with Tempgrady) and the moving frame (Shift = 1) follow the same main: on a
 * restart the frame position is read back from DATA/shift.dat and the gradient offset moved by the elapsed time and
 * that position (GPs/microsim_gp.c:56-74); every snapshot carries the Temperature block and its DATA/shift.dat line
 * the frame offset.
 *
 * Environment: GP_PRECISION=fp32c selects the compensated FP32 mode, fp32 the plain FP32 mode (default fp64,
 * bit-compatible; the non-isothermal and moving-frame runs are FP64 only).
 * Differences from the reference, by construction: the progress statistics are taken over interior cells
 * (the reference's sums also visit halo cells holding stale values), and unsupported inputs (noise,
 * Dirichlet/free sides) stop with an error instead of running.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include "gp_hip.h"
#include "gp_host.h"

static void die(const char *what, const char *msg) {
  fprintf(stderr, "microsim_gp: %s: %s\n", what, msg);
  exit(1);
}

/* DATA/<out>_<t>.vtk; with WRITEHDF5 = 1 the MPI driver's DATA/<out>_<t>.h5 plus its XDMF description
 * DATA/<out>_<t>.xmf (the reference writes the latter with its separate write_xdmf tool); else, with a wx x wy
 * layout (wx > 0), the MPI driver's per-rank files */
static void write_snapshot(const gp_input *in, const char *out, int64_t t, const double *phi, const double *mu,
                           const double *T_rows, int wx, int wy) {
  char name[1200], h5[1100];
  if (!in->writehdf5 && wx > 0) {
    if (gp_write_rank_vtk(in, "DATA", out, t, wx, wy, phi, mu, T_rows)) die("write", "DATA/Processor_*");
    return;
  }
  if (!in->writehdf5) {
    snprintf(name, sizeof name, "DATA/%s_%ld.vtk", out, (long)t);
    if (gp_write_vtk(in, name, phi, mu, T_rows)) die("write", name);
    return;
  }
  snprintf(h5, sizeof h5, "%s_%ld.h5", out, (long)t);
  snprintf(name, sizeof name, "DATA/%s", h5);
  if (gp_write_h5(in, name, phi, mu, T_rows)) die("write", name);
  snprintf(name, sizeof name, "DATA/%s_%ld.xmf", out, (long)t);
  if (gp_write_xdmf(in, name, h5)) die("write", name);
}

int main(int argc, char *argv[]) {
  if (argc < 4) {
    fprintf(stderr, "usage: %s Input.in Filling.in Output [workers_x workers_y]\n", argv[0]);
    return 2;
  }
  mkdir("DATA", 0777);
  static gp_input in;
  char err[512];
  if (gp_read_input(argv[1], &in, 0, err, sizeof err)) die("Input", err);
  if (gp_validate_input(&in, err, sizeof err)) die("Input", err);
  gp_write_echo(&in, argv[1]);
  const gp_params *p = &in.p;
  int wx = 0, wy = 0;
  if (argc >= 6) {
    wx = atoi(argv[4]);
    wy = atoi(argv[5]);
    if (gp_rank_block(p, wx, wy, 0, NULL, NULL)) die("workers", "the decomposition leaves a worker without rows");
    printf("workers %d x %d requested; the domain runs on one GPU\n", wx, wy);
  }
  /* the MPI driver's main: its snapshot files, its STARTTIME writes */
  const int mpi_main = wx > 0 || in.writehdf5;
  const int N = p->nphases, NC = p->ncomp - 1;
  const int64_t n = gp_rows_x(p) * gp_rows_y(p);
  double *phi = calloc((size_t)N * n, sizeof(double));
  double *mu = calloc((size_t)NC * n, sizeof(double));
  double *T_rows = calloc((size_t)gp_rows_y(p), sizeof(double));
  if (!phi || !mu || !T_rows) die("host", "out of memory");

  char name[1200];
  int64_t shift_position = 0;
  if (in.starttime == 0 && in.restart == 0) {
    if (gp_fill_domain(&in, argv[2], phi, mu, NULL, 0, err, sizeof err)) die("Filling", err);
  } else {
    snprintf(name, sizeof name, "DATA/%s_%ld.%s", argv[3], (long)in.starttime, in.writehdf5 ? "h5" : "vtk");
    if (in.writehdf5) {
      if (gp_read_h5(&in, name, phi, mu)) die("restart", name);
    } else if (wx > 0) {
      if (gp_read_rank_vtk(&in, "DATA", argv[3], in.starttime, wx * wy, phi, mu)) die("restart", "DATA/Processor_*");
    } else if (gp_read_vtk(&in, name, phi, mu)) {
      die("restart", name);
    }
    if (in.shift) {
      /* the position of line STARTTIME/SAVET (0-based) of DATA/shift.dat; the last line read when it is shorter */
      FILE *fp = fopen("DATA/shift.dat", "r");
      if (!fp) die("restart", "DATA/shift.dat");
      long tf = 0, pos = 0;
      for (int64_t it = 0; it <= in.starttime / (in.savet > 0 ? in.savet : 1); it++)
        if (fscanf(fp, "%ld %ld\n", &tf, &pos) != 2) break;
      fclose(fp);
      shift_position = pos;
    }
    if (p->tempgrady) {
      if (in.shift)
        in.p.tg_offset = (in.p.tg_offset) + floor((in.p.tg_velocity) * (in.starttime * in.p.dt)) -
                         (double)shift_position * in.p.dy;
      else
        in.p.tg_offset = (in.p.tg_offset) + floor((in.p.tg_velocity) * (in.starttime * in.p.dt));
    }
  }
  gp_apply_bc_host(p, phi, mu);
  gp_temperature_rows(p, 0, 0, T_rows);
  if (mpi_main) write_snapshot(&in, argv[3], in.starttime, phi, mu, T_rows, wx, wy);
  else if (in.restart == 0 || in.starttime == 0) write_snapshot(&in, argv[3], 0, phi, mu, T_rows, wx, wy);

  int precision = GP_FP64;
  const char *pe = getenv("GP_PRECISION");
  if (pe && (strcmp(pe, "fp32") == 0 || strcmp(pe, "FP32") == 0)) precision = GP_FP32;
  if (pe && (strcmp(pe, "fp32c") == 0 || strcmp(pe, "FP32C") == 0)) precision = GP_FP32C;
  gp_handle *h = NULL;
  if (gp_create(p, precision, &h)) die("gp_create", gp_last_error(NULL));
  if (gp_upload(h, phi, mu)) die("gp_upload", gp_last_error(h));
  if (in.nsmooth > 1) {
    if (gp_smooth(h, in.nsmooth - 1)) die("gp_smooth", gp_last_error(h));
    if (mpi_main) {
      if (gp_download(h, phi, mu, NULL)) die("gp_download", gp_last_error(h));
      write_snapshot(&in, argv[3], in.starttime, phi, mu, T_rows, wx, wy);
    }
  }

  const int64_t saveT = in.savet > 0 ? in.savet : in.ntimesteps + 1;
  const int64_t track = in.track_progress > 0 ? in.track_progress : in.ntimesteps + 1;
  int64_t t = 0;
  while (t < in.ntimesteps) {
    int64_t next = in.ntimesteps;
    int64_t ns = (t / saveT + 1) * saveT, nt = (t / track + 1) * track;
    if (ns < next) next = ns;
    if (nt < next) next = nt;
    const int report = (next % track == 0);
    gp_stats st; /* every step accumulates the cumulative max/min, as the reference loop does */
    int rc = gp_step(h, next - t, &st);
    if (rc) die("gp_step", gp_last_error(h));
    t = next;
    if (t % saveT == 0) {
      int64_t shift_offset = 0;
      if (gp_download(h, phi, mu, NULL)) die("gp_download", gp_last_error(h));
      if (gp_get_frame(h, &shift_offset, T_rows)) die("gp_get_frame", gp_last_error(h));
      write_snapshot(&in, argv[3], t + in.starttime, phi, mu, T_rows, wx, wy);
      FILE *fp = fopen("DATA/shift.dat", "a");
      if (fp) {
        fprintf(fp, "%ld %ld\n", (long)(t + in.starttime), (long)(shift_offset + shift_position));
        fclose(fp);
      }
    }
    if (report) {
      fprintf(stdout, "Time=%le\n", (double)t * p->dt);
      for (int b = 0; b < N; b++)
        fprintf(stdout, "%*s, Max = %le, Min = %le, Relative_Change=%le\n", in.max_length, in.phases[b],
                st.phi_max[b], st.phi_min[b], sqrt(st.rel_change_phi[b]));
      for (int k = 0; k < NC; k++)
        fprintf(stdout, "%*s, Max = %le, Min = %le, Relative_Change=%le\n", in.max_length, in.components[k],
                st.mu_max[k], st.mu_min[k], sqrt(st.rel_change_mu[k]));
      fprintf(stdout, "\n");
    }
  }
  gp_destroy(h);
  free(phi);
  free(mu);
  free(T_rows);
  return 0;
}
