/*
 * gp_host.h -- host side of the drop-in: reads the serial driver's Input.in / Filling.in, builds the initial
 * state and the boundary index maps, evaluates the FP64 free-energy coefficients, and reads/writes the
 * driver's VTK files. CPU only (C99 + libm); produces the gp_params consumed by libgp_hip (gp_hip.h).
 *
 * Reference routines replaced (GPs = Grand_potential_Finite_difference_2D_serial):
 *   gp_read_input       <- reading_input_parameters  (GPs/functions/reading_input_parameters.h:1-517)
 *                          + read_boundary_conditions (GPs/functions/read_boundary_conditions.h:4-57)
 *   gp_property_matrices<- initialize_variables + init_propertymatrices
 *                          (GPs/functions/initialize_variables.h:4-20, GPs/functions/functionF_01.h:169-230)
 *   gp_fill_domain      <- fill_domain + fill_phase_* + fill_composition_cube
 *                          (GPs/functions/fill_domain.h:4-151, GPs/functions/filling.h:6-270)
 *   gp_write_vtk        <- writetofile_serial2D[_binary] (GPs/solverloop/file_writer.h:23-255)
 *   gp_read_vtk         <- readfromfile_serial2D[_binary] (GPs/solverloop/file_writer.h:47-70, 256-416)
 *   gp_write_h5 / gp_read_h5 <- writetofile_mpi2D_hdf5 / readfromfile_mpi2D_hdf5 (GPm/solverloop/file_writer.h)
 *   gp_write_xdmf       <- write_xdmf_xml (GPm/write_xdmf.c)
 *   gp_rank_block       <- the MPI driver's decomposition (GPm/solverloop/mpiinfo_xy.h:37-56, 323-326)
 *   gp_write_rank_vtk / gp_read_rank_vtk <- writetofile_mpi2D[_binary] / readfromfile_mpi2D[_binary]
 *                          (GPm/solverloop/file_writer.h:52-75, 119-142, 339-533, 731-869)
 *   gp_reconstruct      <- the reconstruct tool (GPm/reconstruct.c)
 *   gp_write_echo       <- the <infile>.out / <infile>.bd echo files
 *                          (GPs/functions/reading_input_parameters.h:310-516, read_boundary_conditions.h:195-207)
 */
#ifndef GP_HOST_H_
#define GP_HOST_H_

#include "gp_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define GP_NAMELEN 51

/* Everything Input.in sets. p holds what the solver loop reads; the rest drives the host. */
typedef struct gp_input {
  gp_params p;
  int32_t dimension;
  int64_t mesh_z;
  double dz;
  int64_t ntimesteps, savet, nsmooth, starttime, restart, track_progress;
  int32_t ascii;                 /* WRITEFORMAT = ASCII */
  int32_t writecomposition;
  int32_t isothermal;
  int32_t noise_phasefield;
  double amp_noise_phase;
  int32_t shift;
  int64_t shiftj;
  int32_t fn_f;                  /* Function_F */
  int32_t binary, ternary, dilute;
  double R, V, Tfill, temperature_scale, tilt_angle, rtheta;
  char phases[GP_MAXP][GP_NAMELEN];
  char components[GP_MAXNC + 1][GP_NAMELEN];
  int32_t max_length;            /* widest phase/component name, for the progress lines */
  /* thermodynamic tables, [a][b][k] */
  double ceq[GP_MAXP][GP_MAXP][GP_MAXNC];
  double cfill[GP_MAXP][GP_MAXP][GP_MAXNC];
  double slopes[GP_MAXP][GP_MAXP][GP_MAXNC];
  /* full 3x3 rotation matrices as the reference keeps them (after its in-place LU) */
  double rot3[GP_MAXP][GP_MAXP][3][3];
  double rot3_inv[GP_MAXP][GP_MAXP][3][3];
  /* boundary tables: [side 0 X-,1 X+,2 Y+,3 Y-,4 Z+,5 Z-][field 0 phi,1 mu,2 T] */
  int32_t bc_type[6][3];
  int32_t bc_points[6][3][3];
  int32_t bc_proxy[6][3][3];
  int32_t bc_proxy_set[6][3];
  /* bookkeeping of which keys were seen (validation) */
  uint32_t seen;
  uint8_t seen_A[GP_MAXP], seen_D[GP_MAXP], seen_rot[GP_MAXP][GP_MAXP];
  double eigen_dummy;
  int32_t writehdf5;             /* WRITEHDF5 (the MPI driver's key): snapshots as DATA/<out>_<t>.h5 */
  int32_t reserved_h5;
} gp_input;

/* Parse Input.in (both passes: parameters, then BOUNDARY / BOUNDARY_VALUE lines), compute the property
 * matrices and rotation matrices, and fill in->p. Unknown keys are reported on stdout like the reference
 * ("Unrecongized parameter : ..."), unless quiet. Returns GP_OK or GP_E_ARG/GP_E_IO with err filled. */
int gp_read_input(const char *input_path, gp_input *in, int quiet, char *err, size_t errlen);

/* Check the parameters are in the supported set (2-D, F=1, Neumann/periodic sides, no noise; ISOTHERMAL = 0 needs
 * Tempgrady; Shiftj inside the allocated rows). */
int gp_validate_input(const gp_input *in, char *err, size_t errlen);

/* init_propertymatrices(T) on the host: cmu, B(T), C(T), Beq, dBbdT into in->p. */
void gp_property_matrices(gp_input *in);

/* Halo-inclusive sizes. */
int64_t gp_rows_x(const gp_params *p);
int64_t gp_rows_y(const gp_params *p);

/* fill_domain(argv) into halo-inclusive SoA arrays phi[N][rows], mu[NC][rows] (rows = rows_x*rows_y).
 * dphi may be NULL. Filling.in keys: the serial driver's FILLCUBE, FILLCYLINDER, FILLSPHERE, FILLELLIPSE
 * (GPs/functions/fill_domain.h) and the MPI driver's FILLCYLINDERRANDOM, FILLSPHERERANDOM, FILLVORONOI2D,
 * FILLCUBEPATTERN, FILLCUBERANDOM (GPm/functions/fill_domain.h:146-332). The drand48-driven ones start from the
 * POSIX initial state as the reference does; the GSL-driven ones seed MT19937 with GP_FILL_SEED when that
 * environment variable is set, else with time(0) as the reference does. */
int gp_fill_domain(const gp_input *in, const char *filling_path, double *phi, double *mu, double *dphi,
                   int quiet, char *err, size_t errlen);

/* Copy halos from the interior with the boundary maps (apply_boundary_conditions(0)). */
void gp_apply_bc_host(const gp_params *p, double *phi, double *mu);

/* DATA/<out>_<t>.vtk in the serial layout; binary unless in->ascii. T_rows (the temperature of every allocated
 * row, gp_get_frame or gp_temperature_rows) is read when ISOTHERMAL = 0: it fills the Temperature block and the
 * compositions; it may be NULL otherwise. The reader takes phi and mu; a Temperature block is left unread, since
 * the non-isothermal loop sets T from the gradient before the first step. */
int gp_write_vtk(const gp_input *in, const char *path, const double *phi, const double *mu, const double *T_rows);
int gp_read_vtk(const gp_input *in, const char *path, double *phi, double *mu);

/* Temperature of every allocated row (rows_y values) as the serial driver sets it: before the loop (t = 0,
 * with the boundary copies of T) or after step t, with the moving frame shifted by shift_offset rows
 * (GPs/microsim_gp.c:77-84, 108-111). Isothermal runs: T everywhere. */
void gp_temperature_rows(const gp_params *p, int64_t t, int64_t shift_offset, double *T_rows);

/* <stem>.out and <stem>.bd echo files. */
int gp_write_echo(const gp_input *in, const char *input_path);

/* Composition field sum_b c_b(mu) h_b(phi) of one cell (the Composition_<k> output), at the parameters' T or at T. */
double gp_composition(const gp_params *p, const double *phi_cell, const double *mu_cell, int k);
double gp_composition_at(const gp_params *p, const double *phi_cell, const double *mu_cell, int k, double T);

/* The MPI driver's HDF5 snapshot DATA/<out>_<t>.h5 (writetofile_mpi2D_hdf5 / write_cells_hdf5_2D_mpi,
 * GPm/solverloop/file_writer.h:76-95, 534-633): root-group datasets /<phase>, /Mu_<c>, /Composition_<c> and, when
 * ISOTHERMAL = 0, /T, each [rows_x][rows_y] binary64 over the halo-inclusive grid. Written in the HDF5 file format
 * (version 0 superblock, symbol-table root group, contiguous datasets) without the HDF5 library. T_rows as for
 * gp_write_vtk. The reader (readfromfile_mpi2D_hdf5, :97-117 and :637-725) takes the phi and mu datasets of such a
 * file into halo-inclusive SoA arrays. */
int gp_write_h5(const gp_input *in, const char *path, const double *phi, const double *mu, const double *T_rows);
int gp_read_h5(const gp_input *in, const char *path, double *phi, double *mu);

/* The XDMF description of an HDF5 snapshot (write_xdmf_xml, GPm/write_xdmf.c:66-146, 2-D form): h5_name is the
 * file name the XML points at, relative to the XML file. */
int gp_write_xdmf(const gp_input *in, const char *path, const char *h5_name);

/* Block of rank `rank` in the MPI driver's workers_x x workers_y decomposition of the halo-inclusive grid
 * (GPm/solverloop/mpiinfo_xy.h:37-56, 323-326): rows[X], rows[Y] and their offsets; rank r is at
 * (r / workers_y, r % workers_y). GP_E_ARG when a block would be empty. rows or offset may be NULL. */
int gp_rank_block(const gp_params *p, int workers_x, int workers_y, int rank, int64_t rows[2], int64_t offset[2]);

/* The MPI driver's per-rank snapshot <dir>/Processor_<r>/<out>_<t>.vtk for every r < workers_x * workers_y
 * (writetofile_mpi2D[_binary] / write_cells_vtk_2D_mpi[binary], GPm/solverloop/file_writer.h:52-75, 339-533),
 * creating the Processor_<r> folders; ASCII unless in->ascii is 0. T_rows as for gp_write_vtk. The reader
 * (readfromfile_mpi2D[_binary] / read_cells_vtk_2D_mpi[binary], :119-142, 731-869) takes the phi and mu of nranks such files, placed by each file's own
 * counts and offsets; every halo-inclusive cell must come from exactly one file. */
int gp_write_rank_vtk(const gp_input *in, const char *dir, const char *out, int64_t t, int workers_x, int workers_y,
                      const double *phi, const double *mu, const double *T_rows);
int gp_read_rank_vtk(const gp_input *in, const char *dir, const char *out, int64_t t, int nranks, double *phi,
                     double *mu);

/* The reconstruct tool for one time (GPm/reconstruct.c:88-275): joins the nranks per-rank files of time t into
 * <dir>/<out>_<t>.vtk in its ASCII or binary layout (:277-533). */
int gp_reconstruct(const gp_input *in, const char *dir, const char *out, int nranks, int64_t t);

/* The same for every cell of a halo-inclusive SoA state (phi [N][n], mu [NC][n]) into c [NC][n], n = ncells. */
void gp_composition_grid(const gp_params *p, const double *phi, const double *mu, double *c, int64_t ncells);

/* sizeof checks for foreign-function mirrors of these structs. */
size_t gp_sizeof_input(void);
size_t gp_sizeof_params(void);
size_t gp_sizeof_stats(void);

#ifdef __cplusplus
}
#endif
#endif /* GP_HOST_H_ */
