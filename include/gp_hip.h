/*
 * gp_hip.h -- C ABI of libgp_hip, the MI355X solver-loop library for MicroSim's 2-D grand-potential (GP)
 * finite-difference model.
 *
 * The library replaces the serial driver's solver-loop routines (reference paths relative to the MicroSim tree,
 * GPs = Grand_potential_Finite_difference_2D_serial):
 *
 *   gp_step    <- solverloop_phasefield(start,end); apply_boundary_conditions(0);
 *                 solverloop_concentration(start,end); apply_boundary_conditions(0);
 *                 (GPs/microsim_gp.c:102-106, GPs/solverloop/solverloop.h:21-71,
 *                  GPs/solverloop/boundary_serial.h:176-203)
 *   gp_smooth  <- smooth(start,end); apply_boundary_conditions(0);      (GPs/microsim_gp.c:97-100,
 *                 GPs/solverloop/solverloop.h:72-90)
 *   gp_upload / gp_download
 *              <- the AoS array `struct fields *gridinfo` the host owns (GPs/functions/global_vars.h:327-332),
 *                 indexed x*layer_size + y with 3 halo cells per side (GPs/solverloop/serialinfo_xy.h:7-46)
 *   gp_params  <- every global the loop reads (GPs/functions/global_vars.h, set by
 *                 GPs/functions/reading_input_parameters.h and init_propertymatrices in
 *                 GPs/functions/functionF_01.h:169-230)
 *   gp_stats   <- struct max_min global_max_min (GPs/functions/global_vars.h:130-139), filled in
 *                 GPs/solverloop/calculate_divergence_concentration.h:128-151
 *
 * Conventions
 *   - Plain C, no exceptions cross the boundary. Every int-returning call returns GP_OK (0) or a negative
 *     GP_E* code; gp_last_error() gives the message for the last failing call on that handle.
 *   - Host arrays use the reference's halo-inclusive layout: field-major SoA, each component an array of
 *     rows_x*rows_y doubles (rows = mesh + 6), index x*rows_y + y (y fastest). Interior cells are x,y in
 *     [3, mesh+2]. The caller owns host buffers; the library owns its device buffers.
 *   - One HIP stream per handle. A handle is not thread-safe; distinct handles are independent.
 *   - Boundary conditions: halo cells are copies of interior cells given by the index maps in gp_params
 *     (the reference's proxy tables, GPs/functions/utility_functions.h:471-650). Only Neumann (1) and
 *     periodic (3) sides exist in these maps; the reference's Dirichlet/free sides copy nothing
 *     (GPs/solverloop/boundary_serial.h:15) and are rejected by the host with an error.
 */
#ifndef GP_HIP_H_
#define GP_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GP_MAXP 6            /* maximum NUMPHASES */
#define GP_MAXNC 4           /* maximum NUMCOMPONENTS-1 */
#define GP_HALO 3            /* halo width of the reference layout */

/* status codes */
#define GP_OK 0
#define GP_E_ARG (-1)        /* invalid argument / unsupported parameter combination */
#define GP_E_HIP (-2)        /* HIP runtime error */
#define GP_E_OOM (-3)        /* device allocation failed */
#define GP_E_NAN (-4)        /* non-finite value detected by gp_step's checks */
#define GP_E_IO (-5)         /* file I/O (host library) */
#define GP_E_NODEV (-6)      /* no usable GPU */

/* precision modes */
#define GP_FP64 0            /* FP64 storage and arithmetic, same operation order as the reference */
#define GP_FP32 1            /* FP32 phi/mu storage and stencil arithmetic (bulk mu increments in direct form),
                                FP64 statistics and set-up */
#define GP_FP32C 2           /* compensated FP32: phi/mu stored as FP32 plus an 8-bit extension of each value
                                (about 31 significant bits); bulk cells step in FP32 arithmetic with the increment
                                added to the extended value, interface cells step in FP64 arithmetic from the
                                extended values; FP64 statistics and set-up. The cell-classified step only. */

/* closed-form solve of d(mu) = (dc/dmu)^-1 dc (GPs/solverloop/calculate_divergence_concentration.h:93-125) */
#define GP_SOLVE_LU 0
#define GP_SOLVE_BINARY 1
#define GP_SOLVE_TERNARY 2
#define GP_SOLVE_DILUTE 3

/* Flat snapshot of the globals read by the solver loop. Index conventions:
 *   pair (a,b)        -> a*GP_MAXP + b
 *   triple (a,b,c)    -> (a*GP_MAXP + b)*GP_MAXP + c
 *   [phase][i][j]     -> (phase*GP_MAXNC + i)*GP_MAXNC + j
 *   [phase][i]        -> phase*GP_MAXNC + i
 *   rotation (a,b)    -> ((a*GP_MAXP + b)*2 + i)*2 + j   (2-D block of the 3x3 matrix)        */
typedef struct gp_params {
  int32_t nphases;             /* NUMPHASES; the last phase is the liquid/matrix */
  int32_t ncomp;               /* NUMCOMPONENTS; NC = ncomp-1 mu fields */
  int64_t mesh_x, mesh_y;      /* MESH_X, MESH_Y (interior cells) */
  double dx, dy, dt;           /* DELTA_X, DELTA_Y, DELTA_t */
  double epsilon;              /* epsilon */
  double tau;                  /* tau: returned by FunctionTau when sum phi_a phi_b == 0 */
  double T, Teq;               /* T (isothermal), Equilibrium_temperature */
  int32_t fn_w;                /* Function_W: 1 obstacle, 2 well */
  int32_t fn_aniso;            /* Function_anisotropy: 0 isotropic (A_00), 1 anisotropic (A_01) */
  int32_t fold;                /* Anisotropy_type (4) */
  int32_t solve;               /* GP_SOLVE_* from BINARY/TERNARY/DILUTE flags */
  int32_t jat_skip_row;        /* allocated y row whose y-face anti-trapping flux is skipped:
                                  MESH_Y-2 in the serial driver (calculate_fluxes_concentration.h:65); -1 none */
  int32_t reserved0;
  int32_t halo_src_x[2][6];    /* [field 0 phi / 1 mu][k]: source x of allocated x = 0,1,2, mx+3, mx+4, mx+5 */
  int32_t halo_src_y[2][6];    /* same for y */
  double tau_ab[GP_MAXP * GP_MAXP];
  double gamma[GP_MAXP * GP_MAXP];
  double gamma_abc[GP_MAXP * GP_MAXP * GP_MAXP];
  double dab[GP_MAXP * GP_MAXP];
  double rot[GP_MAXP * GP_MAXP * 4];       /* Rotation_matrix[a][b] (2-D block) */
  double rot_inv[GP_MAXP * GP_MAXP * 4];   /* Inv_Rotation_matrix[a][b] (2-D block) */
  double diffusivity[GP_MAXP * GP_MAXNC * GP_MAXNC];
  double A[GP_MAXP * GP_MAXNC * GP_MAXNC]; /* A[a][i][j] as given (f = sum_{i<=j} A_ij c_i c_j + ...) */
  double cmu[GP_MAXP * GP_MAXNC * GP_MAXNC]; /* (d mu/d c)^-1 per phase */
  double B[GP_MAXP * GP_MAXNC];            /* B(T) */
  double C[GP_MAXP];                       /* C(T) */
  double Beq[GP_MAXP * GP_MAXNC];          /* B(Teq) */
  double dBbdT[GP_MAXP * GP_MAXNC];        /* dB/dT */
  /* Non-isothermal runs and the moving frame. ISOTHERMAL = 0 turns on TEMPGRADY in the serial driver
   * (GPs/functions/reading_input_parameters.h:183-188): T depends on y only, is set from the Tempgrady key after
   * every step (GPs/microsim_gp.c:108-113, GPs/functions/Temperature_gradient.h:4-29) and enters c(mu), Mu(c)
   * through Beq + dBbdT (T - Teq) and the free energy through B(T), C(T) (GPs/functions/functionF_01.h:5-104).
   * Shift = 1 moves the frame along y every 100 steps (GPs/microsim_gp.c:115-126, GPs/functions/shift.h). */
  int32_t tempgrady;                       /* 1: ISOTHERMAL = 0 (temperature gradient along y) */
  int32_t shift;                           /* Shift */
  int64_t shiftj;                          /* Shiftj: the interface row the frame keeps */
  double tg_base_temp, tg_deltaT, tg_distance, tg_velocity; /* Tempgrady = {base_temp, DeltaT, Distance, -, velocity} */
  double tg_offset;                        /* Tempgrady gradient_OFFSET, after the restart adjustment
                                              (GPs/microsim_gp.c:68-75) */
  int32_t halo_src_T_y[6];                 /* source y of the halo columns for T (boundary[.][2]); a column that
                                              copies nothing maps to itself */
  int32_t reserved1[2];
  double ceq_sol[GP_MAXP * GP_MAXNC];      /* ceq[a][a][k]: B(T), C(T) of phase a < liquid */
  double ceq_liq[GP_MAXP * GP_MAXNC];      /* ceq[a][L][k] */
  double delta_c[GP_MAXP * GP_MAXNC];      /* DELTA_C[a][k] */
  double delta_T_sol[GP_MAXP];             /* DELTA_T[a][a] */
  double delta_T_liq[GP_MAXP];             /* DELTA_T[a][L] */
  double dcbdT_phase[GP_MAXP * GP_MAXNC];  /* dcbdT_phase of the dc/dT source (the serial driver leaves it 0) */
  double mu_fill[GP_MAXNC];                /* Mu(cfill[L][L], Tfill, L): mu of the liquid rows a shift adds */
} gp_params;

/* Diagnostics. max/min are cumulative from the last gp_upload with the reference's initial values
 * (phi_max 1, phi_min 0, mu_max = mu_min = 1, GPs/solverloop/serialinfo_xy.h:52-59); sums are over interior
 * cells of the last step. solute/energy are evaluated on the state after the last step. */
typedef struct gp_stats {
  double phi_max[GP_MAXP], phi_min[GP_MAXP];
  double mu_max[GP_MAXNC], mu_min[GP_MAXNC];
  double rel_change_phi[GP_MAXP];          /* sum (delta phi_a)^2 */
  double rel_change_mu[GP_MAXNC];          /* sum (delta mu_k)^2 */
  double solute[GP_MAXNC];                 /* sum_cells c_k, c_k = sum_a h_a(phi) c_a^k(mu), times dx*dy */
  double energy;                           /* sum_cells [psi + eps*a(grad phi) + w(phi)/eps] dx*dy */
  int64_t step;                            /* steps taken since gp_upload */
  int64_t nonfinite;                       /* interior values that are not finite */
} gp_stats;

/* AoS cell of the reference host (struct fields, GPs/functions/global_vars.h:327-332) for the adapters. */
typedef struct gp_cell {
  double *phia;
  double *compi;
  double *deltaphi;
  double temperature;
} gp_cell;

typedef struct gp_handle gp_handle;

/* Create a solver for the given parameters on the current HIP device. precision: GP_FP64 or GP_FP32.
 * Replaces the set-up the loop relies on: serialinfo_xy (GPs/solverloop/serialinfo_xy.h:1-60) for the grid,
 * initialize_functions_solverloop (GPs/solverloop/initialize_functions_solverloop.h:4-57) for the model choice. */
int gp_create(const gp_params *params, int precision, gp_handle **out);

/* Copy the state in (halo-inclusive SoA, see above); halos are rebuilt from the interior with the maps,
 * as apply_boundary_conditions(0) does before the loop (GPs/microsim_gp.c:84). Resets the statistics. */
int gp_upload(gp_handle *h, const double *phi, const double *mu);

/* Same from the reference's AoS array (ncells = rows_x*rows_y). */
int gp_upload_fields(gp_handle *h, const gp_cell *cells, int64_t ncells);

/* Copy the state out. dphi may be NULL; it receives the last step's projected delta-phi on the cells the
 * interior update reads (x,y in [2, mesh+3]) and 0 elsewhere. */
int gp_download(gp_handle *h, double *phi, double *mu, double *dphi);

/* Same into the reference's AoS array (phia, compi, deltaphi and temperature of every cell). */
int gp_download_fields(gp_handle *h, gp_cell *cells, int64_t ncells);

/* nsteps pre-conditioning steps (the NSMOOTH-1 loop): no driving force, no anti-trapping current. */
int gp_smooth(gp_handle *h, int64_t nsteps);

/* nsteps explicit time steps (phi sweep, BC, mu sweep, BC). stats may be NULL. When given, every one of these
 * steps folds its max/min into the cumulative ones on the device (as the reference does every step), the sums
 * of the last step are returned with them (one stream synchronisation at the end), and GP_E_NAN is returned if
 * a step produced a non-finite value. Steps of a call without stats skip the statistics work entirely; the
 * cumulative max/min then cover only the steps of calls that passed stats. */
int gp_step(gp_handle *h, int64_t nsteps, gp_stats *stats);

/* Diagnostics of the current state without stepping (max/min, solute, energy; rel_change of the last step). */
int gp_get_stats(gp_handle *h, gp_stats *stats);

/* Frame of a non-isothermal / moving-frame run: shift_offset (may be NULL) receives the cells the frame has moved
 * since gp_upload (shift_OFFSET of GPs/microsim_gp.c:122; DATA/shift.dat adds the restart's shift_position),
 * T_rows (may be NULL) the temperature of every allocated y row (rows_y values; the reference's
 * gridinfo[].temperature, which depends on y only) as it stands after the last step. Isothermal runs report
 * offset 0 and T everywhere. */
int gp_get_frame(gp_handle *h, int64_t *shift_offset, double *T_rows);

/* Solute and energy (the gp_stats fields solute[NC] and energy, as gp_get_stats computes them) of a caller's
 * device-resident state, with no device allocation and no synchronisation.
 * No counterpart in the reference. It is for callers that cannot allocate device memory: code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc.
 *   phi, mu  device arrays in the reference's halo-inclusive layout (as gp_upload takes on the host: each component
 *            rows_x*rows_y elements, index x*rows_y + y, rows = mesh + 6), of the handle's element type (double for
 *            GP_FP64 and for the decoded values of GP_FP32C, float for GP_FP32); the halo cells must hold the boundary images (gp_download gives such arrays)
 *   out      device array of NC + 1 doubles: solute[0..NC-1], then energy
 *   stream   hipStream_t to enqueue on, or NULL for the handle's stream
 * The per-block sums go to a device array the library holds for the life of the process, one per GPU, of
 * GP_NOALLOC_SCRATCH_DOUBLES doubles: a grid needs (NC + 1) * ceil(mesh_y/64) * ceil(mesh_x/4) of them (8192^2 with
 * NC = 2 fits), larger grids are rejected with GP_E_ARG. Every call on one GPU shares that array: calls that overlap
 * in time on one GPU (other streams or handles) race and must be ordered by the caller. The handle's device must be
 * current. */
#define GP_NOALLOC_SCRATCH_DOUBLES (1 << 20)
int gp_invariants_noalloc(gp_handle *h, const void *phi, const void *mu, double *out, void *stream);

/* Block until all work queued on the handle's stream is done. */
int gp_synchronize(gp_handle *h);

/* Message of the last failing call on h (or of the last failing gp_create when h is NULL). */
const char *gp_last_error(const gp_handle *h);

void gp_destroy(gp_handle *h);

/* Library version string, e.g. "gp_hip 0.1 gfx950". */
const char *gp_version(void);

/* Measurement support (no counterpart in the reference; for benchmarks and profilers). In GP_PROFILE_LAUNCHES mode
 * every kernel launch issued by gp_step / gp_smooth is bracketed by HIP events recorded on the handle's stream; in
 * GP_PROFILE_CALLS mode only each whole gp_step / gp_smooth call is (kind GP_KERNEL_CALL), which adds nothing
 * between the launches of a call. gp_profile_read synchronises, returns the mean duration in ms and the count per
 * kind (arrays of GP_NKERNELS) accumulated since the previous read, and resets them. */
#define GP_PROFILE_OFF 0
#define GP_PROFILE_LAUNCHES 1
#define GP_PROFILE_CALLS 2
#define GP_KERNEL_PHI 0      /* delta-phi kernel */
#define GP_KERNEL_MU 1       /* anti-trapping + mu/phi update kernel */
#define GP_KERNEL_STATS 2    /* statistics fold */
#define GP_KERNEL_FUSED 3    /* single-pass step kernel */
#define GP_KERNEL_FUSED_STATS 4 /* single-pass step kernel with the statistics reduction */
#define GP_KERNEL_BULK 5     /* single-pass step kernel of the tiles inside one pure phase */
#define GP_KERNEL_DPHI 6     /* delta-phi of the interface candidate cells */
#define GP_KERNEL_ADJ 7      /* one transposed step of gp_adjoint_backward (its segment recompute not included) */
#define GP_KERNEL_CBULK 8    /* update of the bulk cells (cell-classified step) */
#define GP_KERNEL_CIFACE 9   /* update of the interface cells (cell-classified step, k_ilist) */
#define GP_KERNEL_CALL 10    /* one whole gp_step / gp_smooth call (GP_PROFILE_CALLS mode) */
#define GP_KERNEL_HALO 11    /* halo copies of the new state (cell-classified step, k_halo) */
#define GP_NKERNELS 12
/* mode: GP_PROFILE_OFF, GP_PROFILE_LAUNCHES (any other nonzero value) or GP_PROFILE_CALLS */
int gp_profile_enable(gp_handle *h, int mode);
int gp_profile_read(gp_handle *h, double *avg_ms, int64_t *count);
const char *gp_kernel_name(int kind);

/* Adjoint of the solver loop (no counterpart in the reference; for gradient-based calibration).
 *
 * The differentiable parameters theta are, in this order: for every phase pair a < b (a outer, b inner)
 * dab[a][b], tau_ab[a][b] and Gamma[a][b] (each entering the model symmetrically in a and b); the liquid's
 * diagonal diffusivities D[N-1][k][k], k < NC; and tau (the relaxation time used where sum phi_a phi_b = 0).
 * gp_theta_count gives their number, gp_theta_values their current values. */
int gp_theta_count(const gp_handle *h);
int gp_theta_values(const gp_handle *h, double *theta);

/* nsteps steps from the current state, as gp_step does, keeping a device checkpoint of the state every `every`
 * steps (every <= 0: about sqrt(nsteps)). Checkpoints hold the handle's precision: an FP32 handle records and
 * recomputes its FP32 trajectory. */
int gp_adjoint_forward(gp_handle *h, int64_t nsteps, int64_t every);

/* Reverse sweep of the trajectory recorded by gp_adjoint_forward. Given the gradient of a scalar loss with
 * respect to the final state (gphi_T, gmu_T: halo-inclusive host arrays as in gp_upload), returns its gradient
 * with respect to the state passed to gp_upload (gphi_0, gmu_0; the halo entries, which gp_upload overwrites
 * from the interior, get 0) and with respect to theta. Segments between checkpoints are recomputed; each step is
 * transposed exactly: a closed-form gather for bulk cells and, for the other cells, reverse-mode passes through
 * the step's own per-cell arithmetic. On an FP32 or compensated FP32 handle the Jacobians are those of the FP64
 * model at the mode's (decoded) states, widened to FP64 for each transposed step; adjoint fields and parameter
 * gradients are accumulated in FP64 in every precision. The handle holds the initial state afterwards.
 * Environment: GP_ADJOINT_CHECK=<n> scans the adjoint after every transposed step, returns GP_E_NAN at the first
 * non-finite value (naming the step, the cell and the state there) and prints the largest |adjoint| every n
 * steps to stderr. Non-isothermal and moving-frame runs are rejected by gp_adjoint_forward. */
int gp_adjoint_backward(gp_handle *h, const double *gphi_T, const double *gmu_T, double *gphi_0, double *gmu_0,
                        double *gtheta);

#ifdef __cplusplus
}
#endif
#endif /* GP_HIP_H_ */
