/*
 * gp_oracle.h -- FP64 CPU restatement of the serial grand-potential solver loop.
 *
 * TEST INFRASTRUCTURE ONLY. This is the checker that libgp_hip is compared against; it is never linked into
 * the product. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
 *
 * State arrays use the reference layout (halo-inclusive SoA, index x*rows_y + y, rows = mesh + 6).
 */
#ifndef GP_ORACLE_H_
#define GP_ORACLE_H_

#include "gp_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Copy the halo cells from the interior with the parameter maps (apply_boundary_conditions). */
void oracle_apply_bc(const gp_params *p, double *phi, double *mu);

/* nsteps time steps in place. dphi (N*rows) receives the projected delta-phi of the last step on the cells
 * x,y in [2, mesh+3] (others untouched). stats (may be NULL) accumulates max/min like the reference and gets
 * the sums of the last step (interior cells). nthreads <= 0: OpenMP default. */
int oracle_step(const gp_params *p, double *phi, double *mu, double *dphi, int64_t nsteps, gp_stats *stats,
                int nthreads);

/* nsteps smoothing steps (smooth(): no driving force, no anti-trapping). */
int oracle_smooth(const gp_params *p, double *phi, double *mu, int64_t nsteps, int nthreads);

/* Reset stats to the reference's initial values (phi_max 1, phi_min 0, mu_max = mu_min = 1). */
void oracle_stats_init(gp_stats *st);

/* Global solute and grand-potential energy of the current state (interior cells), FP64. */
void oracle_invariants(const gp_params *p, const double *phi, const double *mu, double *solute, double *energy);

#ifdef __cplusplus
}
#endif
#endif
