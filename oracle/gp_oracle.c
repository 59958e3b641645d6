/*
 * gp_oracle.c -- FP64 CPU restatement of the serial grand-potential solver loop.
 *
 * TEST INFRASTRUCTURE ONLY: the checker libgp_hip is compared against. Never part of the product.
 *
 * Restates, in the same floating-point operation order (no FMA contraction: build with -ffp-contract=off),
 * the arithmetic of the serial driver (GPs = Grand_potential_Finite_difference_2D_serial):
 *   face quantities     GPs/solverloop/calculate_gradients.h:92-185
 *   phi update          GPs/solverloop/calculate_divergence_phasefield.h:4-62
 *   simplex projection  GPs/solverloop/simplex_projection.h:1-90
 *   anti-trapping flux  GPs/solverloop/calculate_fluxes_concentration.h:4-145
 *   mu update           GPs/solverloop/calculate_divergence_concentration.h:4-153
 *   smoothing           GPs/solverloop/solverloop.h:72-90, calculate_divergence_phasefield.h:63-,
 *                       calculate_divergence_concentration.h:154-253
 *   model functions     GPs/functions/{function_A_00.h, function_A_01.h, functionQ.h, anisotropy_01.h,
 *                       functionW_01.h, functionW_02.h, functionF_01.h, functionH.h, functionTau.h,
 *                       functionD.h, matrix.h}
 *   boundary copies     GPs/solverloop/boundary_serial.h:176-203
 *
 * Structure differs from the reference on purpose: instead of the 4-layer ring buffer the face quantities
 * live in full-grid arrays, and the two sweeps are evaluated Jacobi-style over the whole grid. The reference
 * sweeps are Jacobi as well (every face quantity is formed from values not yet updated), so the per-cell
 * arithmetic is identical. The set of cells computed is the one the interior update depends on:
 * delta-phi on [2, mesh+3]^2, mu and phi on the interior [3, mesh+2]^2; halos are then copied from the
 * interior (only Neumann / periodic sides exist). Statistics are taken over interior cells.
 *
 * pow(x,-1) in the small LU (GPs/functions/matrix.h:121,127) is libm's pow here, as in the reference.
 *
 * Pinning: tests/test_oracle_golden.py compares this file against golden vectors written by the reference
 * binary built from its own sources (oracle/Makefile "ref"), at steps 0/1/10/N for eleven input decks covering
 * every branch (BINARY/TERNARY/DILUTE/LU solves, obstacle/well, isotropic/anisotropic with rotation, Neumann/
 * periodic/mixed sides, undercooling, smoothing, 4 phases with third-order terms). The invariants
 * (solute, energy) are not in the reference: they are pinned by closed forms in tests/test_invariants.py
 * (uniform pure phases; a linear two-phase profile with a known gradient, anisotropy and potential term).
 */
#include "gp_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define P GP_MAXP
#define NCM GP_MAXNC
#define X 0
#define Y 1

typedef struct {
  const gp_params *p;
  int N, NC, L;
  int64_t rx, ry, n, ex, ey; /* rows and last interior index (mesh+2) */
  double *phi, *mu, *dphi;
  double *gx, *gy, *sx, *sy, *cx, *cy; /* [a][idx] */
  double *jx, *jy;                     /* [k][idx] */
  double Bmu[P][NCM];                  /* Beq + dBbdT*(T-Teq) */
  int obstacle, well;
} ctx_t;

#define IDX(c, x, y) ((x) * (c)->ry + (y))
#define F(arr, c, a, x, y) ((arr)[(int64_t)(a) * (c)->n + IDX(c, x, y)])

/* ---------------------------------------------------------------- model functions (FP64, reference order) */
static double A_(const gp_params *p, int a, int i, int j) { return p->A[(a * NCM + i) * NCM + j]; }
static double CMU(const gp_params *p, int a, int i, int j) { return p->cmu[(a * NCM + i) * NCM + j]; }
static double DIF(const gp_params *p, int a, int i, int j) { return p->diffusivity[(a * NCM + i) * NCM + j]; }
static double G2(const double *m, int a, int b) { return m[a * P + b]; }
static double R2(const double *m, int a, int b, int i, int j) { return m[((a * P + b) * 2 + i) * 2 + j]; }

static double hphi(int N, const double *phi, int a) {
  double sum1 = 0.0;
  double sum = 3.0 * phi[a] * phi[a] - 2.0 * phi[a] * phi[a] * phi[a];
  for (int b = 0; b < N; b++)
    for (int c = 0; c < N; c++)
      if (b != a && c != a && b < c) sum1 += phi[b] * phi[c];
  sum1 *= 2.0 * phi[a];
  return sum + sum1;
}

/* dh_b / d phi_a */
static double dhphi(int N, const double *phi, int b, int a) {
  double sum = 0.0;
  if (a == b) {
    sum = 6.0 * phi[a] * (1.0 - phi[a]);
    for (int e = 0; e < N; e++)
      for (int f = 0; f < N; f++)
        if (e != b && f != b && e < f) sum += 2.0 * phi[e] * phi[f];
  } else {
    for (int e = 0; e < N; e++)
      if (e != b && e != a) sum += 2.0 * phi[e];
    sum *= phi[b];
  }
  return sum;
}

static double c_mu(const ctx_t *c, const double *mu, int a, int i) {
  double s = 0.0;
  for (int k = 0; k < c->NC; k++) s += CMU(c->p, a, i, k) * (mu[k] - c->Bmu[a][k]);
  return s;
}

static double Mu(const ctx_t *c, const double *cc, int a, int i) {
  double s = 2.0 * A_(c->p, a, i, i) * cc[i] + c->Bmu[a][i];
  for (int j = 0; j < c->NC; j++)
    if (i != j) s += A_(c->p, a, i, j) * cc[j];
  return s;
}

static double free_energy(const ctx_t *c, const double *cc, int a) {
  double s = 0.0;
  for (int i = 0; i < c->NC; i++) {
    for (int j = 0; j < c->NC; j++)
      if (i <= j) s += A_(c->p, a, i, j) * cc[i] * cc[j];
    s += c->p->B[a * NCM + i] * cc[i];
  }
  s += c->p->C[a];
  return s;
}

static double dpsi(const ctx_t *c, const double *mu, const double *phi, int a) {
  double sum = 0.0, cc[NCM];
  for (int b = 0; b < c->N; b++) {
    double psi = 0.0;
    for (int k = 0; k < c->NC; k++) {
      cc[k] = c_mu(c, mu, b, k);
      psi -= mu[k] * cc[k];
    }
    psi += free_energy(c, cc, b);
    sum += psi * dhphi(c->N, phi, b, a);
  }
  return sum;
}

static double FunctionTau(const ctx_t *c, const double *phi) {
  double sum = 0.0, sum1 = 0.0;
  for (int a = 0; a < c->N; a++)
    for (int b = 0; b < c->N; b++)
      if (a < b) {
        sum += c->p->tau_ab[a * P + b] * phi[a] * phi[b];
        sum1 += phi[a] * phi[b];
      }
  if (sum1) return sum / sum1;
  return c->p->tau;
}

/* multiply(R, v, out, 2) */
static void rot2(const double *R, int a, int b, const double *v, double *o) {
  for (int i = 0; i < 2; i++) {
    double s = 0;
    for (int j = 0; j < 2; j++) s = s + R2(R, a, b, i, j) * v[j];
    o[i] = s;
  }
}

static double function_ac(const gp_params *p, const double *q, int a, int b) {
  double qx2 = q[X] * q[X], qx4 = qx2 * qx2;
  double qy2 = q[Y] * q[Y], qy4 = qy2 * qy2;
  double qz2 = 0.0 * 0.0, qz4 = qz2 * qz2;
  double q2 = qx2 + qy2 + qz2;
  if (fabs(q2) > 1.0e-15) return 1.0 - G2(p->dab, a, b) * (3.0 - 4.0 * (qx4 + qy4 + qz4) / (q2 * q2));
  return 1.0;
}

static void dAdq(const gp_params *p, const double *q, double *d, int a, int b) {
  double qz = 0.0;
  double qx3 = q[X] * q[X] * q[X], qy3 = q[Y] * q[Y] * q[Y], qz3 = qz * qz * qz;
  double q2 = q[X] * q[X] + q[Y] * q[Y] + qz * qz;
  double q22 = q2 * q2, q23 = q2 * q2 * q2;
  double q4 = qx3 * q[X] + qy3 * q[Y] + qz3 * qz;
  if (fabs(q2) > 1.0e-15) {
    double dab = G2(p->dab, a, b);
    d[X] = 16.0 * dab * (qx3 / q22 - q[X] * q4 / q23);
    d[Y] = 16.0 * dab * (qy3 / q22 - q[Y] * q4 / q23);
  } else {
    d[X] = 0.0;
    d[Y] = 0.0;
  }
}

static double qnorm2(const double *q) { return q[X] * q[X] + q[Y] * q[Y] + 0.0 * 0.0; }

/* q on the x-face between (x,y) and (x+1,y) (q_divx) */
static void q_divx(const ctx_t *c, int64_t x, int64_t y, int a, int b, double *q) {
  double r[2];
  r[X] = (F(c->sx, c, a, x, y) * F(c->gx, c, b, x, y) - F(c->sx, c, b, x, y) * F(c->gx, c, a, x, y));
  r[Y] = 0.5 * (F(c->sx, c, a, x, y) * (F(c->cy, c, b, x, y) + F(c->cy, c, b, x + 1, y)) -
                F(c->sx, c, b, x, y) * (F(c->cy, c, a, x, y) + F(c->cy, c, a, x + 1, y)));
  rot2(c->p->rot, a, b, r, q);
}

/* q on the y-face between (x,y) and (x,y+1) (q_divy) */
static void q_divy(const ctx_t *c, int64_t x, int64_t y, int a, int b, double *q) {
  double r[2];
  r[X] = 0.5 * (F(c->sy, c, a, x, y) * (F(c->cx, c, b, x, y) + F(c->cx, c, b, x, y + 1)) -
                F(c->sy, c, b, x, y) * (F(c->cx, c, a, x, y) + F(c->cx, c, a, x, y + 1)));
  r[Y] = (F(c->sy, c, a, x, y) * F(c->gy, c, b, x, y) - F(c->sy, c, b, x, y) * F(c->gy, c, a, x, y));
  rot2(c->p->rot, a, b, r, q);
}

static double dAdphi_A01(const ctx_t *c, const double *phi, int64_t x, int64_t y, int a) {
  double sum1 = 0.0, sum2 = 0.0;
  for (int b = 0; b < c->N; b++) {
    if (b == a) continue;
    double r[2], q[2], dadq[2], dq[2], dqr[2];
    r[X] = (phi[a] * F(c->cx, c, b, x, y) - phi[b] * F(c->cx, c, a, x, y));
    r[Y] = (phi[a] * F(c->cy, c, b, x, y) - phi[b] * F(c->cy, c, a, x, y));
    rot2(c->p->rot, a, b, r, q);
    dAdq(c->p, q, dadq, a, b);
    double qab2 = qnorm2(q);
    double ac1 = function_ac(c->p, q, a, b), ac2 = ac1 * ac1;
    dq[X] = F(c->cx, c, b, x, y);
    dq[Y] = F(c->cy, c, b, x, y);
    rot2(c->p->rot, a, b, dq, dqr);
    double s1 = 0.0, s2 = 0.0;
    for (int d = 0; d < 2; d++) s1 += q[d] * dqr[d];
    for (int d = 0; d < 2; d++) s2 += dadq[d] * dqr[d];
    s1 *= G2(c->p->gamma, a, b) * ac2;
    s2 *= G2(c->p->gamma, a, b) * ac1 * qab2;
    sum1 += s1;
    sum2 += s2;
  }
  return 2.0 * (sum1 + sum2);
}

static double divdAdgradphi_A01(const ctx_t *c, int64_t x, int64_t y, int a) {
  double sum1 = 0.0, sum2 = 0.0;
  const double dx = c->p->dx, dy = c->p->dy;
  for (int b = 0; b < c->N; b++) {
    if (b == a) continue;
    double qf[2], qb[2], qr[2], ql[2];
    q_divx(c, x, y, a, b, qf);
    q_divx(c, x - 1, y, a, b, qb);
    q_divy(c, x, y, a, b, qr);
    q_divy(c, x, y - 1, a, b, ql);
    double acf = function_ac(c->p, qf, a, b), acb = function_ac(c->p, qb, a, b);
    double acr = function_ac(c->p, qr, a, b), acl = function_ac(c->p, ql, a, b);
    double acf2 = acf * acf, acb2 = acb * acb, acr2 = acr * acr, acl2 = acl * acl;
    double sxa = F(c->sx, c, a, x, y), sxb = F(c->sx, c, b, x, y), gxa = F(c->gx, c, a, x, y), gxb = F(c->gx, c, b, x, y);
    double Bsxa = F(c->sx, c, a, x - 1, y), Bsxb = F(c->sx, c, b, x - 1, y);
    double Bgxa = F(c->gx, c, a, x - 1, y), Bgxb = F(c->gx, c, b, x - 1, y);
    double sya = F(c->sy, c, a, x, y), syb = F(c->sy, c, b, x, y), gya = F(c->gy, c, a, x, y), gyb = F(c->gy, c, b, x, y);
    double Lsya = F(c->sy, c, a, x, y - 1), Lsyb = F(c->sy, c, b, x, y - 1);
    double Lgya = F(c->gy, c, a, x, y - 1), Lgyb = F(c->gy, c, b, x, y - 1);
    double d1 = ((acf2 * (sxa * gxb - sxb * gxa) * (-sxb)) - (acb2 * (Bsxa * Bgxb - Bsxb * Bgxa) * (-Bsxb))) / dx;
    d1 += ((acr2 * (sya * gyb - syb * gya) * (-syb)) - (acl2 * (Lsya * Lgyb - Lsyb * Lgya) * (-Lsyb))) / dy;
    d1 *= G2(c->p->gamma, a, b);
    sum1 += d1;
    double df[2], db[2], dr[2], dl[2], rf[2], rb[2], rr[2], rl[2];
    dAdq(c->p, qf, df, a, b);
    rot2(c->p->rot_inv, a, b, df, rf);
    dAdq(c->p, qb, db, a, b);
    rot2(c->p->rot_inv, a, b, db, rb);
    dAdq(c->p, qr, dr, a, b);
    rot2(c->p->rot_inv, a, b, dr, rr);
    dAdq(c->p, ql, dl, a, b);
    rot2(c->p->rot_inv, a, b, dl, rl);
    double qf2 = qnorm2(qf), qb2 = qnorm2(qb), qr2 = qnorm2(qr), ql2 = qnorm2(ql);
    double d2 = (acf * qf2 * rf[X] * (-sxb) - acb * qb2 * rb[X] * (-Bsxb)) / dx;
    d2 += (acr * qr2 * rr[Y] * (-syb) - acl * ql2 * rl[Y] * (-Lsyb)) / dy;
    d2 *= G2(c->p->gamma, a, b);
    sum2 += d2;
  }
  return (2.0 * (sum1 + sum2) + 0.0);
}

static double dAdphi_A00(const ctx_t *c, const double *phi, int64_t x, int64_t y, int a) {
  double sum = 0.0;
  for (int b = 0; b < c->N; b++) {
    if (b == a) continue;
    double s = 0.0;
    s += (phi[a] * F(c->cx, c, b, x, y) - phi[b] * F(c->cx, c, a, x, y)) * F(c->cx, c, b, x, y);
    s += (phi[a] * F(c->cy, c, b, x, y) - phi[b] * F(c->cy, c, a, x, y)) * F(c->cy, c, b, x, y);
    s *= G2(c->p->gamma, a, b);
    sum += s;
  }
  return 2.0 * sum;
}

static double divdAdgradphi_A00(const ctx_t *c, int64_t x, int64_t y, int a) {
  double sum = 0.0;
  const double dx = c->p->dx, dy = c->p->dy;
  for (int b = 0; b < c->N; b++) {
    if (b == a) continue;
    double sxa = F(c->sx, c, a, x, y), sxb = F(c->sx, c, b, x, y), gxa = F(c->gx, c, a, x, y), gxb = F(c->gx, c, b, x, y);
    double Bsxa = F(c->sx, c, a, x - 1, y), Bsxb = F(c->sx, c, b, x - 1, y);
    double Bgxa = F(c->gx, c, a, x - 1, y), Bgxb = F(c->gx, c, b, x - 1, y);
    double sya = F(c->sy, c, a, x, y), syb = F(c->sy, c, b, x, y), gya = F(c->gy, c, a, x, y), gyb = F(c->gy, c, b, x, y);
    double Lsya = F(c->sy, c, a, x, y - 1), Lsyb = F(c->sy, c, b, x, y - 1);
    double Lgya = F(c->gy, c, a, x, y - 1), Lgyb = F(c->gy, c, b, x, y - 1);
    double d = ((sxa * gxb - sxb * gxa) * (-sxb) - ((Bsxa * Bgxb - Bsxb * Bgxa) * (-Bsxb))) / dx;
    d += ((sya * gyb - syb * gya) * (-syb) - ((Lsya * Lgyb - Lsyb * Lgya) * (-Lsyb))) / dy;
    d *= G2(c->p->gamma, a, b);
    sum += d;
  }
  return 2.0 * sum;
}

/* A_01 smoothing variants (function_A_01.h:73-100, 220-259): the isotropic forms plus a zero sum2 */
static double dAdphi_A01_smooth(const ctx_t *c, const double *phi, int64_t x, int64_t y, int a) {
  double sum1 = 0.0, sum2 = 0.0;
  for (int b = 0; b < c->N; b++) {
    if (b == a) continue;
    double s = 0.0;
    s += (phi[a] * F(c->cx, c, b, x, y) - phi[b] * F(c->cx, c, a, x, y)) * F(c->cx, c, b, x, y);
    s += (phi[a] * F(c->cy, c, b, x, y) - phi[b] * F(c->cy, c, a, x, y)) * F(c->cy, c, b, x, y);
    s *= G2(c->p->gamma, a, b);
    sum1 += s;
  }
  return 2.0 * (sum1 + sum2);
}

static double divdAdgradphi_A01_smooth(const ctx_t *c, int64_t x, int64_t y, int a) {
  double sum1 = 0.0, sum2 = 0.0;
  const double dx = c->p->dx, dy = c->p->dy;
  for (int b = 0; b < c->N; b++) {
    if (b == a) continue;
    double sxa = F(c->sx, c, a, x, y), sxb = F(c->sx, c, b, x, y), gxa = F(c->gx, c, a, x, y), gxb = F(c->gx, c, b, x, y);
    double Bsxa = F(c->sx, c, a, x - 1, y), Bsxb = F(c->sx, c, b, x - 1, y);
    double Bgxa = F(c->gx, c, a, x - 1, y), Bgxb = F(c->gx, c, b, x - 1, y);
    double sya = F(c->sy, c, a, x, y), syb = F(c->sy, c, b, x, y), gya = F(c->gy, c, a, x, y), gyb = F(c->gy, c, b, x, y);
    double Lsya = F(c->sy, c, a, x, y - 1), Lsyb = F(c->sy, c, b, x, y - 1);
    double Lgya = F(c->gy, c, a, x, y - 1), Lgyb = F(c->gy, c, b, x, y - 1);
    double d = (((sxa * gxb - sxb * gxa) * (-sxb)) - ((Bsxa * Bgxb - Bsxb * Bgxa) * (-Bsxb))) / dx;
    d += (((sya * gyb - syb * gya) * (-syb)) - ((Lsya * Lgyb - Lsyb * Lgya) * (-Lsyb))) / dy;
    d *= G2(c->p->gamma, a, b);
    sum1 += d;
  }
  return 2.0 * (sum1 + sum2);
}

static double dwdphi(const ctx_t *c, const double *phi, const double *divphi, int a, int smooth) {
  const gp_params *p = c->p;
  const int N = c->N;
  double sum = 0.0;
  if (c->obstacle) {
    for (int b = 0; b < N; b++)
      if (b != a && fabs(divphi[b]) > 0.0) {
        if (smooth) sum += G2(p->gamma, a, b) * phi[b];
        else if ((phi[a] * phi[b]) >= 0.00000) sum += G2(p->gamma, a, b) * (phi[b]);
        else sum += -G2(p->gamma, a, b) * (phi[b]);
      }
    sum *= 16.0 / (M_PI * M_PI);
  } else {
    for (int b = 0; b < N; b++)
      if (b != a && fabs(divphi[b]) > 0.0) sum += 2.0 * G2(p->gamma, a, b) * (phi[b] * phi[b] * phi[a]);
    sum *= 9.0;
  }
  for (int b = 0; b < N; b++)
    for (int cc = 0; cc < N; cc++)
      if (b != a && cc != a && b < cc && fabs(divphi[b]) > 0.0 && fabs(divphi[cc]) > 0.0) {
        double g = p->gamma_abc[(a * P + b) * P + cc];
        if (smooth) {
          sum += g * phi[b] * phi[cc];
        } else {
          double pbc = phi[b] * phi[cc];
          if (phi[a] * pbc >= 0.00000) sum += g * pbc;
          else sum += -g * pbc;
        }
      }
  return sum;
}

/* projection_on_simplex (simplex_projection.h:1-49) */
static void projection(int N, const double *phi, const double *divphi, double *dphi) {
  double Deltaphi = 0.0, sum_phib = 0.0;
  long count = 0;
  for (int a = 0; a < N; a++)
    if ((fabs(divphi[a]) > 0.0) && (phi[a] + dphi[a]) > 0.0 && (phi[a] + dphi[a]) < 1.0) {
      count++;
      sum_phib += phi[a];
    }
  for (int a = 0; a < N; a++)
    if ((phi[a] + dphi[a]) < 0.0) {
      double d = fabs(dphi[a] + phi[a]);
      dphi[a] += d;
      Deltaphi += d;
    }
  if (Deltaphi > 0.0)
    for (int b = 0; b < N; b++)
      if ((fabs(divphi[b]) > 0.0) && ((phi[b] + dphi[b]) > 0.0) && ((phi[b] + dphi[b]) < 1.0)) {
        if (fabs(sum_phib) > 0.0) dphi[b] -= Deltaphi * (phi[b]) / (sum_phib);
        else dphi[b] -= Deltaphi / (count);
      }
  for (int a = 0; a < N; a++)
    if (fabs(divphi[a]) > 0.0 && (phi[a] + dphi[a]) > 1.0) {
      dphi[a] = (1.0 - phi[a]);
      for (int b = 0; b < N; b++)
        if ((fabs(divphi[b]) > 0.0) && b != a) dphi[b] = -phi[b];
      break;
    }
}

/* projection_on_simplex_without_weights (simplex_projection.h:50-90) */
static void projection_ww(int N, const double *phi, const double *divphi, double *dphi) {
  long count = 0;
  for (int a = 0; a < N; a++)
    if ((fabs(divphi[a]) > 0.0) && (phi[a] + dphi[a]) > 0.0 && (phi[a] + dphi[a]) < 1.0) count++;
  for (int a = 0; a < N; a++) {
    if (fabs(divphi[a]) > 0.0) {
      if ((phi[a] + dphi[a]) > 1.0) {
        dphi[a] = (1.0 - phi[a]);
        for (int b = 0; b < N; b++)
          if ((fabs(divphi[b]) > 0.0) && b != a) dphi[b] = -phi[b];
        break;
      }
      if ((phi[a] + dphi[a]) < 0.0) {
        double d = fabs(dphi[a] + phi[a]);
        dphi[a] += d;
        for (int b = 0; b < N; b++)
          if (((fabs(divphi[b]) > 0.0) && b != a) && (phi[b] + dphi[b]) > 0.0 && (phi[b] + dphi[b]) < 1.0)
            dphi[b] -= d / (count);
      }
    }
  }
}

/* D(): M_ij = sum_b sum_k D_b,ik cmu_b,kj phi_b (functionD.h:1-16) */
static double Dfun(const ctx_t *c, int64_t x, int64_t y, int i, int j) {
  double s = 0.0;
  for (int b = 0; b < c->N; b++)
    for (int k = 0; k < c->NC; k++) s += DIF(c->p, b, i, k) * CMU(c->p, b, k, j) * F(c->phi, c, b, x, y);
  return s;
}

/* ---------------------------------------------------------------- grid passes */
static void load_cell(const ctx_t *c, const double *base, int ncomp, int64_t x, int64_t y, double *out) {
  for (int a = 0; a < ncomp; a++) out[a] = base[(int64_t)a * c->n + IDX(c, x, y)];
}

static void face_pass(ctx_t *c) {
  const double dx = c->p->dx, dy = c->p->dy;
  int64_t x;
#pragma omp parallel for schedule(static)
  for (x = 0; x < c->rx; x++)
    for (int64_t y = 0; y <= c->ey + 2; y++)
      for (int a = 0; a < c->N; a++) {
        if (x <= c->ex + 2) {
          F(c->gx, c, a, x, y) = (F(c->phi, c, a, x + 1, y) - F(c->phi, c, a, x, y)) / dx;
          F(c->sx, c, a, x, y) = 0.5 * (F(c->phi, c, a, x + 1, y) + F(c->phi, c, a, x, y));
        } else {
          F(c->gx, c, a, x, y) = 0.0;
          F(c->sx, c, a, x, y) = (F(c->phi, c, a, x, y));
        }
        F(c->gy, c, a, x, y) = (F(c->phi, c, a, x, y + 1) - F(c->phi, c, a, x, y)) / dy;
        F(c->sy, c, a, x, y) = 0.5 * (F(c->phi, c, a, x, y + 1) + F(c->phi, c, a, x, y));
      }
#pragma omp parallel for schedule(static)
  for (x = 0; x < c->rx; x++)
    for (int64_t y = 0; y <= c->ey + 2; y++)
      for (int a = 0; a < c->N; a++) {
        if (y > 0 && y <= c->ey + 2)
          F(c->cy, c, a, x, y) = (F(c->phi, c, a, x, y + 1) - F(c->phi, c, a, x, y - 1)) / (2.0 * dy);
        else
          F(c->cy, c, a, x, y) = F(c->gy, c, a, x, y);
        if (x > 0) F(c->cx, c, a, x, y) = 0.5 * (F(c->gx, c, a, x, y) + F(c->gx, c, a, x - 1, y));
        else F(c->cx, c, a, x, y) = 0.0;
      }
}

/* delta-phi of one cell (calculate_divergence_phasefield_2D / _smooth_2D body) */
static void phi_cell(const ctx_t *c, int64_t x, int64_t y, int smooth, double *dphi_out) {
  const gp_params *p = c->p;
  const int N = c->N;
  double phi[P], mu[NCM], divphi[P], lambda[P], dph[P];
  load_cell(c, c->phi, N, x, y, phi);
  load_cell(c, c->mu, c->NC, x, y, mu);
  for (int a = 0; a < N; a++) {
    divphi[a] = (F(c->gx, c, a, x, y) - F(c->gx, c, a, x - 1, y)) / p->dx;
    divphi[a] += (F(c->gy, c, a, x, y) - F(c->gy, c, a, x, y - 1)) / p->dy;
  }
  double sum_lambda = 0.0;
  long active = 0;
  for (int a = 0; a < N; a++) {
    if (fabs(divphi[a]) > 0.0) {
      if (!smooth) {
        double dA = p->fn_aniso ? dAdphi_A01(c, phi, x, y, a) : dAdphi_A00(c, phi, x, y, a);
        double dD = p->fn_aniso ? divdAdgradphi_A01(c, x, y, a) : divdAdgradphi_A00(c, x, y, a);
        lambda[a] = p->epsilon * (-dA + dD) - (1.0 / p->epsilon) * dwdphi(c, phi, divphi, a, 0) - dpsi(c, mu, phi, a);
      } else if (p->fn_aniso) {
        lambda[a] = p->epsilon * (-dAdphi_A01_smooth(c, phi, x, y, a) + divdAdgradphi_A01_smooth(c, x, y, a)) -
                    (1.0 / p->epsilon) * dwdphi(c, phi, divphi, a, 1);
      } else {
        lambda[a] = p->epsilon * (-dAdphi_A00(c, phi, x, y, a) + divdAdgradphi_A00(c, x, y, a)) -
                    (1.0 / p->epsilon) * dwdphi(c, phi, divphi, a, 0);
      }
      sum_lambda += lambda[a];
      active++;
    }
  }
  if (active) sum_lambda /= active;
  for (int a = 0; a < N; a++) {
    if (fabs(divphi[a]) > 0.0) dph[a] = p->dt * (lambda[a] - sum_lambda) / (FunctionTau(c, phi) * p->epsilon);
    else dph[a] = 0.0;
  }
  if (smooth) projection_ww(N, phi, divphi, dph);
  else projection(N, phi, divphi, dph);
  for (int a = 0; a < N; a++) dphi_out[a] = dph[a];
}

/* anti-trapping current on the faces of (x,y) (calculate_fluxes_concentration_2D body) */
static void jat_cell(const ctx_t *c, int64_t x, int64_t y) {
  const gp_params *p = c->p;
  const int N = c->N, L = c->L, NC = c->NC;
  double *jx = c->jx, *jy = c->jy;
  for (int k = 0; k < NC; k++) {
    F(jx, c, k, x, y) = 0.0;
    F(jy, c, k, x, y) = 0.0;
  }
  double phr[P], phc[P], phf[P], mur[NCM], muc[NCM], muf[NCM];
  load_cell(c, c->phi, N, x, y + 1, phr);
  load_cell(c, c->phi, N, x, y, phc);
  load_cell(c, c->phi, N, x + 1, y, phf);
  load_cell(c, c->mu, NC, x, y + 1, mur);
  load_cell(c, c->mu, NC, x, y, muc);
  load_cell(c, c->mu, NC, x + 1, y, muf);
  double gxl = 0.5 * (F(c->cx, c, L, x, y) + F(c->cx, c, L, x, y + 1));
  double nyl = sqrt(gxl * gxl + F(c->gy, c, L, x, y) * F(c->gy, c, L, x, y));
  double gyl = 0.5 * (F(c->cy, c, L, x, y) + F(c->cy, c, L, x + 1, y));
  double nxl = sqrt(gyl * gyl + F(c->gx, c, L, x, y) * F(c->gx, c, L, x, y));
  const double w = 0.5 / sqrt(2.0);
  for (int a = 0; a < N - 1; a++) {
    double gpx = 0.5 * (F(c->cx, c, a, x, y) + F(c->cx, c, a, x, y + 1));
    double gya = F(c->gy, c, a, x, y);
    double ng = sqrt(gpx * gpx + gya * gya);
    double s_r = 0, s_c = 0, s_f = 0;
    if (c->obstacle) {
      s_r = (phr[a] * (1.0 - phr[a]) > 0.0) ? phr[a] * (1.0 - hphi(N, phr, a)) / (sqrt(phr[a] * (1.0 - phr[a]))) : 0.0;
      s_c = (phc[a] * (1.0 - phc[a]) > 0.0) ? phc[a] * (1.0 - hphi(N, phc, a)) / (sqrt(phc[a] * (1.0 - phc[a]))) : 0.0;
      s_f = (phf[a] * (1.0 - phf[a]) > 0.0) ? phf[a] * (1.0 - hphi(N, phf, a)) / (sqrt(phf[a] * (1.0 - phf[a]))) : 0.0;
    }
    double dr = F(c->dphi, c, a, x, y + 1), dc = F(c->dphi, c, a, x, y), df = F(c->dphi, c, a, x + 1, y);
    for (int k = 0; k < NC; k++) {
      if (ng != 0.0 && y != 0 && y != p->jat_skip_row) {
        double cl_r = c_mu(c, mur, L, k), cl_c = c_mu(c, muc, L, k);
        double cs_r = c_mu(c, mur, a, k), cs_c = c_mu(c, muc, a, k);
        double sp = gya * F(c->gy, c, L, x, y);
        sp += gpx * gxl;
        if (nyl > 0.0) sp /= (nyl * ng);
        double fac = (1.0 - DIF(p, a, k, k) / DIF(p, L, k, k));
        if (c->well) F(jy, c, k, x, y) += fac * fabs(sp) * w * ((cl_r - cs_r) * dr + (cl_c - cs_c) * dc) * gya / ng;
        if (c->obstacle) F(jy, c, k, x, y) += fac * fabs(sp) * 0.5 * (s_r * (cl_r - cs_r) * dr + s_c * (cl_c - cs_c) * dc) * gya / ng;
      }
    }
    double gpy = 0.5 * (F(c->cy, c, a, x, y) + F(c->cy, c, a, x + 1, y));
    double gxa = F(c->gx, c, a, x, y);
    ng = sqrt(gpy * gpy + gxa * gxa);
    for (int k = 0; k < NC; k++) {
      if (ng != 0.0) {
        double cl_f = c_mu(c, muf, L, k), cl_c = c_mu(c, muc, L, k);
        double cs_f = c_mu(c, muf, a, k), cs_c = c_mu(c, muc, a, k);
        double sp = gxa * F(c->gx, c, L, x, y);
        sp += gpy * gyl;
        if (nxl > 0.0) sp /= (nxl * ng);
        double fac = (1.0 - DIF(p, a, k, k) / DIF(p, L, k, k));
        if (c->well) F(jx, c, k, x, y) += fac * fabs(sp) * w * ((cl_f - cs_f) * df + (cl_c - cs_c) * dc) * (gxa) / ng;
        if (c->obstacle) F(jx, c, k, x, y) += fac * fabs(sp) * 0.5 * (s_f * (cl_f - cs_f) * df + s_c * (cl_c - cs_c) * dc) * (gxa) / ng;
      }
    }
  }
}

/* Dmid and grad mu on the faces of (x,y) */
static void mu_faces(const ctx_t *c, int64_t x, int64_t y, double Dx[NCM][NCM], double gmx[NCM], double Dy[NCM][NCM],
                     double gmy[NCM]) {
  for (int k = 0; k < c->NC; k++) {
    gmx[k] = (x <= c->ex + 2) ? (F(c->mu, c, k, x + 1, y) - F(c->mu, c, k, x, y)) / c->p->dx : 0.0;
    gmy[k] = (F(c->mu, c, k, x, y + 1) - F(c->mu, c, k, x, y)) / c->p->dy;
    for (int l = 0; l < c->NC; l++) {
      Dx[k][l] = (x <= c->ex + 2) ? (Dfun(c, x + 1, y, k, l) + Dfun(c, x, y, k, l)) * 0.5 : Dfun(c, x, y, k, l);
      Dy[k][l] = (Dfun(c, x, y + 1, k, l) + Dfun(c, x, y, k, l)) * 0.5;
    }
  }
}

static void small_inverse(int n, double a[][NCM], double inv[][NCM]) {
  /* matinvnew (matrix.h:50-114); the caller's matrix is scratch afterwards */
  double factor[NCM][NCM] = {{0}}, inv1[NCM][NCM];
  int tag[NCM];
  for (int k = 0; k < n; k++) tag[k] = k;
  for (int k = 0; k < n; k++) {
    double big = fabs(a[k][k]);
    int t1 = k;
    for (int i = k + 1; i < n; i++)
      if (fabs(a[i][k]) > big) {
        t1 = i;
        big = a[i][k];
      }
    int ts = tag[k];
    tag[k] = tag[t1];
    tag[t1] = ts;
    for (int i = 0; i < n; i++) {
      double s = a[k][i];
      a[k][i] = a[t1][i];
      a[t1][i] = s;
    }
    for (int i = 0; i < k; i++) {
      double s = factor[k][i];
      factor[k][i] = factor[t1][i];
      factor[t1][i] = s;
    }
    for (int i = k + 1; i < n; i++) {
      double fact = -a[i][k] / a[k][k];
      factor[i][k] = -fact;
      for (int j = k; j <= n - 1; j++) a[i][j] = fact * a[k][j] + a[i][j];
    }
  }
  for (int col = 0; col < n; col++) {
    double v1[NCM], v[NCM];
    v1[0] = (col == 0) ? 1.0 : 0.0;
    for (int i = 1; i < n; i++) {
      double s = 0;
      for (int j = 0; j < i; j++) s = s - factor[i][j] * v1[j];
      v1[i] = ((i == col) ? 1.0 : 0.0) + s;
    }
    v[n - 1] = v1[n - 1] * pow(a[n - 1][n - 1], -1);
    for (int i = n - 2; i >= 0; i--) {
      double s = 0;
      for (int j = i + 1; j < n; j++) s = s - a[i][j] * v[j];
      v[i] = (v1[i] + s) * pow(a[i][i], -1);
    }
    for (int j = 0; j < n; j++) inv1[j][col] = v[j];
  }
  for (int j = 0; j < n; j++)
    for (int q = 0; q < n; q++) inv[q][tag[j]] = inv1[q][j];
}

/* mu and phi update of one interior cell (calculate_divergence_concentration_2D / _smooth_2D body). */
static void mu_cell(const ctx_t *c, int64_t x, int64_t y, int smooth, const double *dph, double *phi_new,
                    double *mu_new, double *dmu_out) {
  const gp_params *p = c->p;
  const int N = c->N, NC = c->NC;
  double phi[P], mu[NCM], sum[NCM] = {0}, divflux[NCM], divjat[NCM] = {0}, deltamu[NCM] = {0};
  double dcdmu[NCM][NCM] = {{0}}, inv[NCM][NCM] = {{0}};
  load_cell(c, c->phi, N, x, y, phi);
  load_cell(c, c->mu, NC, x, y, mu);
  if (!smooth) {
    for (int a = 0; a < N; a++) {
      double sdh = 0.0;
      for (int b = 0; b < N; b++) sdh += dhphi(N, phi, a, b) * dph[b];
      for (int k = 0; k < NC; k++) sum[k] += c_mu(c, mu, a, k) * sdh;
    }
  }
  double Dx[NCM][NCM], gmx[NCM], Dy[NCM][NCM], gmy[NCM];
  double Dxb[NCM][NCM], gmxb[NCM], Dyl[NCM][NCM], gmyl[NCM], dum[NCM][NCM], dumv[NCM];
  mu_faces(c, x, y, Dx, gmx, Dy, gmy);
  mu_faces(c, x - 1, y, Dxb, gmxb, dum, dumv);
  mu_faces(c, x, y - 1, dum, dumv, Dyl, gmyl);
  for (int k = 0; k < NC; k++) {
    divflux[k] = 0.0;
    for (int l = 0; l < NC; l++) {
      divflux[k] += (Dx[k][l] * gmx[l] - Dxb[k][l] * gmxb[l]) / p->dx;
      divflux[k] += (Dy[k][l] * gmy[l] - Dyl[k][l] * gmyl[l]) / p->dy;
    }
    if (!smooth) {
      if (c->obstacle) {
        divjat[k] = (F(c->jx, c, k, x, y) - F(c->jx, c, k, x - 1, y)) * (-0.25 * M_PI * p->epsilon) / p->dx;
        divjat[k] += (F(c->jy, c, k, x, y) - F(c->jy, c, k, x, y - 1)) * (-0.25 * M_PI * p->epsilon) / p->dy;
      }
      if (c->well) {
        divjat[k] = (F(c->jx, c, k, x, y) - F(c->jx, c, k, x - 1, y)) * (-1.0 * p->epsilon) / p->dx;
        divjat[k] += (F(c->jy, c, k, x, y) - F(c->jy, c, k, x, y - 1)) * (-1.0 * p->epsilon) / p->dy;
      }
    }
  }
  int interface = 1, bulk = 0;
  for (int a = 0; a < N; a++)
    if (phi[a] == 1.0) {
      bulk = a;
      interface = 0;
      break;
    }
  double munew[NCM];
  for (int k = 0; k < NC; k++) munew[k] = mu[k];
  if (!interface) {
    double c_old[NCM], c_new[NCM];
    for (int k = 0; k < NC; k++) {
      c_old[k] = c_mu(c, mu, bulk, k);
      c_new[k] = smooth ? c_old[k] + p->dt * (divflux[k]) : c_old[k] + p->dt * (divflux[k]) - divjat[k];
    }
    for (int k = 0; k < NC; k++) {
      deltamu[k] = Mu(c, c_new, bulk, k) - munew[k];
      munew[k] = Mu(c, c_new, bulk, k);
    }
  } else {
    double deltac[NCM];
    for (int k = 0; k < NC; k++) {
      deltac[k] = smooth ? p->dt * (divflux[k]) : p->dt * (divflux[k]) - divjat[k] - sum[k];
      for (int l = 0; l < NC; l++)
        for (int a = 0; a < N; a++) dcdmu[k][l] += CMU(p, a, k, l) * hphi(N, phi, a);
    }
    int mode = smooth ? GP_SOLVE_LU : p->solve;
    if (mode == GP_SOLVE_BINARY) {
      inv[0][0] = 1.0 / dcdmu[0][0];
      deltamu[0] = deltac[0] * inv[0][0];
      munew[0] += deltamu[0];
    } else if (mode == GP_SOLVE_TERNARY) {
      double DET = dcdmu[0][0] * dcdmu[1][1] - dcdmu[0][1] * dcdmu[1][0];
      inv[0][0] = dcdmu[1][1] / DET;
      inv[1][1] = dcdmu[0][0] / DET;
      inv[0][1] = -dcdmu[0][1] / DET;
      inv[1][0] = -dcdmu[1][0] / DET;
      for (int k = 0; k < NC; k++) {
        deltamu[k] = 0.0;
        for (int l = 0; l < NC; l++) deltamu[k] += inv[k][l] * deltac[l];
      }
      for (int k = 0; k < NC; k++) munew[k] = deltamu[k] + munew[k];
    } else if (mode == GP_SOLVE_DILUTE) {
      for (int k = 0; k < NC; k++) {
        inv[k][k] = 1.0 / dcdmu[k][k];
        deltamu[k] = deltac[k] * inv[k][k];
        munew[k] += deltamu[k];
      }
    } else {
      small_inverse(NC, dcdmu, inv);
      for (int i = 0; i < NC; i++) {
        double s = 0;
        for (int j = 0; j < NC; j++) s = s + inv[i][j] * deltac[j];
        deltamu[i] = s;
      }
      for (int k = 0; k < NC; k++) munew[k] = deltamu[k] + munew[k];
    }
  }
  for (int b = 0; b < N; b++) phi_new[b] = phi[b] + dph[b];
  for (int k = 0; k < NC; k++) {
    mu_new[k] = munew[k];
    dmu_out[k] = deltamu[k];
  }
}

static int ctx_init(ctx_t *c, const gp_params *p, double *phi, double *mu) {
  memset(c, 0, sizeof(*c));
  c->p = p;
  c->N = p->nphases;
  c->NC = p->ncomp - 1;
  c->L = c->N - 1;
  c->rx = p->mesh_x + 6;
  c->ry = p->mesh_y + 6;
  c->n = c->rx * c->ry;
  c->ex = p->mesh_x + 2;
  c->ey = p->mesh_y + 2;
  c->phi = phi;
  c->mu = mu;
  c->obstacle = p->fn_w == 1;
  c->well = p->fn_w == 2;
  for (int a = 0; a < c->N; a++)
    for (int k = 0; k < c->NC; k++) c->Bmu[a][k] = (p->Beq[a * NCM + k] + p->dBbdT[a * NCM + k] * (p->T - p->Teq));
  size_t fa = sizeof(double) * (size_t)c->N * c->n, fk = sizeof(double) * (size_t)c->NC * c->n;
  c->gx = calloc(1, fa); c->gy = calloc(1, fa); c->sx = calloc(1, fa);
  c->sy = calloc(1, fa); c->cx = calloc(1, fa); c->cy = calloc(1, fa);
  c->jx = calloc(1, fk); c->jy = calloc(1, fk);
  c->dphi = calloc(1, fa);
  if (!c->gx || !c->gy || !c->sx || !c->sy || !c->cx || !c->cy || !c->jx || !c->jy || !c->dphi) return GP_E_OOM;
  return GP_OK;
}

static void ctx_free(ctx_t *c) {
  free(c->gx); free(c->gy); free(c->sx); free(c->sy); free(c->cx); free(c->cy);
  free(c->jx); free(c->jy); free(c->dphi);
}

void oracle_apply_bc(const gp_params *p, double *phi, double *mu) {
  const int64_t rx = p->mesh_x + 6, ry = p->mesh_y + 6, n = rx * ry;
  const int64_t hx[6] = {0, 1, 2, rx - 3, rx - 2, rx - 1}, hy[6] = {0, 1, 2, ry - 3, ry - 2, ry - 1};
  for (int f = 0; f < 2; f++) {
    double *base = f == 0 ? phi : mu;
    int nc = f == 0 ? p->nphases : p->ncomp - 1;
    for (int k = 0; k < nc; k++) {
      double *v = base + (int64_t)k * n;
      for (int s = 0; s < 6; s++)
        for (int64_t y = 0; y < ry; y++) v[hx[s] * ry + y] = v[(int64_t)p->halo_src_x[f][s] * ry + y];
      for (int s = 0; s < 6; s++)
        for (int64_t x = 0; x < rx; x++) v[x * ry + hy[s]] = v[x * ry + p->halo_src_y[f][s]];
    }
  }
}

void oracle_stats_init(gp_stats *st) {
  memset(st, 0, sizeof(*st));
  for (int a = 0; a < P; a++) {
    st->phi_max[a] = 1.0;
    st->phi_min[a] = 0.0;
  }
  for (int k = 0; k < NCM; k++) {
    st->mu_max[k] = 1.0;
    st->mu_min[k] = 1.0;
  }
}

static int run(const gp_params *p, double *phi, double *mu, double *dphi_out, int64_t nsteps, gp_stats *st,
               int smooth, int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  ctx_t c;
  if (ctx_init(&c, p, phi, mu)) {
    ctx_free(&c);
    return GP_E_OOM;
  }
  const int N = c.N, NC = c.NC;
  const int64_t MX = p->mesh_x, MY = p->mesh_y;
  double *phin = malloc(sizeof(double) * (size_t)N * MX * MY);
  double *mun = malloc(sizeof(double) * (size_t)(NC > 0 ? NC : 1) * MX * MY);
  double *dmu = malloc(sizeof(double) * (size_t)(NC > 0 ? NC : 1) * MX * MY);
  if (!phin || !mun || !dmu) {
    free(phin); free(mun); free(dmu); ctx_free(&c);
    return GP_E_OOM;
  }
  oracle_apply_bc(p, phi, mu);
  for (int64_t t = 0; t < nsteps; t++) {
    face_pass(&c);
    int64_t x;
    /* delta-phi on [2, mesh+3]^2 */
#pragma omp parallel for schedule(dynamic, 4)
    for (x = 2; x <= c.ex + 1; x++)
      for (int64_t y = 2; y <= c.ey + 1; y++) {
        double d[P];
        phi_cell(&c, x, y, smooth, d);
        for (int a = 0; a < N; a++) F(c.dphi, &c, a, x, y) = d[a];
      }
    if (!smooth) {
#pragma omp parallel for schedule(dynamic, 4)
      for (x = 2; x <= c.ex; x++)
        for (int64_t y = 2; y <= c.ey; y++) jat_cell(&c, x, y);
    }
#pragma omp parallel for schedule(dynamic, 4)
    for (x = 3; x <= c.ex; x++)
      for (int64_t y = 3; y <= c.ey; y++) {
        double d[P], pn[P], mn[NCM], dm[NCM];
        for (int a = 0; a < N; a++) d[a] = F(c.dphi, &c, a, x, y);
        mu_cell(&c, x, y, smooth, d, pn, mn, dm);
        int64_t i = (x - 3) * MY + (y - 3);
        for (int a = 0; a < N; a++) phin[(int64_t)a * MX * MY + i] = pn[a];
        for (int k = 0; k < NC; k++) {
          mun[(int64_t)k * MX * MY + i] = mn[k];
          dmu[(int64_t)k * MX * MY + i] = dm[k];
        }
      }
    /* statistics over interior cells, row-major order (deterministic) */
    if (st && !smooth) {
      for (int a = 0; a < N; a++) st->rel_change_phi[a] = 0.0;
      for (int k = 0; k < NC; k++) st->rel_change_mu[k] = 0.0;
      for (int64_t xx = 0; xx < MX; xx++)
        for (int64_t yy = 0; yy < MY; yy++) {
          int64_t i = xx * MY + yy;
          for (int a = 0; a < N; a++) {
            double v = phin[(int64_t)a * MX * MY + i], d = F(c.dphi, &c, a, xx + 3, yy + 3);
            if (v > st->phi_max[a]) st->phi_max[a] = v;
            if (v < st->phi_min[a]) st->phi_min[a] = v;
            st->rel_change_phi[a] += d * d;
            if (!isfinite(v)) st->nonfinite++;
          }
          for (int k = 0; k < NC; k++) {
            double v = mun[(int64_t)k * MX * MY + i], d = dmu[(int64_t)k * MX * MY + i];
            if (v > st->mu_max[k]) st->mu_max[k] = v;
            if (v < st->mu_min[k]) st->mu_min[k] = v;
            st->rel_change_mu[k] += d * d;
            if (!isfinite(v)) st->nonfinite++;
          }
        }
      st->step++;
    }
    for (int64_t xx = 0; xx < MX; xx++)
      for (int64_t yy = 0; yy < MY; yy++) {
        int64_t i = xx * MY + yy, g = (xx + 3) * c.ry + (yy + 3);
        for (int a = 0; a < N; a++) phi[(int64_t)a * c.n + g] = phin[(int64_t)a * MX * MY + i];
        for (int k = 0; k < NC; k++) mu[(int64_t)k * c.n + g] = mun[(int64_t)k * MX * MY + i];
      }
    oracle_apply_bc(p, phi, mu);
  }
  if (dphi_out) {
    for (int a = 0; a < N; a++)
      for (int64_t x = 2; x <= c.ex + 1; x++)
        for (int64_t y = 2; y <= c.ey + 1; y++) F(dphi_out, &c, a, x, y) = F(c.dphi, &c, a, x, y);
  }
  free(phin); free(mun); free(dmu);
  ctx_free(&c);
  return GP_OK;
}

int oracle_step(const gp_params *p, double *phi, double *mu, double *dphi, int64_t nsteps, gp_stats *stats,
                int nthreads) {
  return run(p, phi, mu, dphi, nsteps, stats, 0, nthreads);
}

int oracle_smooth(const gp_params *p, double *phi, double *mu, int64_t nsteps, int nthreads) {
  return run(p, phi, mu, NULL, nsteps, NULL, 1, nthreads);
}

/* ---------------------------------------------------------------- invariants (not in the reference)
 * solute_k = sum_cells sum_a h_a(phi) c_a^k(mu) dx dy
 * energy   = sum_cells [ sum_a h_a psi_a(mu) + eps * sum_{a<b} Gamma_ab ac(q_ab)^2 |q_ab|^2
 *                        + (1/eps) w(phi) ] dx dy,
 * psi_a(mu) = f_a(c_a(mu)) - mu . c_a(mu), q_ab from the centred gradients, w the obstacle
 * (16/pi^2) sum_{a<b} Gamma_ab phi_a phi_b + sum_{a<b<c} Gamma_abc phi_a phi_b phi_c or the well
 * 9 sum_{a<b} Gamma_ab phi_a^2 phi_b^2 (+ the same third-order term). */
void oracle_invariants(const gp_params *p, const double *phi, const double *mu, double *solute, double *energy) {
  ctx_t c;
  memset(&c, 0, sizeof c);
  c.p = p; c.N = p->nphases; c.NC = p->ncomp - 1; c.L = c.N - 1;
  c.rx = p->mesh_x + 6; c.ry = p->mesh_y + 6; c.n = c.rx * c.ry;
  for (int a = 0; a < c.N; a++)
    for (int k = 0; k < c.NC; k++) c.Bmu[a][k] = (p->Beq[a * NCM + k] + p->dBbdT[a * NCM + k] * (p->T - p->Teq));
  const int N = c.N, NC = c.NC;
  double sol[NCM] = {0}, en = 0.0;
  const double area = p->dx * p->dy;
  for (int64_t x = 3; x <= p->mesh_x + 2; x++)
    for (int64_t y = 3; y <= p->mesh_y + 2; y++) {
      double ph[P], m[NCM], h[P], gxc[P], gyc[P];
      for (int a = 0; a < N; a++) {
        ph[a] = phi[(int64_t)a * c.n + x * c.ry + y];
        gxc[a] = (phi[(int64_t)a * c.n + (x + 1) * c.ry + y] - phi[(int64_t)a * c.n + (x - 1) * c.ry + y]) / (2.0 * p->dx);
        gyc[a] = (phi[(int64_t)a * c.n + x * c.ry + y + 1] - phi[(int64_t)a * c.n + x * c.ry + y - 1]) / (2.0 * p->dy);
      }
      for (int k = 0; k < NC; k++) m[k] = mu[(int64_t)k * c.n + x * c.ry + y];
      for (int a = 0; a < N; a++) h[a] = hphi(N, ph, a);
      double e = 0.0;
      for (int a = 0; a < N; a++) {
        double cc[NCM], psi = 0.0;
        for (int k = 0; k < NC; k++) {
          cc[k] = c_mu(&c, m, a, k);
          sol[k] += h[a] * cc[k] * area;
          psi -= m[k] * cc[k];
        }
        psi += free_energy(&c, cc, a);
        e += h[a] * psi;
      }
      double grad = 0.0, w = 0.0;
      for (int a = 0; a < N; a++)
        for (int b = a + 1; b < N; b++) {
          double r[2] = {ph[a] * gxc[b] - ph[b] * gxc[a], ph[a] * gyc[b] - ph[b] * gyc[a]}, q[2];
          rot2(p->rot, a, b, r, q);
          double ac = p->fn_aniso ? function_ac(p, q, a, b) : 1.0;
          grad += G2(p->gamma, a, b) * ac * ac * qnorm2(q);
          if (p->fn_w == 1) w += (16.0 / (M_PI * M_PI)) * G2(p->gamma, a, b) * ph[a] * ph[b];
          else w += 9.0 * G2(p->gamma, a, b) * ph[a] * ph[a] * ph[b] * ph[b];
          for (int cc = b + 1; cc < N; cc++) w += p->gamma_abc[(a * P + b) * P + cc] * ph[a] * ph[b] * ph[cc];
        }
      e += p->epsilon * grad + w / p->epsilon;
      en += e * area;
    }
  for (int k = 0; k < NC; k++) solute[k] = sol[k];
  *energy = en;
}
