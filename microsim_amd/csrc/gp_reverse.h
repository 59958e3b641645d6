/*
 * gp_reverse.h -- reverse-mode (vector-Jacobian) forms of the per-cell functions of gp_model.h, for the adjoint.
 *
 * A step's two per-cell stages (gp_adjoint.h: P = phi_delta, Q = jat_x / jat_y / mu_update) are transposed here in
 * one pass per cell: the stage's forward values are recomputed with the functions of gp_model.h (so every branch,
 * gate and clamp takes the same decision as the forward step), then the incoming adjoint of the stage's outputs is
 * carried back through each operation to the stencil's inputs and to theta. The result is the derivative of the
 * branch the forward takes (at the kinks of the simplex projection and of |x| the one-sided derivative of that
 * branch), the same quantity the forward-mode dual numbers of gp_dual.h evaluate one direction at a time; the test
 * suite checks the two against each other on every golden deck.
 *
 * Reference arithmetic transposed (GPs = Grand_potential_Finite_difference_2D_serial):
 *   delta-phi            GPs/solverloop/calculate_divergence_phasefield.h:4-62
 *   simplex projection   GPs/solverloop/simplex_projection.h:1-49
 *   gradient energy      GPs/functions/function_A_01.h:4-219, function_A_00.h:4-58, functionQ.h, anisotropy_01.h
 *   potentials           GPs/functions/functionW_01.h, functionW_02.h
 *   dpsi, h, tau         GPs/functions/functionF_01.h:5-57, functionH.h:1-37, functionTau.h:1-17
 *   anti-trapping        GPs/solverloop/calculate_fluxes_concentration.h:4-145
 *   mu update            GPs/solverloop/calculate_divergence_concentration.h:4-153, functionD.h:1-16
 *
 * FP64 only (the adjoint of an FP32 trajectory evaluates the FP64 model at the FP32 states). Header-only and
 * __host__ __device__, so the host test build compiles the identical source.
 */
#pragma once

#include "gp_model.h"

namespace gpk {

template <int N>
constexpr int kPairs = N * (N - 1) / 2;
/* theta: per unordered pair (a < b, pair order) dab, tau_ab, Gamma; the liquid's diagonal diffusivities; tau */
template <int N, int NC>
constexpr int kThetaN = 3 * kPairs<N> + NC + 1;
inline int theta_count(int N, int NC) { return 3 * (N * (N - 1) / 2) + NC + 1; }

/* index of the unordered pair {a, b} in pair order (a < b, a outer) */
GP_HD int pair_index(int N, int a, int b) {
  if (a > b) {
    const int t = a;
    a = b;
    b = t;
  }
  return a * N - a * (a + 1) / 2 + (b - a - 1);
}

/* Adjoint of one cell's stencil inputs: phi on the 3x3 neighbourhood (p[a][i+1][j+1]), mu and delta-phi on the
 * 5-point one (slot five_slot(i, j)), and theta. */
template <int N, int NC>
struct RevAcc {
  double p[N][3][3];
  double u[NC][5];
  double d[N][5];
  double th[kThetaN<N, NC>];
  GP_HD void clear() {
    for (int a = 0; a < N; a++)
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) p[a][i][j] = 0.0;
    for (int k = 0; k < NC; k++)
      for (int o = 0; o < 5; o++) u[k][o] = 0.0;
    for (int a = 0; a < N; a++)
      for (int o = 0; o < 5; o++) d[a][o] = 0.0;
    for (int t = 0; t < kThetaN<N, NC>; t++) th[t] = 0.0;
  }
  GP_HD double &ph(int a, int i, int j) { return p[a][i + 1][j + 1]; }
  GP_HD double &m(int k, int i, int j) { return u[k][five_slot(i, j)]; }
  GP_HD double &dp(int a, int i, int j) { return d[a][five_slot(i, j)]; }
  GP_HD double &dab(int a, int b) { return th[3 * pair_index(N, a, b)]; }
  GP_HD double &tau_ab(int a, int b) { return th[3 * pair_index(N, a, b) + 1]; }
  GP_HD double &gamma(int a, int b) { return th[3 * pair_index(N, a, b) + 2]; }
  GP_HD double &dliq(int k) { return th[3 * kPairs<N> + k]; }
  GP_HD double &tau() { return th[3 * kPairs<N> + NC]; }
};

/* ---------------------------------------------------------------- stencil primitives transposed
 * gx(i,j) = (phi(i+1,j) - phi(i,j)) / dx, gy(i,j) = (phi(i,j+1) - phi(i,j)) / dy,
 * cy(i,j) = (phi(i,j+1) - phi(i,j-1)) / (2 dy), cx(i,j) = (gx(i,j) + gx(i-1,j)) / 2 */
template <int N, int NC>
GP_HD void add_gx(const Model<double> &m, RevAcc<N, NC> &g, int a, int i, int j, double v) {
  g.ph(a, i + 1, j) += v * m.inv_dx;
  g.ph(a, i, j) -= v * m.inv_dx;
}
template <int N, int NC>
GP_HD void add_gy(const Model<double> &m, RevAcc<N, NC> &g, int a, int i, int j, double v) {
  g.ph(a, i, j + 1) += v * m.inv_dy;
  g.ph(a, i, j) -= v * m.inv_dy;
}
template <int N, int NC>
GP_HD void add_cy(const Model<double> &m, RevAcc<N, NC> &g, int a, int i, int j, double v) {
  g.ph(a, i, j + 1) += v * m.inv_2dy;
  g.ph(a, i, j - 1) -= v * m.inv_2dy;
}
template <int N, int NC>
GP_HD void add_cx(const Model<double> &m, RevAcc<N, NC> &g, int a, int i, int j, double v) {
  add_gx(m, g, a, i, j, 0.5 * v);
  add_gx(m, g, a, i - 1, j, 0.5 * v);
}
/* s = (phi(i+1,j) + phi(i,j)) / 2 and the y form */
template <int N, int NC>
GP_HD void add_sx(RevAcc<N, NC> &g, int a, int i, int j, double v) {
  g.ph(a, i + 1, j) += 0.5 * v;
  g.ph(a, i, j) += 0.5 * v;
}
template <int N, int NC>
GP_HD void add_sy(RevAcc<N, NC> &g, int a, int i, int j, double v) {
  g.ph(a, i, j + 1) += 0.5 * v;
  g.ph(a, i, j) += 0.5 * v;
}

/* ---------------------------------------------------------------- h(phi) */
/* d^2 h_b / (d phi_a d phi_e) */
template <int N>
GP_HD double d2hphi(const double *phi, int b, int a, int e) {
  if (a == b && e == b) return 6.0 - 12.0 * phi[b];
  if (a == b || e == b) {
    const int o = a == b ? e : a; /* the other index */
    double s = 0.0;
    for (int f = 0; f < N; f++)
      if (f != b && f != o) s += phi[f];
    return 2.0 * s;
  }
  if (a == e) return 0.0;
  return 2.0 * phi[b];
}

/* ---------------------------------------------------------------- anisotropy: a_c and its derivatives
 * a_c = 1 - dab (3 - 4 Q4 / Q2^2), Q2 = |q|^2, Q4 = qx^4 + qy^4 (1 where Q2 <= 1e-15, as aniso_ac);
 * G = d a_c / d q (aniso_dadq), G0 = G / dab, H = d G / d q (symmetric), dac = d a_c / d dab. */
struct AcTerms {
  double ac, dac;
  double G[2], G0[2], H[2][2];
};
GP_HD AcTerms ac_terms(double dab, const double *q) {
  AcTerms t;
  const double q2 = q[0] * q[0] + q[1] * q[1];
  if (fabs(q2) > 1.0e-15) {
    const double q4 = q[0] * q[0] * q[0] * q[0] + q[1] * q[1] * q[1] * q[1];
    const double i2 = 1.0 / (q2 * q2), i3 = i2 / q2, i4 = i2 * i2;
    t.dac = -(3.0 - 4.0 * q4 * i2);
    t.ac = 1.0 + dab * t.dac;
    for (int i = 0; i < 2; i++) {
      t.G0[i] = 16.0 * (q[i] * q[i] * q[i] * i2 - q[i] * q4 * i3);
      t.G[i] = dab * t.G0[i];
    }
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2; j++) {
        double h = -4.0 * (q[i] * q[i] * q[i] * q[j] + q[i] * q[j] * q[j] * q[j]) * i3 + 6.0 * q[i] * q[j] * q4 * i4;
        if (i == j) h += 3.0 * q[i] * q[i] * i2 - q4 * i3;
        t.H[i][j] = dab * 16.0 * h;
      }
  } else {
    t.ac = 1.0;
    t.dac = 0.0;
    for (int i = 0; i < 2; i++) {
      t.G[i] = t.G0[i] = 0.0;
      for (int j = 0; j < 2; j++) t.H[i][j] = 0.0;
    }
  }
  return t;
}

/* y = R^T x for a 2x2 R */
GP_HD void rot2t(const double (*R)[2], const double *x, double *y) {
  y[0] = R[0][0] * x[0] + R[1][0] * x[1];
  y[1] = R[0][1] * x[0] + R[1][1] * x[1];
}

/* ---------------------------------------------------------------- dA/dphi (centre term) */
/* d/d(inputs) of dAdphi_A01(phi, cxc, cyc, a) times gres: into gphi, gcx, gcy (centre values) and theta */
template <int N, int NC>
GP_HD void dAdphi_A01_vjp(const Model<double> &m, const double *phi, const double *cxc, const double *cyc, int a,
                          double gres, double *gphi, double *gcx, double *gcy, RevAcc<N, NC> &acc) {
  const double g = 2.0 * gres;
  for (int b = 0; b < N; b++) {
    if (b == a) continue;
    const double(*R)[2] = m.rot[a][b];
    const double u[2] = {cxc[b], cyc[b]}, w[2] = {cxc[a], cyc[a]};
    double r[2], q[2], Ru[2];
    r[0] = (phi[a] * cxc[b] - phi[b] * cxc[a]);
    r[1] = (phi[a] * cyc[b] - phi[b] * cyc[a]);
    rot2(R, r, q);
    rot2(R, u, Ru);
    const AcTerms t = ac_terms(m.dab[a][b], q);
    const double gam = m.gamma[a][b];
    const double q2 = q[0] * q[0] + q[1] * q[1];
    const double qRu = q[0] * Ru[0] + q[1] * Ru[1];
    const double GRu = t.G[0] * Ru[0] + t.G[1] * Ru[1];
    const double G0Ru = t.G0[0] * Ru[0] + t.G0[1] * Ru[1];
    /* s1 = gam ac^2 (q . Ru), s2 = gam ac |q|^2 (G . Ru) */
    double gq[2], gRu[2];
    for (int j = 0; j < 2; j++) {
      const double HRu = t.H[j][0] * Ru[0] + t.H[j][1] * Ru[1];
      gq[j] = g * gam *
              (2.0 * t.ac * t.G[j] * qRu + t.ac * t.ac * Ru[j] + t.G[j] * q2 * GRu + t.ac * 2.0 * q[j] * GRu +
               t.ac * q2 * HRu);
      gRu[j] = g * gam * (t.ac * t.ac * q[j] + t.ac * q2 * t.G[j]);
    }
    acc.gamma(a, b) += g * (t.ac * t.ac * qRu + t.ac * q2 * GRu);
    acc.dab(a, b) += g * gam * (2.0 * t.ac * t.dac * qRu + t.dac * q2 * GRu + t.ac * q2 * G0Ru);
    double gr[2], gu[2];
    rot2t(R, gq, gr);
    rot2t(R, gRu, gu);
    /* r = phi_a u - phi_b w */
    gphi[a] += gr[0] * u[0] + gr[1] * u[1];
    gphi[b] -= gr[0] * w[0] + gr[1] * w[1];
    gcx[b] += phi[a] * gr[0] + gu[0];
    gcy[b] += phi[a] * gr[1] + gu[1];
    gcx[a] -= phi[b] * gr[0];
    gcy[a] -= phi[b] * gr[1];
  }
}

/* isotropic centre term: dAdphi_A00 */
template <int N>
GP_HD void dAdphi_A00_vjp(const Model<double> &m, const double *phi, const double *cxc, const double *cyc, int a,
                          double gres, double *gphi, double *gcx, double *gcy) {
  const double g = 2.0 * gres;
  for (int b = 0; b < N; b++) {
    if (b == a) continue;
    const double gg = g * m.gamma[a][b];
    /* s = (phi_a cx_b - phi_b cx_a) cx_b + (phi_a cy_b - phi_b cy_a) cy_b */
    const double rx = phi[a] * cxc[b] - phi[b] * cxc[a], ry = phi[a] * cyc[b] - phi[b] * cyc[a];
    gphi[a] += gg * (cxc[b] * cxc[b] + cyc[b] * cyc[b]);
    gphi[b] -= gg * (cxc[a] * cxc[b] + cyc[a] * cyc[b]);
    gcx[b] += gg * (phi[a] * cxc[b] + rx);
    gcy[b] += gg * (phi[a] * cyc[b] + ry);
    gcx[a] -= gg * phi[b] * cxc[b];
    gcy[a] -= gg * phi[b] * cyc[b];
  }
}

/* isotropic centre term's gamma derivative: the pair sum without gamma, times 2 */
template <int N, int NC>
GP_HD void dAdphi_A00_theta(const Model<double> &, const double *phi, const double *cxc, const double *cyc, int a,
                            double gres, RevAcc<N, NC> &acc) {
  for (int b = 0; b < N; b++) {
    if (b == a) continue;
    const double s = (phi[a] * cxc[b] - phi[b] * cxc[a]) * cxc[b] + (phi[a] * cyc[b] - phi[b] * cyc[a]) * cyc[b];
    acc.gamma(a, b) += 2.0 * gres * s;
  }
}

/* ---------------------------------------------------------------- div d(a)/d(grad phi) (face terms) */
/* One face of divdAdgradphi_A01 for the pair (a, b): normal component n (0: x-face, 1: y-face); sa, sb face
 * averages; ga, gb normal differences (divided by the spacing); Ca, Cb the sums of the two tangential centred
 * differences across the face. r_n = sa gb - sb ga, r_t = (sa Cb - sb Ca) / 2, q = R r,
 * T1 = ac^2 r_n (-sb), T2 = ac |q|^2 (Rinv G)_n (-sb). With aniso = false: T1 = r_n (-sb), T2 = 0 (the isotropic
 * form). Adds gT1 dT1 + gT2 dT2 to gf and gdab; returns T1, T2. */
struct FaceIn {
  double sa, sb, ga, gb, Ca, Cb;
};
GP_HD void face_vjp(bool aniso, double dab, const double (*R)[2], const double (*Ri)[2], int n, const FaceIn &f,
                    double gT1, double gT2, FaceIn &gf, double &gdab, double &T1, double &T2) {
  const int t = 1 - n;
  double r[2];
  r[n] = f.sa * f.gb - f.sb * f.ga;
  r[t] = 0.5 * (f.sa * f.Cb - f.sb * f.Ca);
  const double ms = -f.sb;
  double gr[2] = {0.0, 0.0};
  if (!aniso) {
    T1 = r[n] * ms;
    T2 = 0.0;
    gr[n] = gT1 * ms;
    gf.sb -= gT1 * r[n];
  } else {
    double q[2];
    rot2(R, r, q);
    const AcTerms a = ac_terms(dab, q);
    double rg[2], rg0[2];
    rot2(Ri, a.G, rg);
    rot2(Ri, a.G0, rg0);
    const double q2 = q[0] * q[0] + q[1] * q[1];
    T1 = a.ac * a.ac * r[n] * ms;
    T2 = a.ac * q2 * rg[n] * ms;
    double gq[2];
    for (int j = 0; j < 2; j++) {
      const double HRi = a.H[j][0] * Ri[n][0] + a.H[j][1] * Ri[n][1];
      gq[j] = gT1 * 2.0 * a.ac * a.G[j] * r[n] * ms +
              gT2 * (a.G[j] * q2 * rg[n] + a.ac * 2.0 * q[j] * rg[n] + a.ac * q2 * HRi) * ms;
    }
    rot2t(R, gq, gr);
    gr[n] += gT1 * a.ac * a.ac * ms;
    gf.sb -= gT1 * a.ac * a.ac * r[n] + gT2 * a.ac * q2 * rg[n];
    gdab += gT1 * 2.0 * a.ac * a.dac * r[n] * ms + gT2 * (a.dac * q2 * rg[n] + a.ac * q2 * rg0[n]) * ms;
  }
  gf.sa += gr[n] * f.gb + 0.5 * gr[t] * f.Cb;
  gf.gb += gr[n] * f.sa;
  gf.sb -= gr[n] * f.ga + 0.5 * gr[t] * f.Ca;
  gf.ga -= gr[n] * f.sb;
  gf.Cb += 0.5 * gr[t] * f.sa;
  gf.Ca -= 0.5 * gr[t] * f.sb;
}

/* divdAdgradphi_A01 (aniso) or divdAdgradphi_iso (isotropic) of phase a, transposed: gres times its derivative into
 * the stencil (acc.p) and theta (dab, Gamma). The face inputs are the forward's (gx, sx, cy, ... of gp_model.h). */
template <int N, int NC, class V>
GP_HD void divdA_vjp(const Model<double> &m, const V &v, int a, bool aniso, double gres, RevAcc<N, NC> &acc) {
  const double g = 2.0 * gres;
  for (int b = 0; b < N; b++) {
    if (b == a) continue;
    const double gam = m.gamma[a][b];
    const double dab = m.dab[a][b];
    const double(*R)[2] = m.rot[a][b];
    const double(*Ri)[2] = m.rot_inv[a][b];
    /* faces: f = +x at (0,0), b = +x at (-1,0), r = +y at (0,0), l = +y at (0,-1) */
    FaceIn fx[2], fy[2], gfx[2], gfy[2];
    double T1[4], T2[4];
    double gdab = 0.0;
    for (int s = 0; s < 2; s++) {
      const int i = s == 0 ? 0 : -1;
      fx[s] = FaceIn{sx(m, v, a, i, 0), sx(m, v, b, i, 0), gx(m, v, a, i, 0), gx(m, v, b, i, 0),
                     cy(m, v, a, i, 0) + cy(m, v, a, i + 1, 0), cy(m, v, b, i, 0) + cy(m, v, b, i + 1, 0)};
      fy[s] = FaceIn{sy(m, v, a, 0, i), sy(m, v, b, 0, i), gy(m, v, a, 0, i), gy(m, v, b, 0, i),
                     cx(m, v, a, 0, i) + cx(m, v, a, 0, i + 1), cx(m, v, b, 0, i) + cx(m, v, b, 0, i + 1)};
      gfx[s] = FaceIn{0, 0, 0, 0, 0, 0};
      gfy[s] = FaceIn{0, 0, 0, 0, 0, 0};
    }
    /* d = [(T(f) - T(b)) / dx + (T(r) - T(l)) / dy] gam for T1 and T2 */
    const double gx0 = g * gam * m.inv_dx, gy0 = g * gam * m.inv_dy;
    face_vjp(aniso, dab, R, Ri, 0, fx[0], gx0, gx0, gfx[0], gdab, T1[0], T2[0]);
    face_vjp(aniso, dab, R, Ri, 0, fx[1], -gx0, -gx0, gfx[1], gdab, T1[1], T2[1]);
    face_vjp(aniso, dab, R, Ri, 1, fy[0], gy0, gy0, gfy[0], gdab, T1[2], T2[2]);
    face_vjp(aniso, dab, R, Ri, 1, fy[1], -gy0, -gy0, gfy[1], gdab, T1[3], T2[3]);
    acc.gamma(a, b) += g * ((T1[0] - T1[1]) * m.inv_dx + (T1[2] - T1[3]) * m.inv_dy + (T2[0] - T2[1]) * m.inv_dx +
                            (T2[2] - T2[3]) * m.inv_dy);
    if (aniso) acc.dab(a, b) += gdab;
    for (int s = 0; s < 2; s++) {
      const int i = s == 0 ? 0 : -1;
      /* x-face at (i, 0) */
      add_sx(acc, a, i, 0, gfx[s].sa);
      add_sx(acc, b, i, 0, gfx[s].sb);
      add_gx(m, acc, a, i, 0, gfx[s].ga);
      add_gx(m, acc, b, i, 0, gfx[s].gb);
      add_cy(m, acc, a, i, 0, gfx[s].Ca);
      add_cy(m, acc, a, i + 1, 0, gfx[s].Ca);
      add_cy(m, acc, b, i, 0, gfx[s].Cb);
      add_cy(m, acc, b, i + 1, 0, gfx[s].Cb);
      /* y-face at (0, i) */
      add_sy(acc, a, 0, i, gfy[s].sa);
      add_sy(acc, b, 0, i, gfy[s].sb);
      add_gy(m, acc, a, 0, i, gfy[s].ga);
      add_gy(m, acc, b, 0, i, gfy[s].gb);
      add_cx(m, acc, a, 0, i, gfy[s].Ca);
      add_cx(m, acc, a, 0, i + 1, gfy[s].Ca);
      add_cx(m, acc, b, 0, i, gfy[s].Cb);
      add_cx(m, acc, b, 0, i + 1, gfy[s].Cb);
    }
  }
}

/* ---------------------------------------------------------------- potentials */
/* dwdphi(phi, divphi, a, smooth = false) transposed: gres times its derivative into gphi and Gamma */
template <int N, int NC>
GP_HD void dwdphi_vjp(const Model<double> &m, const double *phi, const double *divphi, int a, double gres,
                      double *gphi, RevAcc<N, NC> &acc) {
  if (m.fn_w == 1) {
    for (int b = 0; b < N; b++)
      if (b != a && fabs(divphi[b]) > 0.0) {
        const double sg = (phi[a] * phi[b]) >= 0.0 ? 1.0 : -1.0;
        gphi[b] += gres * m.w_obst * sg * m.gamma[a][b];
        acc.gamma(a, b) += gres * m.w_obst * sg * phi[b];
      }
  } else {
    for (int b = 0; b < N; b++)
      if (b != a && fabs(divphi[b]) > 0.0) {
        const double c = 9.0 * 2.0;
        gphi[b] += gres * c * m.gamma[a][b] * 2.0 * phi[b] * phi[a];
        gphi[a] += gres * c * m.gamma[a][b] * phi[b] * phi[b];
        acc.gamma(a, b) += gres * c * phi[b] * phi[b] * phi[a];
      }
  }
  for (int b = 0; b < N; b++)
    for (int c = 0; c < N; c++)
      if (b != a && c != a && b < c && fabs(divphi[b]) > 0.0 && fabs(divphi[c]) > 0.0) {
        const double gabc = m.gamma_abc[a][b][c];
        const double pbc = phi[b] * phi[c];
        const double sg = (phi[a] * pbc >= 0.0) ? 1.0 : -1.0;
        gphi[b] += gres * sg * gabc * phi[c];
        gphi[c] += gres * sg * gabc * phi[b];
      }
}

/* ---------------------------------------------------------------- tau(phi) */
template <int N, int NC>
GP_HD void tau_vjp(const Model<double> &m, const double *phi, double gtau, double *gphi, RevAcc<N, NC> &acc) {
  double s = 0.0, s1 = 0.0;
  for (int a = 0; a < N; a++)
    for (int b = a + 1; b < N; b++) {
      s += m.tau_ab[a][b] * phi[a] * phi[b];
      s1 += phi[a] * phi[b];
    }
  if (s1 != 0.0) {
    const double gs = gtau / s1, gs1 = -gtau * s / (s1 * s1);
    for (int a = 0; a < N; a++)
      for (int b = a + 1; b < N; b++) {
        acc.tau_ab(a, b) += gs * phi[a] * phi[b];
        gphi[a] += (gs * m.tau_ab[a][b] + gs1) * phi[b];
        gphi[b] += (gs * m.tau_ab[a][b] + gs1) * phi[a];
      }
  } else {
    acc.tau() += gtau;
  }
}

/* ---------------------------------------------------------------- psi_b(mu) */
/* psi_b = f_b(c_b) - mu . c_b, c_b = cmu_b (mu - Bmu_b): gpsi[b] times d psi_b / d mu into gmu */
template <int N, int NC>
GP_HD void psi_vjp(const Model<double> &m, const double *mu, const double *gpsi, double *gmu) {
  for (int b = 0; b < N; b++) {
    if (gpsi[b] == 0.0) continue;
    double c[NC];
    for (int k = 0; k < NC; k++) c[k] = c_mu<NC>(m, mu, b, k);
    double gc[NC];
    for (int i = 0; i < NC; i++) {
      double df = m.B[b][i];
      for (int j = 0; j < NC; j++) {
        if (i <= j) df += m.A[b][i][j] * c[j];
        if (j <= i) df += m.A[b][j][i] * c[j];
      }
      gc[i] = gpsi[b] * (df - mu[i]);
      gmu[i] -= gpsi[b] * c[i];
    }
    for (int i = 0; i < NC; i++)
      for (int j = 0; j < NC; j++) gmu[j] += gc[i] * m.cmu[b][i][j];
  }
}

/* ---------------------------------------------------------------- simplex projection */
/* project_simplex transposed. d0: delta-phi before the projection; g: in, the adjoint of the projected values;
 * out, the adjoint of d0. The adjoint with respect to phi (centre) is added to gphi. */
template <int N>
GP_HD void project_simplex_vjp(const double *phi, const double *divphi, const double *d0, double *g, double *gphi) {
  bool act[N], inA[N], clamp[N], inC[N];
  double d1[N], d2[N];
  int count = 0;
  double sum_phib = 0.0;
  for (int a = 0; a < N; a++) {
    act[a] = fabs(divphi[a]) > 0.0;
    inA[a] = act[a] && (phi[a] + d0[a]) > 0.0 && (phi[a] + d0[a]) < 1.0;
    if (inA[a]) {
      count++;
      sum_phib += phi[a];
    }
  }
  double delta = 0.0;
  for (int a = 0; a < N; a++) {
    clamp[a] = (phi[a] + d0[a]) < 0.0;
    if (clamp[a]) {
      const double dd = fabs(d0[a] + phi[a]);
      d1[a] = d0[a] + dd;
      delta += dd;
    } else {
      d1[a] = d0[a];
    }
  }
  const bool wsum = fabs(sum_phib) > 0.0;
  for (int b = 0; b < N; b++) {
    inC[b] = delta > 0.0 && act[b] && (phi[b] + d1[b]) > 0.0 && (phi[b] + d1[b]) < 1.0;
    d2[b] = d1[b];
    if (inC[b]) d2[b] -= wsum ? qdiv(delta * phi[b], sum_phib) : qdiv(delta, (double)count);
  }
  int astar = -1;
  for (int a = 0; a < N; a++)
    if (astar < 0 && act[a] && (phi[a] + d2[a]) > 1.0) astar = a;
  /* D: the first active phase above 1 is set to 1 - phi, the other active phases to -phi */
  if (astar >= 0)
    for (int b = 0; b < N; b++)
      if (b == astar || act[b]) {
        gphi[b] -= g[b];
        g[b] = 0.0;
      }
  /* C: d2 = d1 - delta phi_b / sum_phib (or delta / count). A value that D overwrote has no adjoint left and
   * contributes nothing; skipping it also keeps 0 / count = 0 / 0 out when no phase lay strictly inside (0, 1)
   * before the clamping (count = 0: the forward's delta / count is an infinity that D then replaces by -phi). */
  double gdelta = 0.0, gsum = 0.0;
  for (int b = 0; b < N; b++)
    if (inC[b] && g[b] != 0.0) {
      if (wsum) {
        gdelta -= g[b] * phi[b] / sum_phib;
        gphi[b] -= g[b] * delta / sum_phib;
        gsum += g[b] * delta * phi[b] / (sum_phib * sum_phib);
      } else {
        gdelta -= g[b] / (double)count;
      }
    }
  if (wsum)
    for (int a = 0; a < N; a++)
      if (inA[a]) gphi[a] += gsum;
  /* B: a clamped phase: d1 = d0 + |d0 + phi| = -phi, delta += -(d0 + phi) */
  for (int a = 0; a < N; a++)
    if (clamp[a]) {
      gphi[a] -= g[a] + gdelta;
      g[a] = -gdelta;
    }
}

/* ---------------------------------------------------------------- stage P: phi_delta transposed */
/* gout[N]: the adjoint of the (projected) delta-phi of the centre cell. Adds its transposed Jacobian to acc: phi on
 * the 3x3 stencil, mu at the centre (acc.m(k, 0, 0)), theta. Non-smoothing steps only. */
template <int N, int NC, class V>
GP_HD void phi_delta_vjp(const Model<double> &m, const V &v, const double *gout, RevAcc<N, NC> &acc) {
  double phi[N], mu[NC > 0 ? NC : 1], divphi[N], lambda[N], out[N];
  for (int a = 0; a < N; a++) phi[a] = v.ph(a, 0, 0);
  for (int k = 0; k < NC; k++) mu[k] = v.m(k, 0, 0);
  bool any = false;
  for (int a = 0; a < N; a++) {
    divphi[a] = divx<V>(m, (gx(m, v, a, 0, 0) - gx(m, v, a, -1, 0)));
    divphi[a] += divy<V>(m, (gy(m, v, a, 0, 0) - gy(m, v, a, 0, -1)));
    any = any || (fabs(divphi[a]) > 0.0);
  }
  double g[N], gphi[N];
  for (int a = 0; a < N; a++) {
    g[a] = gout[a];
    gphi[a] = 0.0;
    out[a] = 0.0;
  }
  if (!any) {
    project_simplex_vjp<N>(phi, divphi, out, g, gphi);
    for (int a = 0; a < N; a++) acc.ph(a, 0, 0) += gphi[a];
    return;
  }
  double cxc[N], cyc[N];
  for (int a = 0; a < N; a++) {
    cxc[a] = cx(m, v, a, 0, 0);
    cyc[a] = cy(m, v, a, 0, 0);
  }
  double psi[N];
  psi_all<N, NC>(m, mu, psi);
  double sum_lambda = 0.0;
  int active = 0;
  for (int a = 0; a < N; a++)
    if (fabs(divphi[a]) > 0.0) {
      double dA, dD;
      if (m.fn_aniso) {
        dA = dAdphi_A01<N>(m, v, phi, cxc, cyc, a);
        dD = divdAdgradphi_A01<N>(m, v, a);
      } else {
        dA = dAdphi_A00<N>(m, phi, cxc, cyc, a);
        dD = divdAdgradphi_iso<N>(m, v, a);
      }
      double dp = 0.0;
      for (int b = 0; b < N; b++) dp += psi[b] * dhphi<N>(phi, b, a);
      lambda[a] = m.eps * (-dA + dD) - m.inv_eps * dwdphi<N>(m, phi, divphi, a, false) - dp;
      sum_lambda += lambda[a];
      active++;
    }
  if (active) sum_lambda = qdiv(sum_lambda, (double)active);
  const double tauv = tau_of<N>(m, phi);
  for (int a = 0; a < N; a++)
    if (fabs(divphi[a]) > 0.0) out[a] = qdiv(m.dt * (lambda[a] - sum_lambda), (tauv * m.eps));
  /* projection */
  project_simplex_vjp<N>(phi, divphi, out, g, gphi);
  /* out_a = dt (lambda_a - mean) / (tau eps) */
  double glam[N], gmean = 0.0, gtau = 0.0;
  const double sc = m.dt / (tauv * m.eps);
  for (int a = 0; a < N; a++) {
    glam[a] = 0.0;
    if (fabs(divphi[a]) > 0.0) {
      glam[a] = g[a] * sc;
      gmean -= g[a] * sc;
      gtau -= g[a] * out[a] / tauv;
    }
  }
  tau_vjp<N, NC>(m, phi, gtau, gphi, acc);
  for (int a = 0; a < N; a++)
    if (fabs(divphi[a]) > 0.0) glam[a] += gmean / (double)active;
  /* lambda_a = eps (-dA + dD) - w' / eps - sum_b psi_b dh_b/dphi_a */
  double gcx[N], gcy[N], gpsi[N], gmu[NC > 0 ? NC : 1];
  for (int a = 0; a < N; a++) gcx[a] = gcy[a] = gpsi[a] = 0.0;
  for (int k = 0; k < NC; k++) gmu[k] = 0.0;
  for (int a = 0; a < N; a++) {
    if (!(fabs(divphi[a]) > 0.0) || glam[a] == 0.0) continue;
    const double gl = glam[a];
    if (m.fn_aniso) {
      dAdphi_A01_vjp<N, NC>(m, phi, cxc, cyc, a, -m.eps * gl, gphi, gcx, gcy, acc);
      divdA_vjp<N, NC>(m, v, a, true, m.eps * gl, acc);
    } else {
      dAdphi_A00_vjp<N>(m, phi, cxc, cyc, a, -m.eps * gl, gphi, gcx, gcy);
      dAdphi_A00_theta<N, NC>(m, phi, cxc, cyc, a, -m.eps * gl, acc);
      divdA_vjp<N, NC>(m, v, a, false, m.eps * gl, acc);
    }
    dwdphi_vjp<N, NC>(m, phi, divphi, a, -m.inv_eps * gl, gphi, acc);
    for (int b = 0; b < N; b++) {
      gpsi[b] -= gl * dhphi<N>(phi, b, a);
      for (int e = 0; e < N; e++) gphi[e] -= gl * psi[b] * d2hphi<N>(phi, b, a, e);
    }
  }
  psi_vjp<N, NC>(m, mu, gpsi, gmu);
  for (int k = 0; k < NC; k++) acc.m(k, 0, 0) += gmu[k];
  for (int a = 0; a < N; a++) {
    acc.ph(a, 0, 0) += gphi[a];
    add_cx(m, acc, a, 0, 0, gcx[a]);
    add_cy(m, acc, a, 0, 0, gcy[a]);
  }
}

/* ---------------------------------------------------------------- anti-trapping current */
/* s_obst(phi, a) transposed: gs times d s / d phi into gphi[N] (phi of that cell) */
template <int N>
GP_HD void s_obst_vjp(const double *phi, int a, double gs, double *gphi) {
  const double u = phi[a] * (1.0 - phi[a]);
  if (!(u > 0.0) || gs == 0.0) return;
  const double h = hphi<N>(phi, a);
  const double dn = sqrt(u), nn = phi[a] * (1.0 - h);
  const double ddn = (1.0 - 2.0 * phi[a]) / (2.0 * dn);
  gphi[a] += gs * ((1.0 - h) / dn - nn * ddn / (dn * dn));
  const double gh = -gs * phi[a] / dn;
  for (int e = 0; e < N; e++) gphi[e] += gh * dhphi<N>(phi, a, e);
}

/* jat_x (DIR 0: the +x face of the cell at offset (i, j), between it and (i+1, j)) or jat_y (DIR 1: the +y face,
 * between (i, j) and (i, j+1); yglob is that cell's allocated y) transposed: gj[NC] times the derivative of the face
 * current into acc (phi, mu, delta-phi of the two face cells and their tangential neighbours; the liquid's
 * diffusivities). */
template <int N, int NC, int DIR, class V>
GP_HD void jat_vjp(const Model<double> &m, const V &v0, int i, int j, int64_t yglob, const double *gj,
                   RevAcc<N, NC> &acc) {
  const int L = N - 1;
  const int di = DIR == 0 ? 1 : 0, dj = DIR == 0 ? 0 : 1; /* offset of the face's second cell */
  if (DIR == 1 && (yglob == 0 || yglob == m.jat_skip_row)) return;
  bool gz = true;
  for (int k = 0; k < NC; k++) gz = gz && gj[k] == 0.0;
  if (gz) return;
  const V v = v0.shift(i, j);
  bool any = false;
  for (int a = 0; a < N - 1; a++) any = any || (v.dp(a, 0, 0) != 0.0) || (v.dp(a, di, dj) != 0.0);
  if (!any) return;
  double phc[N], phf[N], muc[NC], muf[NC];
  for (int a = 0; a < N; a++) {
    phc[a] = v.ph(a, 0, 0);
    phf[a] = v.ph(a, di, dj);
  }
  for (int k = 0; k < NC; k++) {
    muc[k] = v.m(k, 0, 0);
    muf[k] = v.m(k, di, dj);
  }
  /* tangential (t) and normal (nn) gradients of a phase at the face, as jat_x / jat_y form them */
  auto tang = [&](int a) {
    return DIR == 0 ? 0.5 * (cy(m, v, a, 0, 0) + cy(m, v, a, 1, 0)) : 0.5 * (cx(m, v, a, 0, 0) + cx(m, v, a, 0, 1));
  };
  auto norm = [&](int a) { return DIR == 0 ? gx(m, v, a, 0, 0) : gy(m, v, a, 0, 0); };
  auto add_tang = [&](int a, double gv) {
    if (DIR == 0) {
      add_cy(m, acc, a, i, j, 0.5 * gv);
      add_cy(m, acc, a, i + 1, j, 0.5 * gv);
    } else {
      add_cx(m, acc, a, i, j, 0.5 * gv);
      add_cx(m, acc, a, i, j + 1, 0.5 * gv);
    }
  };
  auto add_norm = [&](int a, double gv) {
    if (DIR == 0) add_gx(m, acc, a, i, j, gv);
    else add_gy(m, acc, a, i, j, gv);
  };
  const double tl = tang(L), nL = norm(L);
  const double nxl = sqrt(tl * tl + nL * nL);
  double g_tl = 0.0, g_nL = 0.0, g_nxl = 0.0;
  double gphc[N], gphf[N];
  for (int a = 0; a < N; a++) gphc[a] = gphf[a] = 0.0;
  for (int a = 0; a < N - 1; a++) {
    const double ta = tang(a), na = norm(a);
    const double ng = sqrt(ta * ta + na * na);
    if (!(ng != 0.0)) continue;
    double s_c = 0.0, s_f = 0.0;
    if (m.fn_w == 1) {
      s_c = s_obst<N>(phc, a);
      s_f = s_obst<N>(phf, a);
    }
    const double dc = v.dp(a, 0, 0), df = v.dp(a, di, dj);
    double S = na * nL;
    S += ta * tl;
    const double sp = nxl > 0.0 ? qdiv(S, (nxl * ng)) : S;
    const double asp = fabs(sp), sgn = sp < 0.0 ? -1.0 : 1.0;
    double g_sp = 0.0, g_na = 0.0, g_ng = 0.0, g_sc = 0.0, g_sf = 0.0, g_dc = 0.0, g_df = 0.0;
    for (int k = 0; k < NC; k++) {
      const double gt = gj[k];
      if (gt == 0.0) continue;
      const double Df = c_mu<NC>(m, muf, L, k) - c_mu<NC>(m, muf, a, k);
      const double Dc = c_mu<NC>(m, muc, L, k) - c_mu<NC>(m, muc, a, k);
      double P, gDf, gDc;
      const double dr = m.dratio[a][k];
      if (m.fn_w == 2) {
        P = m.jat_well * (Df * df + Dc * dc);
      } else {
        P = 0.5 * (s_f * Df * df + s_c * Dc * dc);
      }
      const double term = dr * asp * P * na / ng;
      /* dratio = 1 - D_a / D_L */
      acc.dliq(k) += gt * asp * P * na / ng * (m.D[a][k][k] / (m.D[L][k][k] * m.D[L][k][k]));
      g_sp += gt * dr * sgn * P * na / ng;
      const double gP = gt * dr * asp * na / ng;
      g_na += gt * dr * asp * P / ng;
      g_ng -= gt * term / ng;
      if (m.fn_w == 2) {
        gDf = gP * m.jat_well * df;
        gDc = gP * m.jat_well * dc;
        g_df += gP * m.jat_well * Df;
        g_dc += gP * m.jat_well * Dc;
      } else {
        gDf = gP * 0.5 * s_f * df;
        gDc = gP * 0.5 * s_c * dc;
        g_sf += gP * 0.5 * Df * df;
        g_sc += gP * 0.5 * Dc * dc;
        g_df += gP * 0.5 * s_f * Df;
        g_dc += gP * 0.5 * s_c * Dc;
      }
      for (int q = 0; q < NC; q++) {
        const double w = m.cmu[L][k][q] - m.cmu[a][k][q];
        acc.m(q, i + di, j + dj) += gDf * w;
        acc.m(q, i, j) += gDc * w;
      }
    }
    /* sp = S / (nxl ng) when nxl > 0 */
    double gS = g_sp;
    if (nxl > 0.0) {
      gS = g_sp / (nxl * ng);
      g_nxl -= g_sp * sp / nxl;
      g_ng -= g_sp * sp / ng;
    }
    g_na += gS * nL;
    g_nL += gS * na;
    double g_ta = gS * tl;
    g_tl += gS * ta;
    g_ta += g_ng * ta / ng;
    g_na += g_ng * na / ng;
    acc.dp(a, i, j) += g_dc;
    acc.dp(a, i + di, j + dj) += g_df;
    s_obst_vjp<N>(phc, a, g_sc, gphc);
    s_obst_vjp<N>(phf, a, g_sf, gphf);
    add_tang(a, g_ta);
    add_norm(a, g_na);
  }
  if (nxl > 0.0) {
    g_tl += g_nxl * tl / nxl;
    g_nL += g_nxl * nL / nxl;
  }
  add_tang(L, g_tl);
  add_norm(L, g_nL);
  for (int a = 0; a < N; a++) {
    acc.ph(a, i, j) += gphc[a];
    acc.ph(a, i + di, j + dj) += gphf[a];
  }
}

/* ---------------------------------------------------------------- stage Q: currents and update transposed */
/* gpn[N], gmn[NC]: the adjoint of the new phi and mu of the centre cell. Adds the transposed Jacobian of
 * mu_update and of the four face currents to acc: phi on the 3x3 stencil, mu and delta-phi on the 5-point one, the
 * liquid's diffusivities. y is the centre cell's allocated y. Non-smoothing steps, FP64 model. */
template <int N, int NC, class V>
GP_HD void cell_update_vjp(const Model<double> &m, const V &v, int64_t y, const double *gpn, const double *gmn,
                           RevAcc<N, NC> &acc) {
  const int L = N - 1;
  double phi[N], mu[NC], dph[N];
  for (int a = 0; a < N; a++) {
    phi[a] = v.ph(a, 0, 0);
    dph[a] = v.dp(a, 0, 0);
  }
  for (int k = 0; k < NC; k++) mu[k] = v.m(k, 0, 0);
  /* phi_new = phi + dphi */
  for (int a = 0; a < N; a++) {
    acc.ph(a, 0, 0) += gpn[a];
    acc.dp(a, 0, 0) += gpn[a];
  }
  bool gz = true;
  for (int k = 0; k < NC; k++) gz = gz && gmn[k] == 0.0;
  if (gz) return;
  /* forward values of the mu update */
  double jx0[NC], jxb[NC], jy0[NC], jyl[NC];
  jat_x<N, NC>(m, v, 0, 0, jx0);
  jat_x<N, NC>(m, v, -1, 0, jxb);
  jat_y<N, NC>(m, v, 0, 0, y, jy0);
  jat_y<N, NC>(m, v, 0, -1, y - 1, jyl);
  double Dx[NC][NC], gmx[NC], Dxb[NC][NC], gmxb[NC], Dy[NC][NC], gmy[NC], Dyl[NC][NC], gmyl[NC];
  mu_face_x<N, NC>(m, v, 0, 0, Dx, gmx);
  mu_face_x<N, NC>(m, v, -1, 0, Dxb, gmxb);
  mu_face_y<N, NC>(m, v, 0, 0, Dy, gmy);
  mu_face_y<N, NC>(m, v, 0, -1, Dyl, gmyl);
  double divflux[NC], divjat[NC];
  for (int k = 0; k < NC; k++) {
    divflux[k] = 0.0;
    for (int l = 0; l < NC; l++) {
      divflux[k] += divx<V>(m, (Dx[k][l] * gmx[l] - Dxb[k][l] * gmxb[l]));
      divflux[k] += divy<V>(m, (Dy[k][l] * gmy[l] - Dyl[k][l] * gmyl[l]));
    }
    divjat[k] = divx<V>(m, (jx0[k] - jxb[k]) * m.jat_fac);
    divjat[k] += divy<V>(m, (jy0[k] - jyl[k]) * m.jat_fac);
  }
  int bulk = -1;
  for (int a = 0; a < N; a++)
    if (bulk < 0 && phi[a] == 1.0) bulk = a;
  double gdiv[NC], gjat[NC];
  if (bulk >= 0) {
    /* mu_new = Mu_b(c_b(mu) + dt divflux - divjat): Mu_b is affine with Hessian H_b */
    const int b = bulk;
    double gc[NC];
    for (int j = 0; j < NC; j++) {
      gc[j] = 0.0;
      for (int i = 0; i < NC; i++) gc[j] += gmn[i] * (i == j ? 2.0 * m.A[b][i][i] : m.A[b][i][j]);
    }
    for (int j = 0; j < NC; j++) {
      double s = 0.0;
      for (int k = 0; k < NC; k++) s += gc[k] * m.cmu[b][k][j];
      acc.m(j, 0, 0) += s;
    }
    for (int k = 0; k < NC; k++) {
      gdiv[k] = m.dt * gc[k];
      gjat[k] = -gc[k];
    }
  } else {
    double h[N], sdh[N], cm[N][NC], sum[NC], deltac[NC], dcdmu[NC][NC];
    for (int a = 0; a < N; a++) {
      h[a] = hphi<N>(phi, a);
      sdh[a] = 0.0;
      for (int b = 0; b < N; b++) sdh[a] += dhphi<N>(phi, a, b) * dph[b];
      for (int k = 0; k < NC; k++) cm[a][k] = c_mu<NC>(m, mu, a, k);
    }
    for (int k = 0; k < NC; k++) {
      sum[k] = 0.0;
      for (int a = 0; a < N; a++) sum[k] += cm[a][k] * sdh[a];
      deltac[k] = m.dt * (divflux[k]) - divjat[k] - sum[k];
      for (int l = 0; l < NC; l++) {
        dcdmu[k][l] = 0.0;
        for (int a = 0; a < N; a++) dcdmu[k][l] += m.cmu[a][k][l] * h[a];
      }
    }
    /* delta_mu = inv(dcdmu) deltac (the diagonal of dcdmu only in the DILUTE solve) */
    double gdc[NC], dmu[NC], gM[NC][NC];
    if (m.solve == GP_SOLVE_DILUTE) {
      for (int k = 0; k < NC; k++) {
        dmu[k] = deltac[k] / dcdmu[k][k];
        gdc[k] = gmn[k] / dcdmu[k][k];
      }
      for (int k = 0; k < NC; k++)
        for (int l = 0; l < NC; l++) gM[k][l] = k == l ? -gdc[k] * dmu[k] : 0.0;
    } else {
      double a_[NC][NC], inv[NC][NC];
      for (int k = 0; k < NC; k++)
        for (int l = 0; l < NC; l++) a_[k][l] = dcdmu[k][l];
      lu_inverse<NC>(a_, inv);
      for (int k = 0; k < NC; k++) {
        dmu[k] = 0.0;
        gdc[k] = 0.0;
        for (int l = 0; l < NC; l++) {
          dmu[k] += inv[k][l] * deltac[l];
          gdc[k] += inv[l][k] * gmn[l];
        }
      }
      for (int k = 0; k < NC; k++)
        for (int l = 0; l < NC; l++) gM[k][l] = -gdc[k] * dmu[l];
    }
    /* mu_new = mu + delta_mu */
    for (int k = 0; k < NC; k++) acc.m(k, 0, 0) += gmn[k];
    /* dcdmu = sum_a cmu_a h_a */
    double gphi[N];
    for (int e = 0; e < N; e++) gphi[e] = 0.0;
    for (int a = 0; a < N; a++) {
      double gh = 0.0;
      for (int k = 0; k < NC; k++)
        for (int l = 0; l < NC; l++) gh += gM[k][l] * m.cmu[a][k][l];
      for (int e = 0; e < N; e++) gphi[e] += gh * dhphi<N>(phi, a, e);
    }
    /* deltac = dt divflux - divjat - sum, sum_k = sum_a c_a,k(mu) sdh_a, sdh_a = sum_b dh_a/dphi_b dphi_b */
    for (int k = 0; k < NC; k++) {
      gdiv[k] = m.dt * gdc[k];
      gjat[k] = -gdc[k];
    }
    for (int a = 0; a < N; a++) {
      double gs = 0.0;
      for (int k = 0; k < NC; k++) gs -= gdc[k] * cm[a][k];
      for (int j = 0; j < NC; j++) {
        double t = 0.0;
        for (int k = 0; k < NC; k++) t -= gdc[k] * sdh[a] * m.cmu[a][k][j];
        acc.m(j, 0, 0) += t;
      }
      if (gs == 0.0) continue;
      for (int b = 0; b < N; b++) {
        acc.dp(b, 0, 0) += gs * dhphi<N>(phi, a, b);
        for (int e = 0; e < N; e++) gphi[e] += gs * dph[b] * d2hphi<N>(phi, a, b, e);
      }
    }
    for (int e = 0; e < N; e++) acc.ph(e, 0, 0) += gphi[e];
  }
  /* divflux: the face mobilities M(x) = sum_b sum_q DC[b][k][l][q] phi_b(x) and the mu differences */
  double gMc[5][NC][NC];
  for (int o = 0; o < 5; o++)
    for (int k = 0; k < NC; k++)
      for (int l = 0; l < NC; l++) gMc[o][k][l] = 0.0;
  for (int k = 0; k < NC; k++)
    for (int l = 0; l < NC; l++) {
      const double gxk = gdiv[k] * m.inv_dx, gyk = gdiv[k] * m.inv_dy;
      /* +x face (0,0)-(1,0), -x face (-1,0)-(0,0), +y face (0,0)-(0,1), -y face (0,-1)-(0,0) */
      gMc[1][k][l] += 0.5 * gxk * gmx[l];
      gMc[0][k][l] += 0.5 * gxk * gmx[l];
      gMc[0][k][l] -= 0.5 * gxk * gmxb[l];
      gMc[2][k][l] -= 0.5 * gxk * gmxb[l];
      gMc[3][k][l] += 0.5 * gyk * gmy[l];
      gMc[0][k][l] += 0.5 * gyk * gmy[l];
      gMc[0][k][l] -= 0.5 * gyk * gmyl[l];
      gMc[4][k][l] -= 0.5 * gyk * gmyl[l];
      const double gfx = gxk * Dx[k][l], gbx = gxk * Dxb[k][l], gfy = gyk * Dy[k][l], gly = gyk * Dyl[k][l];
      acc.m(l, 1, 0) += gfx * m.inv_dx;
      acc.m(l, 0, 0) -= gfx * m.inv_dx;
      acc.m(l, 0, 0) -= gbx * m.inv_dx;
      acc.m(l, -1, 0) += gbx * m.inv_dx;
      acc.m(l, 0, 1) += gfy * m.inv_dy;
      acc.m(l, 0, 0) -= gfy * m.inv_dy;
      acc.m(l, 0, 0) -= gly * m.inv_dy;
      acc.m(l, 0, -1) += gly * m.inv_dy;
    }
  const int oi[5] = {0, 1, -1, 0, 0}, oj[5] = {0, 0, 0, 1, -1};
  for (int o = 0; o < 5; o++) {
    const double pL = v.ph(L, oi[o], oj[o]);
    for (int k = 0; k < NC; k++)
      for (int l = 0; l < NC; l++) {
        const double gv = gMc[o][k][l];
        if (gv == 0.0) continue;
        for (int b = 0; b < N; b++) {
          double dcs = 0.0;
          for (int q = 0; q < NC; q++) dcs += m.DC[b][k][l][q];
          acc.ph(b, oi[o], oj[o]) += gv * dcs;
        }
        /* DC[L][k][l][k] = D[L][k][k] cmu[L][k][l] */
        acc.dliq(k) += gv * m.cmu[L][k][l] * pL;
      }
  }
  /* divjat = ((jx0 - jxb) / dx + (jy0 - jyl) / dy) jat_fac */
  double g0[NC], g1[NC], g2[NC], g3[NC];
  for (int k = 0; k < NC; k++) {
    g0[k] = gjat[k] * m.jat_fac * m.inv_dx;
    g1[k] = -g0[k];
    g2[k] = gjat[k] * m.jat_fac * m.inv_dy;
    g3[k] = -g2[k];
  }
  jat_vjp<N, NC, 0>(m, v, 0, 0, y, g0, acc);
  jat_vjp<N, NC, 0>(m, v, -1, 0, y, g1, acc);
  jat_vjp<N, NC, 1>(m, v, 0, 0, y, g2, acc);
  jat_vjp<N, NC, 1>(m, v, 0, -1, y - 1, g3, acc);
}

}  // namespace gpk
