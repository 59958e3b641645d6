/*
 * gp_model.h -- per-cell arithmetic of the 2-D grand-potential solver loop for the CDNA4 kernels.
 *
 * Header-only, __host__ __device__ when compiled by hipcc and plain inline C++ otherwise, so the identical
 * source is also compiled for the CPU by the test suite and compared with the FP64 CPU restatement.
 *
 * Every expression keeps the reference's operation order (and the build disables FMA contraction), so in FP64
 * the results are bit-identical to the serial driver (GPs = Grand_potential_Finite_difference_2D_serial):
 *   face gradients / averages   GPs/solverloop/calculate_gradients.h:92-185
 *   delta-phi                   GPs/solverloop/calculate_divergence_phasefield.h:4-62 (smooth: :63-)
 *   gradient energy             GPs/functions/function_A_01.h:4-219, function_A_00.h:4-58, functionQ.h:4-71,
 *                               anisotropy_01.h:9-48
 *   potentials                  GPs/functions/functionW_01.h:6-64, functionW_02.h:4-60
 *   free energy / dpsi          GPs/functions/functionF_01.h:5-57, functionH.h:1-37, functionTau.h:1-17
 *   simplex projection          GPs/solverloop/simplex_projection.h:1-90
 *   anti-trapping current       GPs/solverloop/calculate_fluxes_concentration.h:4-145
 *   mu update                   GPs/solverloop/calculate_divergence_concentration.h:4-153 (smooth: :154-253)
 *   mobility                    GPs/functions/functionD.h:1-16
 *   small LU                    GPs/functions/matrix.h:50-218 (pow(x,-1) taken as the correctly rounded 1/x)
 *
 * Cells are addressed through a View (stencil accessor relative to the centre cell), so the same functions run
 * on global memory or on an LDS tile.
 */
#pragma once

#include <stdint.h>
#include <string.h>

#include <type_traits>

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#define GP_HD __host__ __device__ __forceinline__
#else
#include <math.h>
#define GP_HD inline
#endif

#include "gp_hip.h"

namespace gpk {

constexpr int P = GP_MAXP;
constexpr int NCM = GP_MAXNC;
constexpr double kPi = 3.14159265358979323846;

/* Parameters as the kernels read them (one copy per handle in device memory). Values that the reference
 * recomputes identically per cell (1/eps, 2 dy, -pi eps/4, 1 - D_a/D_l, Beq + dBbdT (T - Teq)) are evaluated
 * once on the host with the same expression, which yields the same bits. */
template <typename R>
struct Model {
  R dx, dy, dt, eps, tau;
  R inv_eps;     /* 1.0/epsilon */
  R two_dy;      /* 2.0*deltay */
  R inv_dx, inv_dy, inv_2dy; /* exact reciprocals when dx, dy are powers of two (pow2 != 0) */
  int32_t pow2, pad2;
  R w_obst;      /* 16.0/(M_PI*M_PI) */
  R jat_well;    /* 0.5/sqrt(2.0) */
  R jat_fac;     /* -0.25*M_PI*epsilon (obstacle) or -1.0*epsilon (well) */
  int32_t N, NC, fn_w, fn_aniso, solve, jat_skip_row, pad0, pad1;
  R tau_ab[P][P], gamma[P][P], gamma_abc[P][P][P], dab[P][P];
  R rot[P][P][2][2], rot_inv[P][P][2][2];
  R D[P][NCM][NCM], A[P][NCM][NCM], cmu[P][NCM][NCM], B[P][NCM], C[P], Bmu[P][NCM];
  R dratio[P][NCM]; /* 1.0 - D[a][k][k]/D[L][k][k] */
  R DC[P][NCM][NCM][NCM]; /* D[b][k][q]*cmu[b][q][l] at [b][k][l][q]: the mobility product, formed once */
  R Mpure[P][NCM][NCM];   /* mobility of a cell that is purely phase b */
  /* FP32 bulk form: a pure-phase-b cell's mu increment is Kbulk[b] . lap(mu), Kbulk[b] = H_b dt Mpure[b] with
   * H_b the Hessian of Mu_lin (2 A[b][i][i] on the diagonal, A[b][i][j] off it), formed in FP64; lap_x = 1/dx^2,
   * lap_y = 1/dy^2 */
  R Kbulk[P][NCM][NCM];
  R lap_x, lap_y;
  /* non-isothermal runs: the dc/dT source of the interface mu update (calculate_divergence_concentration.h:82-85),
   * sum_a dcbdT[a][k] h_a times dt (-G v), with the serial driver's dcbdT_phase and temperature_gradientY.GRADIENT
   * (never assigned there: 0) and the Tempgrady velocity */
  R dcbdT[P][NCM];
  R tg_G, tg_vel;
};

/* Temperature-dependent coefficients of one row of cells (non-isothermal runs, where T depends on y only):
 * Beq + dBbdT (T - Teq), the offset of c(mu) and Mu(c) (functionF_01.h:19-37), and B(T), C(T), the free energy's
 * (function_B / function_C, functionF_01.h:62-104, set by init_propertymatrices(T) per cell). */
template <typename R>
struct RowThermo {
  R bmu[P][NCM];
  R b[P][NCM];
  R c[P];
  R T;
};

/* Stencil accessor over SoA planes: field a at (x+i, y+j) relative to the centre index c.
 * POW2: dx and dy are powers of two, so v/dx == v*(1/dx) exactly (the quotient and the product are the
 * same real number, rounded once) and the divisions by the spacings become multiplications. */
template <typename R, bool POW2 = false, bool TG = false, bool FAST = false>
struct View {
  static constexpr bool kPow2 = POW2;
  /* FAST: the anisotropy terms take one reciprocal of |q|^2 per q instead of the reference's five quotients
   * (aniso_ac_dadq_fast); for the modes whose results are not the reference's bit for bit */
  static constexpr bool kFast = FAST;
  /* TG: non-isothermal, th[y] holds the coefficients of allocated row y and y is the centre's row */
  static constexpr bool kTG = TG;
  const R *phi, *mu, *dphi;
  int64_t fs, ps, c;
  const RowThermo<R> *th = nullptr;
  int64_t y = 0;
  GP_HD R ph(int a, int i, int j) const { return phi[a * fs + c + i * ps + j]; }
  GP_HD R m(int k, int i, int j) const { return mu[k * fs + c + i * ps + j]; }
  GP_HD R dp(int a, int i, int j) const { return dphi[a * fs + c + i * ps + j]; }
  GP_HD View shift(int i, int j) const {
    View v = *this;
    v.c = c + i * ps + j;
    v.y = y + j;
    return v;
  }
};

/* Register-resident stencil of one cell: phi on the 3x3 neighbourhood, mu and delta-phi on the 5-point one,
 * gathered once (independent loads) so that the per-cell functions, which read the same values many times, work
 * from registers. A RegView is a position (si, sj) on the store; after inlining every index is a constant. The
 * per-cell functions read phi within one cell of the centre and mu and delta-phi on the 5-point stencil only
 * (offsets (0,0), (1,0), (-1,0), (0,1), (0,-1): slots 0..4). */
GP_HD int five_slot(int i, int j) {
  return (i == 0 && j == 0) ? 0 : (i == 1 && j == 0) ? 1 : (i == -1 && j == 0) ? 2 : (i == 0 && j == 1) ? 3 : 4;
}
template <int N, int NC, typename R>
struct RegStore {
  R p[N][3][3];
  R u[NC > 0 ? NC : 1][5];
  R d[N][5];
};
template <int N, int NC, typename R, bool POW2, bool FAST = false>
struct RegView {
  static constexpr bool kPow2 = POW2;
  static constexpr bool kTG = false;
  static constexpr bool kFast = FAST;
  const RegStore<N, NC, R> *s;
  int si, sj;
  GP_HD R ph(int a, int i, int j) const { return s->p[a][si + i + 1][sj + j + 1]; }
  GP_HD R m(int k, int i, int j) const { return s->u[k][five_slot(si + i, sj + j)]; }
  GP_HD R dp(int a, int i, int j) const { return s->d[a][five_slot(si + i, sj + j)]; }
  GP_HD RegView shift(int i, int j) const { return RegView{s, si + i, sj + j}; }
};

/* views of the FP32 mode take the fast anisotropy forms (View::kFast) */
template <typename R>
constexpr bool kFastOf = std::is_same<R, float>::value;

/* Quotient and square root of the step arithmetic (the one place that chooses their form). FP64: IEEE,
 * correctly rounded, as the reference computes them, so the FP64 results are bit-identical to it. FP32 (the
 * production mode) takes the device build's fast single-precision forms (hipcc
 * -fno-hip-fp32-correctly-rounded-divide-sqrt: reciprocal-based quotient and hardware square root of about
 * 1-2.5 ulp, with denormal operands rescaled), about half the instructions of the correctly rounded
 * sequences; the FP32 mode's error is set by its FP32 storage, far above that. A bare v_rcp_f32 is not used:
 * it flushes denormal operands, and FP32 denominators such as |q|^6 of the anisotropy derivative reach the
 * denormal range. */
template <typename R>
GP_HD R qdiv(R a, R b) {
  return a / b;
}
template <typename R>
GP_HD R qsqrt(R a) {
  return sqrt(a);
}

/* v/dx, v/dy, v/(2 dy) */
template <class V, typename R>
GP_HD R divx(const Model<R> &m, R v) {
  if constexpr (V::kPow2) return v * m.inv_dx;
  else return qdiv(v, m.dx);
}
template <class V, typename R>
GP_HD R divy(const Model<R> &m, R v) {
  if constexpr (V::kPow2) return v * m.inv_dy;
  else return qdiv(v, m.dy);
}
template <class V, typename R>
GP_HD R div2y(const Model<R> &m, R v) {
  if constexpr (V::kPow2) return v * m.inv_2dy;
  else return qdiv(v, m.two_dy);
}

/* ------------------------------------------------------------------ face quantities at offset (i,j) */
template <typename R, class V>
GP_HD R gx(const Model<R> &m, const V &v, int a, int i, int j) { return divx<V>(m, v.ph(a, i + 1, j) - v.ph(a, i, j)); }
template <typename R, class V>
GP_HD R sx(const Model<R> &, const V &v, int a, int i, int j) { return R(0.5) * (v.ph(a, i + 1, j) + v.ph(a, i, j)); }
template <typename R, class V>
GP_HD R gy(const Model<R> &m, const V &v, int a, int i, int j) { return divy<V>(m, v.ph(a, i, j + 1) - v.ph(a, i, j)); }
template <typename R, class V>
GP_HD R sy(const Model<R> &, const V &v, int a, int i, int j) { return R(0.5) * (v.ph(a, i, j + 1) + v.ph(a, i, j)); }
template <typename R, class V>
GP_HD R cy(const Model<R> &m, const V &v, int a, int i, int j) { return div2y<V>(m, v.ph(a, i, j + 1) - v.ph(a, i, j - 1)); }
template <typename R, class V>
GP_HD R cx(const Model<R> &m, const V &v, int a, int i, int j) { return R(0.5) * (gx(m, v, a, i, j) + gx(m, v, a, i - 1, j)); }

/* ------------------------------------------------------------------ thermodynamics */
template <int N, typename R>
GP_HD R hphi(const R *phi, int a) {
  R s1 = R(0.0);
  R s = R(3.0) * phi[a] * phi[a] - R(2.0) * phi[a] * phi[a] * phi[a];
  for (int b = 0; b < N; b++)
    for (int c = 0; c < N; c++)
      if (b != a && c != a && b < c) s1 += phi[b] * phi[c];
  s1 *= R(2.0) * phi[a];
  return s + s1;
}

/* dh_b / dphi_a */
template <int N, typename R>
GP_HD R dhphi(const R *phi, int b, int a) {
  R s = R(0.0);
  if (a == b) {
    s = R(6.0) * phi[a] * (R(1.0) - phi[a]);
    for (int e = 0; e < N; e++)
      for (int f = 0; f < N; f++)
        if (e != b && f != b && e < f) s += R(2.0) * phi[e] * phi[f];
  } else {
    for (int e = 0; e < N; e++)
      if (e != b && e != a) s += R(2.0) * phi[e];
    s *= phi[b];
  }
  return s;
}

/* c(mu) and Mu(c) with the offsets bmu[a][k] = Beq + dBbdT (T - Teq) of the cell's temperature */
template <int NC, typename R>
GP_HD R c_mu_b(const Model<R> &m, const R (*bmu)[NCM], const R *mu, int a, int i) {
  R s = R(0.0);
  for (int k = 0; k < NC; k++) s += m.cmu[a][i][k] * (mu[k] - bmu[a][k]);
  return s;
}
template <int NC, typename R>
GP_HD R c_mu(const Model<R> &m, const R *mu, int a, int i) {
  return c_mu_b<NC>(m, m.Bmu, mu, a, i);
}

template <int NC, typename R>
GP_HD R Mu_b(const Model<R> &m, const R (*bmu)[NCM], const R *c, int a, int i) {
  R s = R(2.0) * m.A[a][i][i] * c[i] + bmu[a][i];
  for (int j = 0; j < NC; j++)
    if (i != j) s += m.A[a][i][j] * c[j];
  return s;
}
template <int NC, typename R>
GP_HD R Mu(const Model<R> &m, const R *c, int a, int i) {
  return Mu_b<NC>(m, m.Bmu, c, a, i);
}

/* the offsets of the row at y offset j from a view's centre: the row's own (non-isothermal views) or the model's */
template <class V, typename R>
GP_HD auto row_bmu(const Model<R> &m, const V &v, int j) -> const R (*)[NCM] {
  if constexpr (V::kTG) return v.th[v.y + j].bmu;
  else return m.Bmu;
}

/* Linear part of Mu: the change of mu_i for a change dc of the composition in phase a (Mu is affine in c). */
template <int NC, typename R>
GP_HD R Mu_lin(const Model<R> &m, const R *dc, int a, int i) {
  R s = R(2.0) * m.A[a][i][i] * dc[i];
  for (int j = 0; j < NC; j++)
    if (i != j) s += m.A[a][i][j] * dc[j];
  return s;
}

/* Whether the bulk-cell mu increment takes the direct form dmu = Mu_lin(dc). The reference's form
 * (c_new = c(mu) + dc, mu_new = Mu(c_new); calculate_divergence_concentration.h:71-79) is kept in FP64 for
 * bit-exact parity. In FP32 its mu -> c -> mu round trip costs about one ulp of mu per step in every bulk cell,
 * independent of the increment, and those errors random-walk through a long run (about 20x the phi error after
 * 2000 steps of the ternary seed workload); the direct form, identical in exact arithmetic, rounds relative to
 * the increment only. */
template <typename R>
constexpr bool kDirectBulkMu = sizeof(R) < sizeof(double);

template <int NC, typename R>
GP_HD R free_energy_b(const Model<R> &m, const R (*B)[NCM], const R *C, const R *c, int a) {
  R s = R(0.0);
  for (int i = 0; i < NC; i++) {
    for (int j = 0; j < NC; j++)
      if (i <= j) s += m.A[a][i][j] * c[i] * c[j];
    s += B[a][i] * c[i];
  }
  s += C[a];
  return s;
}
template <int NC, typename R>
GP_HD R free_energy(const Model<R> &m, const R *c, int a) {
  return free_energy_b<NC>(m, m.B, m.C, c, a);
}

/* psi_b(mu) = f_b(c_b(mu)) - mu.c_b(mu) for every phase (the phase-independent part of dpsi), with the given
 * temperature-dependent coefficients */
template <int N, int NC, typename R>
GP_HD void psi_all_b(const Model<R> &m, const R (*bmu)[NCM], const R (*B)[NCM], const R *C, const R *mu, R *psi) {
  for (int b = 0; b < N; b++) {
    R c[NC];
    R p = R(0.0);
    for (int k = 0; k < NC; k++) {
      c[k] = c_mu_b<NC>(m, bmu, mu, b, k);
      p -= mu[k] * c[k];
    }
    p += free_energy_b<NC>(m, B, C, c, b);
    psi[b] = p;
  }
}
template <int N, int NC, typename R>
GP_HD void psi_all(const Model<R> &m, const R *mu, R *psi) {
  psi_all_b<N, NC>(m, m.Bmu, m.B, m.C, mu, psi);
}

template <int N, typename R>
GP_HD R tau_of(const Model<R> &m, const R *phi) {
  R s = R(0.0), s1 = R(0.0);
  for (int a = 0; a < N; a++)
    for (int b = 0; b < N; b++)
      if (a < b) {
        s += m.tau_ab[a][b] * phi[a] * phi[b];
        s1 += phi[a] * phi[b];
      }
  if (s1 != R(0.0)) return qdiv(s, s1);
  return m.tau;
}

/* ------------------------------------------------------------------ anisotropy (4-fold) */
template <typename R>
GP_HD void rot2(const R (*Rm)[2], const R *v, R *o) {
  for (int i = 0; i < 2; i++) {
    R s = R(0);
    for (int j = 0; j < 2; j++) s = s + Rm[i][j] * v[j];
    o[i] = s;
  }
}

template <typename R>
GP_HD R aniso_ac(R dab, const R *q) {
  R qx2 = q[0] * q[0], qx4 = qx2 * qx2;
  R qy2 = q[1] * q[1], qy4 = qy2 * qy2;
  R qz2 = R(0.0), qz4 = qz2 * qz2;
  R q2 = qx2 + qy2 + qz2;
  if (fabs(q2) > R(1.0e-15)) return R(1.0) - dab * (R(3.0) - qdiv(R(4.0) * (qx4 + qy4 + qz4), (q2 * q2)));
  return R(1.0);
}

template <typename R>
GP_HD void aniso_dadq(R dab, const R *q, R *d) {
  R qz = R(0.0);
  R qx3 = q[0] * q[0] * q[0], qy3 = q[1] * q[1] * q[1], qz3 = qz * qz * qz;
  R q2 = q[0] * q[0] + q[1] * q[1] + qz * qz;
  R q22 = q2 * q2, q23 = q2 * q2 * q2;
  R q4 = qx3 * q[0] + qy3 * q[1] + qz3 * qz;
  if (fabs(q2) > R(1.0e-15)) {
    d[0] = R(16.0) * dab * (qdiv(qx3, q22) - qdiv(q[0] * q4, q23));
    d[1] = R(16.0) * dab * (qdiv(qy3, q22) - qdiv(q[1] * q4, q23));
  } else {
    d[0] = R(0.0);
    d[1] = R(0.0);
  }
}

/* a_c and d a_c/dq together from one reciprocal of |q|^2: a_c = 1 - dab (3 - 4 Q4 r^2), d a_c/dq_i =
 * 16 dab r^2 q_i (q_i^2 - Q4 r), r = 1/|q|^2 (the same real values as aniso_ac and aniso_dadq, within a few ulp) */
template <typename R>
GP_HD void aniso_ac_dadq_fast(R dab, const R *q, R &ac, R *d) {
  const R qx2 = q[0] * q[0], qy2 = q[1] * q[1];
  const R q2 = qx2 + qy2;
  if (fabs(q2) > R(1.0e-15)) {
    const R r = R(1.0) / q2, r2 = r * r;
    const R q4 = qx2 * qx2 + qy2 * qy2;
    const R t = q4 * r, f = R(16.0) * dab * r2;
    ac = R(1.0) - dab * (R(3.0) - R(4.0) * q4 * r2);
    d[0] = f * q[0] * (qx2 - t);
    d[1] = f * q[1] * (qy2 - t);
  } else {
    ac = R(1.0);
    d[0] = R(0.0);
    d[1] = R(0.0);
  }
}

template <typename R>
GP_HD R qnorm2(const R *q) { return q[0] * q[0] + q[1] * q[1] + R(0.0) * R(0.0); }

/* q_ab on the x-face between offsets (i,j) and (i+1,j), rotated */
template <typename R, class V>
GP_HD void q_face_x(const Model<R> &m, const V &v, int a, int b, int i, int j, R *q) {
  R r[2];
  R sxa = sx(m, v, a, i, j), sxb = sx(m, v, b, i, j);
  r[0] = (sxa * gx(m, v, b, i, j) - sxb * gx(m, v, a, i, j));
  r[1] = R(0.5) * (sxa * (cy(m, v, b, i, j) + cy(m, v, b, i + 1, j)) - sxb * (cy(m, v, a, i, j) + cy(m, v, a, i + 1, j)));
  rot2(m.rot[a][b], r, q);
}

/* q_ab on the y-face between offsets (i,j) and (i,j+1), rotated */
template <typename R, class V>
GP_HD void q_face_y(const Model<R> &m, const V &v, int a, int b, int i, int j, R *q) {
  R r[2];
  R sya = sy(m, v, a, i, j), syb = sy(m, v, b, i, j);
  r[0] = R(0.5) * (sya * (cx(m, v, b, i, j) + cx(m, v, b, i, j + 1)) - syb * (cx(m, v, a, i, j) + cx(m, v, a, i, j + 1)));
  r[1] = (sya * gy(m, v, b, i, j) - syb * gy(m, v, a, i, j));
  rot2(m.rot[a][b], r, q);
}

template <int N, typename R, class V>
GP_HD R dAdphi_A01(const Model<R> &m, const V &v, const R *phi, const R *cxc, const R *cyc, int a) {
  R sum1 = R(0.0), sum2 = R(0.0);
  for (int b = 0; b < N; b++) {
    if (b == a) continue;
    R r[2], q[2], dadq[2], dq[2], dqr[2];
    r[0] = (phi[a] * cxc[b] - phi[b] * cxc[a]);
    r[1] = (phi[a] * cyc[b] - phi[b] * cyc[a]);
    rot2(m.rot[a][b], r, q);
    R ac1;
    if constexpr (V::kFast) {
      aniso_ac_dadq_fast(m.dab[a][b], q, ac1, dadq);
    } else {
      aniso_dadq(m.dab[a][b], q, dadq);
      ac1 = aniso_ac(m.dab[a][b], q);
    }
    R qab2 = qnorm2(q);
    R ac2 = ac1 * ac1;
    dq[0] = cxc[b];
    dq[1] = cyc[b];
    rot2(m.rot[a][b], dq, dqr);
    R s1 = R(0.0), s2 = R(0.0);
    for (int d = 0; d < 2; d++) s1 += q[d] * dqr[d];
    for (int d = 0; d < 2; d++) s2 += dadq[d] * dqr[d];
    s1 *= m.gamma[a][b] * ac2;
    s2 *= m.gamma[a][b] * ac1 * qab2;
    sum1 += s1;
    sum2 += s2;
  }
  return R(2.0) * (sum1 + sum2);
}

template <int N, typename R, class V>
GP_HD R divdAdgradphi_A01(const Model<R> &m, const V &v, int a) {
  R sum1 = R(0.0), sum2 = R(0.0);
  for (int b = 0; b < N; b++) {
    if (b == a) continue;
    R qf[2], qb[2], qr[2], ql[2];
    q_face_x(m, v, a, b, 0, 0, qf);
    q_face_x(m, v, a, b, -1, 0, qb);
    q_face_y(m, v, a, b, 0, 0, qr);
    q_face_y(m, v, a, b, 0, -1, ql);
    const R dab = m.dab[a][b];
    R acf, acb, acr, acl, df[2], db[2], dr[2], dl[2];
    if constexpr (V::kFast) {
      aniso_ac_dadq_fast(dab, qf, acf, df);
      aniso_ac_dadq_fast(dab, qb, acb, db);
      aniso_ac_dadq_fast(dab, qr, acr, dr);
      aniso_ac_dadq_fast(dab, ql, acl, dl);
    } else {
      acf = aniso_ac(dab, qf);
      acb = aniso_ac(dab, qb);
      acr = aniso_ac(dab, qr);
      acl = aniso_ac(dab, ql);
      aniso_dadq(dab, qf, df);
      aniso_dadq(dab, qb, db);
      aniso_dadq(dab, qr, dr);
      aniso_dadq(dab, ql, dl);
    }
    R acf2 = acf * acf, acb2 = acb * acb, acr2 = acr * acr, acl2 = acl * acl;
    R sxa = sx(m, v, a, 0, 0), sxb = sx(m, v, b, 0, 0), gxa = gx(m, v, a, 0, 0), gxb = gx(m, v, b, 0, 0);
    R Bsxa = sx(m, v, a, -1, 0), Bsxb = sx(m, v, b, -1, 0), Bgxa = gx(m, v, a, -1, 0), Bgxb = gx(m, v, b, -1, 0);
    R sya = sy(m, v, a, 0, 0), syb = sy(m, v, b, 0, 0), gya = gy(m, v, a, 0, 0), gyb = gy(m, v, b, 0, 0);
    R Lsya = sy(m, v, a, 0, -1), Lsyb = sy(m, v, b, 0, -1), Lgya = gy(m, v, a, 0, -1), Lgyb = gy(m, v, b, 0, -1);
    R d1 = divx<V>(m, ((acf2 * (sxa * gxb - sxb * gxa) * (-sxb)) - (acb2 * (Bsxa * Bgxb - Bsxb * Bgxa) * (-Bsxb))));
    d1 += divy<V>(m, ((acr2 * (sya * gyb - syb * gya) * (-syb)) - (acl2 * (Lsya * Lgyb - Lsyb * Lgya) * (-Lsyb))));
    d1 *= m.gamma[a][b];
    sum1 += d1;
    R rf[2], rb[2], rr[2], rl[2];
    rot2(m.rot_inv[a][b], df, rf);
    rot2(m.rot_inv[a][b], db, rb);
    rot2(m.rot_inv[a][b], dr, rr);
    rot2(m.rot_inv[a][b], dl, rl);
    R qf2 = qnorm2(qf), qb2 = qnorm2(qb), qr2 = qnorm2(qr), ql2 = qnorm2(ql);
    R d2 = divx<V>(m, (acf * qf2 * rf[0] * (-sxb) - acb * qb2 * rb[0] * (-Bsxb)));
    d2 += divy<V>(m, (acr * qr2 * rr[1] * (-syb) - acl * ql2 * rl[1] * (-Lsyb)));
    d2 *= m.gamma[a][b];
    sum2 += d2;
  }
  return (R(2.0) * (sum1 + sum2) + R(0.0));
}

template <int N, typename R>
GP_HD R dAdphi_A00(const Model<R> &m, const R *phi, const R *cxc, const R *cyc, int a) {
  R sum = R(0.0);
  for (int b = 0; b < N; b++) {
    if (b == a) continue;
    R s = R(0.0);
    s += (phi[a] * cxc[b] - phi[b] * cxc[a]) * cxc[b];
    s += (phi[a] * cyc[b] - phi[b] * cyc[a]) * cyc[b];
    s *= m.gamma[a][b];
    sum += s;
  }
  return R(2.0) * sum;
}

/* isotropic divergence; `smooth` selects the A_01 smoothing form (same value, +0.0 from its zero sum2) */
template <int N, typename R, class V>
GP_HD R divdAdgradphi_iso(const Model<R> &m, const V &v, int a) {
  R sum = R(0.0);
  for (int b = 0; b < N; b++) {
    if (b == a) continue;
    R sxa = sx(m, v, a, 0, 0), sxb = sx(m, v, b, 0, 0), gxa = gx(m, v, a, 0, 0), gxb = gx(m, v, b, 0, 0);
    R Bsxa = sx(m, v, a, -1, 0), Bsxb = sx(m, v, b, -1, 0), Bgxa = gx(m, v, a, -1, 0), Bgxb = gx(m, v, b, -1, 0);
    R sya = sy(m, v, a, 0, 0), syb = sy(m, v, b, 0, 0), gya = gy(m, v, a, 0, 0), gyb = gy(m, v, b, 0, 0);
    R Lsya = sy(m, v, a, 0, -1), Lsyb = sy(m, v, b, 0, -1), Lgya = gy(m, v, a, 0, -1), Lgyb = gy(m, v, b, 0, -1);
    R d = divx<V>(m, ((sxa * gxb - sxb * gxa) * (-sxb) - ((Bsxa * Bgxb - Bsxb * Bgxa) * (-Bsxb))));
    d += divy<V>(m, ((sya * gyb - syb * gya) * (-syb) - ((Lsya * Lgyb - Lsyb * Lgya) * (-Lsyb))));
    d *= m.gamma[a][b];
    sum += d;
  }
  return R(2.0) * sum;
}

/* d w / d phi_a; smooth: the *_smooth forms (no sign test, Gamma_abc phi_b phi_c association) */
template <int N, typename R>
GP_HD R dwdphi(const Model<R> &m, const R *phi, const R *divphi, int a, bool smooth) {
  R sum = R(0.0);
  if (m.fn_w == 1) {
    for (int b = 0; b < N; b++)
      if (b != a && fabs(divphi[b]) > R(0.0)) {
        if (smooth) sum += m.gamma[a][b] * phi[b];
        else if ((phi[a] * phi[b]) >= R(0.0)) sum += m.gamma[a][b] * (phi[b]);
        else sum += -m.gamma[a][b] * (phi[b]);
      }
    sum *= m.w_obst;
  } else {
    for (int b = 0; b < N; b++)
      if (b != a && fabs(divphi[b]) > R(0.0)) sum += R(2.0) * m.gamma[a][b] * (phi[b] * phi[b] * phi[a]);
    sum *= R(9.0);
  }
  for (int b = 0; b < N; b++)
    for (int c = 0; c < N; c++)
      if (b != a && c != a && b < c && fabs(divphi[b]) > R(0.0) && fabs(divphi[c]) > R(0.0)) {
        const R g = m.gamma_abc[a][b][c];
        if (smooth) {
          sum += g * phi[b] * phi[c];
        } else {
          R pbc = phi[b] * phi[c];
          if (phi[a] * pbc >= R(0.0)) sum += g * pbc;
          else sum += -g * pbc;
        }
      }
  return sum;
}

template <int N, typename R>
GP_HD void project_simplex(const R *phi, const R *divphi, R *dphi) {
  R Deltaphi = R(0.0), sum_phib = R(0.0);
  int count = 0;
  for (int a = 0; a < N; a++)
    if ((fabs(divphi[a]) > R(0.0)) && (phi[a] + dphi[a]) > R(0.0) && (phi[a] + dphi[a]) < R(1.0)) {
      count++;
      sum_phib += phi[a];
    }
  for (int a = 0; a < N; a++)
    if ((phi[a] + dphi[a]) < R(0.0)) {
      R d = fabs(dphi[a] + phi[a]);
      dphi[a] += d;
      Deltaphi += d;
    }
  if (Deltaphi > R(0.0))
    for (int b = 0; b < N; b++)
      if ((fabs(divphi[b]) > R(0.0)) && ((phi[b] + dphi[b]) > R(0.0)) && ((phi[b] + dphi[b]) < R(1.0))) {
        if (fabs(sum_phib) > R(0.0)) dphi[b] -= qdiv(Deltaphi * (phi[b]), (sum_phib));
        else dphi[b] -= qdiv(Deltaphi, R(count));
      }
  bool done = false;
  for (int a = 0; a < N; a++)
    if (!done && fabs(divphi[a]) > R(0.0) && (phi[a] + dphi[a]) > R(1.0)) {
      dphi[a] = (R(1.0) - phi[a]);
      for (int b = 0; b < N; b++)
        if ((fabs(divphi[b]) > R(0.0)) && b != a) dphi[b] = -phi[b];
      done = true;
    }
}

template <int N, typename R>
GP_HD void project_simplex_ww(const R *phi, const R *divphi, R *dphi) {
  int count = 0;
  for (int a = 0; a < N; a++)
    if ((fabs(divphi[a]) > R(0.0)) && (phi[a] + dphi[a]) > R(0.0) && (phi[a] + dphi[a]) < R(1.0)) count++;
  bool done = false;
  for (int a = 0; a < N; a++) {
    if (done || !(fabs(divphi[a]) > R(0.0))) continue;
    if ((phi[a] + dphi[a]) > R(1.0)) {
      dphi[a] = (R(1.0) - phi[a]);
      for (int b = 0; b < N; b++)
        if ((fabs(divphi[b]) > R(0.0)) && b != a) dphi[b] = -phi[b];
      done = true;
      continue;
    }
    if ((phi[a] + dphi[a]) < R(0.0)) {
      R d = fabs(dphi[a] + phi[a]);
      dphi[a] += d;
      for (int b = 0; b < N; b++)
        if (((fabs(divphi[b]) > R(0.0)) && b != a) && (phi[b] + dphi[b]) > R(0.0) && (phi[b] + dphi[b]) < R(1.0))
          dphi[b] -= qdiv(d, R(count));
    }
  }
}

/* ------------------------------------------------------------------ delta-phi of the centre cell */
template <int N, int NC, typename R, class V>
GP_HD void phi_delta(const Model<R> &m, const V &v, bool smooth, R *out) {
  R phi[N], mu[NC > 0 ? NC : 1], divphi[N], lambda[N];
  for (int a = 0; a < N; a++) phi[a] = v.ph(a, 0, 0);
  for (int k = 0; k < NC; k++) mu[k] = v.m(k, 0, 0);
  bool any = false;
  for (int a = 0; a < N; a++) {
    divphi[a] = divx<V>(m, (gx(m, v, a, 0, 0) - gx(m, v, a, -1, 0)));
    divphi[a] += divy<V>(m, (gy(m, v, a, 0, 0) - gy(m, v, a, 0, -1)));
    any = any || (fabs(divphi[a]) > R(0.0));
  }
  if (!any) {
    /* no lambda, no tau; the projection still clamps a negative phi (its < 0 test is not gated) */
    for (int a = 0; a < N; a++) out[a] = R(0.0);
    if (smooth) project_simplex_ww<N>(phi, divphi, out);
    else project_simplex<N>(phi, divphi, out);
    return;
  }
  R cxc[N], cyc[N];
  for (int a = 0; a < N; a++) {
    cxc[a] = cx(m, v, a, 0, 0);
    cyc[a] = cy(m, v, a, 0, 0);
  }
  R psi[N];
  if (!smooth) {
    if constexpr (V::kTG) {
      const RowThermo<R> &t = v.th[v.y];
      psi_all_b<N, NC>(m, t.bmu, t.b, t.c, mu, psi);
    } else {
      psi_all<N, NC>(m, mu, psi);
    }
  }
  R sum_lambda = R(0.0);
  int active = 0;
  for (int a = 0; a < N; a++) {
    if (fabs(divphi[a]) > R(0.0)) {
      if (!smooth) {
        R dA, dD;
        if (m.fn_aniso) {
          dA = dAdphi_A01<N>(m, v, phi, cxc, cyc, a);
          dD = divdAdgradphi_A01<N>(m, v, a);
        } else {
          dA = dAdphi_A00<N>(m, phi, cxc, cyc, a);
          dD = divdAdgradphi_iso<N>(m, v, a);
        }
        R dp = R(0.0);
        for (int b = 0; b < N; b++) dp += psi[b] * dhphi<N>(phi, b, a);
        lambda[a] = m.eps * (-dA + dD) - m.inv_eps * dwdphi<N>(m, phi, divphi, a, false) - dp;
      } else {
        /* smoothing: the anisotropic set uses the isotropic forms and the smooth potential */
        R dA = dAdphi_A00<N>(m, phi, cxc, cyc, a);
        R dD = divdAdgradphi_iso<N>(m, v, a);
        lambda[a] = m.eps * (-dA + dD) - m.inv_eps * dwdphi<N>(m, phi, divphi, a, m.fn_aniso != 0);
      }
      sum_lambda += lambda[a];
      active++;
    }
  }
  if (active) sum_lambda = qdiv(sum_lambda, R(active));
  const R tauv = tau_of<N>(m, phi);
  for (int a = 0; a < N; a++) {
    if (fabs(divphi[a]) > R(0.0)) out[a] = qdiv(m.dt * (lambda[a] - sum_lambda), (tauv * m.eps));
    else out[a] = R(0.0);
  }
  if (smooth) project_simplex_ww<N>(phi, divphi, out);
  else project_simplex<N>(phi, divphi, out);
}

/* ------------------------------------------------------------------ anti-trapping current */
template <int N, typename R>
GP_HD R s_obst(const R *phi, int a) {
  if (phi[a] * (R(1.0) - phi[a]) > R(0.0)) return qdiv(phi[a] * (R(1.0) - hphi<N>(phi, a)), (qsqrt(phi[a] * (R(1.0) - phi[a]))));
  return R(0.0);
}

/* j_at on the +x face (between centre and (1,0)) of the cell at offset (i,j) */
template <int N, int NC, typename R, class V>
GP_HD void jat_x(const Model<R> &m, const V &v0, int i, int j, R *jx) {
  const V v = v0.shift(i, j);
  const int L = N - 1;
  for (int k = 0; k < NC; k++) jx[k] = R(0.0);
  /* Every term carries a factor delta-phi_a of one of the two face cells: with both zero each term is a
   * signed zero and j stays +0, so the face is skipped (same value). */
  bool any = false;
  for (int a = 0; a < N - 1; a++) any = any || (v.dp(a, 0, 0) != R(0.0)) || (v.dp(a, 1, 0) != R(0.0));
  if (!any) return;
  R phc[N], phf[N], muc[NC], muf[NC];
  for (int a = 0; a < N; a++) {
    phc[a] = v.ph(a, 0, 0);
    phf[a] = v.ph(a, 1, 0);
  }
  for (int k = 0; k < NC; k++) {
    muc[k] = v.m(k, 0, 0);
    muf[k] = v.m(k, 1, 0);
  }
  R gyl = R(0.5) * (cy(m, v, L, 0, 0) + cy(m, v, L, 1, 0));
  R gxL = gx(m, v, L, 0, 0);
  R nxl = qsqrt(gyl * gyl + gxL * gxL);
  const auto bm = row_bmu(m, v, 0); /* both face cells are in the centre's row */
  for (int a = 0; a < N - 1; a++) {
    R gpy = R(0.5) * (cy(m, v, a, 0, 0) + cy(m, v, a, 1, 0));
    R gxa = gx(m, v, a, 0, 0);
    R ng = qsqrt(gpy * gpy + gxa * gxa);
    if (!(ng != R(0.0))) continue;
    R s_c = R(0), s_f = R(0);
    if (m.fn_w == 1) {
      s_c = s_obst<N>(phc, a);
      s_f = s_obst<N>(phf, a);
    }
    R dc = v.dp(a, 0, 0), df = v.dp(a, 1, 0);
    for (int k = 0; k < NC; k++) {
      R cl_f = c_mu_b<NC>(m, bm, muf, L, k), cl_c = c_mu_b<NC>(m, bm, muc, L, k);
      R cs_f = c_mu_b<NC>(m, bm, muf, a, k), cs_c = c_mu_b<NC>(m, bm, muc, a, k);
      R sp = gxa * gxL;
      sp += gpy * gyl;
      if (nxl > R(0.0)) sp = qdiv(sp, (nxl * ng));
      if (m.fn_w == 2)
        jx[k] += qdiv(m.dratio[a][k] * fabs(sp) * m.jat_well * ((cl_f - cs_f) * df + (cl_c - cs_c) * dc) * (gxa), ng);
      else
        jx[k] += qdiv(m.dratio[a][k] * fabs(sp) * R(0.5) * (s_f * (cl_f - cs_f) * df + s_c * (cl_c - cs_c) * dc) * (gxa), ng);
    }
  }
}

/* j_at on the +y face (between centre and (0,1)) of the cell at offset (i,j); yglob = allocated y of that cell */
template <int N, int NC, typename R, class V>
GP_HD void jat_y(const Model<R> &m, const V &v0, int i, int j, int64_t yglob, R *jy) {
  const V v = v0.shift(i, j);
  const int L = N - 1;
  for (int k = 0; k < NC; k++) jy[k] = R(0.0);
  if (yglob == 0 || yglob == m.jat_skip_row) return;
  bool any = false;
  for (int a = 0; a < N - 1; a++) any = any || (v.dp(a, 0, 0) != R(0.0)) || (v.dp(a, 0, 1) != R(0.0));
  if (!any) return;
  R phc[N], phr[N], muc[NC], mur[NC];
  for (int a = 0; a < N; a++) {
    phc[a] = v.ph(a, 0, 0);
    phr[a] = v.ph(a, 0, 1);
  }
  for (int k = 0; k < NC; k++) {
    muc[k] = v.m(k, 0, 0);
    mur[k] = v.m(k, 0, 1);
  }
  R gxl = R(0.5) * (cx(m, v, L, 0, 0) + cx(m, v, L, 0, 1));
  R gyL = gy(m, v, L, 0, 0);
  R nyl = qsqrt(gxl * gxl + gyL * gyL);
  const auto bmc = row_bmu(m, v, 0), bmr = row_bmu(m, v, 1); /* the face's lower and upper rows */
  for (int a = 0; a < N - 1; a++) {
    R gpx = R(0.5) * (cx(m, v, a, 0, 0) + cx(m, v, a, 0, 1));
    R gya = gy(m, v, a, 0, 0);
    R ng = qsqrt(gpx * gpx + gya * gya);
    if (!(ng != R(0.0))) continue;
    R s_c = R(0), s_r = R(0);
    if (m.fn_w == 1) {
      s_c = s_obst<N>(phc, a);
      s_r = s_obst<N>(phr, a);
    }
    R dc = v.dp(a, 0, 0), dr = v.dp(a, 0, 1);
    for (int k = 0; k < NC; k++) {
      R cl_r = c_mu_b<NC>(m, bmr, mur, L, k), cl_c = c_mu_b<NC>(m, bmc, muc, L, k);
      R cs_r = c_mu_b<NC>(m, bmr, mur, a, k), cs_c = c_mu_b<NC>(m, bmc, muc, a, k);
      R sp = gya * gyL;
      sp += gpx * gxl;
      if (nyl > R(0.0)) sp = qdiv(sp, (nyl * ng));
      if (m.fn_w == 2)
        jy[k] += qdiv(m.dratio[a][k] * fabs(sp) * m.jat_well * ((cl_r - cs_r) * dr + (cl_c - cs_c) * dc) * gya, ng);
      else
        jy[k] += qdiv(m.dratio[a][k] * fabs(sp) * R(0.5) * (s_r * (cl_r - cs_r) * dr + s_c * (cl_c - cs_c) * dc) * gya, ng);
    }
  }
}

/* ------------------------------------------------------------------ mu / phi update */
template <int N, int NC, typename R, class V>
GP_HD R mobility(const Model<R> &m, const V &v, int i, int j, int k, int l) {
  R s = R(0.0);
  for (int b = 0; b < N; b++)
    for (int q = 0; q < NC; q++) s += m.DC[b][k][l][q] * v.ph(b, i, j);
  return s;
}

/* sum_l [ Dmid_kl grad mu_l ] on the +x face of offset (i,j), per k (the divflux x-term numerator parts) */
template <int N, int NC, typename R, class V>
GP_HD void mu_face_x(const Model<R> &m, const V &v, int i, int j, R (*Dm)[NC], R *gm) {
  for (int k = 0; k < NC; k++) {
    gm[k] = divx<V>(m, (v.m(k, i + 1, j) - v.m(k, i, j)));
    for (int l = 0; l < NC; l++) Dm[k][l] = (mobility<N, NC>(m, v, i + 1, j, k, l) + mobility<N, NC>(m, v, i, j, k, l)) * R(0.5);
  }
}

template <int N, int NC, typename R, class V>
GP_HD void mu_face_y(const Model<R> &m, const V &v, int i, int j, R (*Dm)[NC], R *gm) {
  for (int k = 0; k < NC; k++) {
    gm[k] = divy<V>(m, (v.m(k, i, j + 1) - v.m(k, i, j)));
    for (int l = 0; l < NC; l++) Dm[k][l] = (mobility<N, NC>(m, v, i, j + 1, k, l) + mobility<N, NC>(m, v, i, j, k, l)) * R(0.5);
  }
}

/* small inverse with the reference's LU (pivot quirk included); pow(x,-1) -> 1/x */
template <int n, typename R>
GP_HD void lu_inverse(R (*a)[n], R (*inv)[n]) {
  R factor[n][n], inv1[n][n];
  int tag[n];
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) factor[i][j] = R(0);
  for (int k = 0; k < n; k++) tag[k] = k;
  for (int k = 0; k < n; k++) {
    R big = fabs(a[k][k]);
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
      R s = a[k][i];
      a[k][i] = a[t1][i];
      a[t1][i] = s;
    }
    for (int i = 0; i < k; i++) {
      R s = factor[k][i];
      factor[k][i] = factor[t1][i];
      factor[t1][i] = s;
    }
    for (int i = k + 1; i < n; i++) {
      R fact = qdiv(-a[i][k], a[k][k]);
      factor[i][k] = -fact;
      for (int j = k; j <= n - 1; j++) a[i][j] = fact * a[k][j] + a[i][j];
    }
  }
  for (int col = 0; col < n; col++) {
    R v1[n], v[n];
    v1[0] = (col == 0) ? R(1.0) : R(0.0);
    for (int i = 1; i < n; i++) {
      R s = R(0);
      for (int j = 0; j < i; j++) s = s - factor[i][j] * v1[j];
      v1[i] = ((i == col) ? R(1.0) : R(0.0)) + s;
    }
    v[n - 1] = v1[n - 1] * qdiv(R(1.0), a[n - 1][n - 1]);
    for (int i = n - 2; i >= 0; i--) {
      R s = R(0);
      for (int j = i + 1; j < n; j++) s = s - a[i][j] * v[j];
      v[i] = (v1[i] + s) * qdiv(R(1.0), a[i][i]);
    }
    for (int j = 0; j < n; j++) inv1[j][col] = v[j];
  }
  for (int j = 0; j < n; j++)
    for (int q = 0; q < n; q++) inv[q][tag[j]] = inv1[q][j];
}

/* New phi, mu and delta-mu of the centre cell. jx0/jxb: +x / -x face currents, jy0/jyl: +y / -y. */
template <int N, int NC, typename R, class V>
GP_HD void mu_update(const Model<R> &m, const V &v, bool smooth, const R *dph, const R *jx0, const R *jxb,
                     const R *jy0, const R *jyl, R *phi_new, R *mu_new, R *dmu) {
  R phi[N], mu[NC], sum[NC], sum_dcbdT[NC], divflux[NC], divjat[NC], deltamu[NC];
  const auto bm = row_bmu(m, v, 0);
  for (int a = 0; a < N; a++) phi[a] = v.ph(a, 0, 0);
  for (int k = 0; k < NC; k++) {
    mu[k] = v.m(k, 0, 0);
    sum[k] = R(0.0);
    sum_dcbdT[k] = R(0.0);
    divjat[k] = R(0.0);
    deltamu[k] = R(0.0);
  }
  if (!smooth) {
    for (int a = 0; a < N; a++) {
      R sdh = R(0.0);
      for (int b = 0; b < N; b++) sdh += dhphi<N>(phi, a, b) * dph[b];
      for (int k = 0; k < NC; k++) {
        sum[k] += c_mu_b<NC>(m, bm, mu, a, k) * sdh;
        if constexpr (V::kTG) sum_dcbdT[k] += m.dcbdT[a][k] * hphi<N>(phi, a);
      }
    }
  }
  R Dx[NC][NC], gmx[NC], Dxb[NC][NC], gmxb[NC], Dy[NC][NC], gmy[NC], Dyl[NC][NC], gmyl[NC];
  mu_face_x<N, NC>(m, v, 0, 0, Dx, gmx);
  mu_face_x<N, NC>(m, v, -1, 0, Dxb, gmxb);
  mu_face_y<N, NC>(m, v, 0, 0, Dy, gmy);
  mu_face_y<N, NC>(m, v, 0, -1, Dyl, gmyl);
  for (int k = 0; k < NC; k++) {
    divflux[k] = R(0.0);
    for (int l = 0; l < NC; l++) {
      divflux[k] += divx<V>(m, (Dx[k][l] * gmx[l] - Dxb[k][l] * gmxb[l]));
      divflux[k] += divy<V>(m, (Dy[k][l] * gmy[l] - Dyl[k][l] * gmyl[l]));
    }
    if (!smooth) {
      divjat[k] = divx<V>(m, (jx0[k] - jxb[k]) * m.jat_fac);
      divjat[k] += divy<V>(m, (jy0[k] - jyl[k]) * m.jat_fac);
    }
  }
  int bulk = -1;
  for (int a = 0; a < N; a++)
    if (bulk < 0 && phi[a] == R(1.0)) bulk = a;
  R munew[NC];
  for (int k = 0; k < NC; k++) munew[k] = mu[k];
  if (bulk >= 0 && kDirectBulkMu<R>) {
    R dc[NC];
    for (int k = 0; k < NC; k++) dc[k] = smooth ? m.dt * (divflux[k]) : m.dt * (divflux[k]) - divjat[k];
    for (int k = 0; k < NC; k++) {
      deltamu[k] = Mu_lin<NC>(m, dc, bulk, k);
      munew[k] = munew[k] + deltamu[k];
    }
  } else if (bulk >= 0) {
    R c_new[NC];
    for (int k = 0; k < NC; k++) {
      R c_old = c_mu_b<NC>(m, bm, mu, bulk, k);
      c_new[k] = smooth ? c_old + m.dt * (divflux[k]) : c_old + m.dt * (divflux[k]) - divjat[k];
    }
    for (int k = 0; k < NC; k++) {
      R nm = Mu_b<NC>(m, bm, c_new, bulk, k);
      deltamu[k] = nm - munew[k];
      munew[k] = nm;
    }
  } else {
    R deltac[NC], dcdmu[NC][NC];
    for (int k = 0; k < NC; k++)
      for (int l = 0; l < NC; l++) dcdmu[k][l] = R(0.0);
    R h[N];
    for (int a = 0; a < N; a++) h[a] = hphi<N>(phi, a);
    for (int k = 0; k < NC; k++) {
      deltac[k] = smooth ? m.dt * (divflux[k]) : m.dt * (divflux[k]) - divjat[k] - sum[k];
      if constexpr (V::kTG)
        if (!smooth) deltac[k] = deltac[k] - sum_dcbdT[k] * m.dt * (-m.tg_G * m.tg_vel);
      for (int l = 0; l < NC; l++)
        for (int a = 0; a < N; a++) dcdmu[k][l] += m.cmu[a][k][l] * h[a];
    }
    const int mode = smooth ? GP_SOLVE_LU : m.solve;
    if (mode == GP_SOLVE_BINARY && NC == 1) {
      R inv = qdiv(R(1.0), dcdmu[0][0]);
      deltamu[0] = deltac[0] * inv;
      munew[0] += deltamu[0];
    } else if (mode == GP_SOLVE_TERNARY && NC == 2) {
      R DET = dcdmu[0][0] * dcdmu[1 % NC][1 % NC] - dcdmu[0][1 % NC] * dcdmu[1 % NC][0];
      R inv[NC][NC];
      inv[0][0] = qdiv(dcdmu[1 % NC][1 % NC], DET);
      inv[1 % NC][1 % NC] = qdiv(dcdmu[0][0], DET);
      inv[0][1 % NC] = qdiv(-dcdmu[0][1 % NC], DET);
      inv[1 % NC][0] = qdiv(-dcdmu[1 % NC][0], DET);
      for (int k = 0; k < NC; k++) {
        deltamu[k] = R(0.0);
        for (int l = 0; l < NC; l++) deltamu[k] += inv[k][l] * deltac[l];
      }
      for (int k = 0; k < NC; k++) munew[k] = deltamu[k] + munew[k];
    } else if (mode == GP_SOLVE_DILUTE) {
      for (int k = 0; k < NC; k++) {
        R inv = qdiv(R(1.0), dcdmu[k][k]);
        deltamu[k] = deltac[k] * inv;
        munew[k] += deltamu[k];
      }
    } else {
      R inv[NC][NC];
      lu_inverse<NC>(dcdmu, inv);
      for (int i = 0; i < NC; i++) {
        R s = R(0);
        for (int j = 0; j < NC; j++) s = s + inv[i][j] * deltac[j];
        deltamu[i] = s;
      }
      for (int k = 0; k < NC; k++) munew[k] = deltamu[k] + munew[k];
    }
  }
  for (int b = 0; b < N; b++) phi_new[b] = phi[b] + dph[b];
  for (int k = 0; k < NC; k++) {
    mu_new[k] = munew[k];
    dmu[k] = deltamu[k];
  }
}

/* ------------------------------------------------------------------ compensated FP32 storage (GP_FP32C)
 * mu: a value v is stored as hi = fl32(v) (round to nearest) and an 8-bit signed extension q counting units of
 * ulp32(hi) / 128: v ~= hi + q * 2^(E - 157), E the biased exponent of hi (ulp32(hi) = 2^(E - 150) for a normal
 * hi). |v - hi| <= ulp32(hi) / 2 gives |q| <= 64. Zero, subnormal and non-finite hi carry q = 0. The pair holds
 * about 31 significant bits, relative to v.
 * phi: the model forms 1 - phi (h, dh/dphi, the anti-trapping s(phi)), which near phi = 1 needs absolute precision
 * far below that of 31 bits relative to 1, so phi is stored as an unevaluated sum of two FP32 values,
 * v = hi + lo with hi = fl32(v) and lo = fl32(v - hi) (about 48 bits; 1 - phi keeps 24 bits of its own).
 * Every conversion is an exact operation or one correctly rounded one, in FP64, so host and device agree bit for
 * bit. */
GP_HD uint32_t f32_bits(float f) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __float_as_uint(f);
#else
  uint32_t u;
  memcpy(&u, &f, sizeof u);
  return u;
#endif
}
/* 2^k for -1022 <= k <= 1023, exactly */
GP_HD double pow2d(int k) {
  const uint64_t b = (uint64_t)(k + 1023) << 52;
#if defined(__HIP_DEVICE_COMPILE__)
  return __longlong_as_double((long long)b);
#else
  double d;
  memcpy(&d, &b, sizeof d);
  return d;
#endif
}
GP_HD double c32_decode(float hi, int8_t q) {
  if (q == 0) return (double)hi;
  const int E = (int)((f32_bits(hi) >> 23) & 0xffu);
  return (double)hi + (double)q * pow2d(E - 157);
}
GP_HD void c32_encode(double v, float &hi, int8_t &q) {
  hi = (float)v;
  const int E = (int)((f32_bits(hi) >> 23) & 0xffu);
  q = 0;
  if (E == 0 || E == 255) return;
  double s = rint((v - (double)hi) * pow2d(157 - E));
  s = s > 127.0 ? 127.0 : (s < -127.0 ? -127.0 : s);
  q = (int8_t)s;
}
GP_HD double df_decode(float hi, float lo) { return (double)hi + (double)lo; }
GP_HD void df_encode(double v, float &hi, float &lo) {
  hi = (float)v;
  lo = (float)(v - (double)hi);
}

/* Host-side construction of the device parameter block from gp_params (same expressions as the reference). */
template <typename R>
inline int make_model(const gp_params *p, Model<R> *m) {
  const int N = p->nphases, NC = p->ncomp - 1, L = N - 1;
  if (N < 2 || N > P || NC < 1 || NC > NCM) return -1;
  *m = Model<R>();
  m->dx = (R)p->dx;
  m->dy = (R)p->dy;
  m->dt = (R)p->dt;
  m->eps = (R)p->epsilon;
  m->tau = (R)p->tau;
  m->inv_eps = (R)(1.0 / p->epsilon);
  m->two_dy = (R)(2.0 * p->dy);
  {
    int ex, ey;
    const bool px = frexp(p->dx, &ex) == 0.5, py = frexp(p->dy, &ey) == 0.5;
    m->pow2 = (px && py) ? 1 : 0;
    m->inv_dx = (R)(1.0 / p->dx);
    m->inv_dy = (R)(1.0 / p->dy);
    m->inv_2dy = (R)(1.0 / (2.0 * p->dy));
  }
  m->w_obst = (R)(16.0 / (kPi * kPi));
  m->jat_well = (R)(0.5 / sqrt(2.0));
  m->jat_fac = p->fn_w == 1 ? (R)(-0.25 * kPi * p->epsilon) : (R)(-1.0 * p->epsilon);
  m->N = N;
  m->NC = NC;
  m->fn_w = p->fn_w;
  m->fn_aniso = p->fn_aniso;
  m->solve = p->solve;
  m->jat_skip_row = p->jat_skip_row;
  for (int a = 0; a < P; a++)
    for (int b = 0; b < P; b++) {
      m->tau_ab[a][b] = (R)p->tau_ab[a * P + b];
      m->gamma[a][b] = (R)p->gamma[a * P + b];
      m->dab[a][b] = (R)p->dab[a * P + b];
      for (int c = 0; c < P; c++) m->gamma_abc[a][b][c] = (R)p->gamma_abc[(a * P + b) * P + c];
      for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) {
          m->rot[a][b][i][j] = (R)p->rot[((a * P + b) * 2 + i) * 2 + j];
          m->rot_inv[a][b][i][j] = (R)p->rot_inv[((a * P + b) * 2 + i) * 2 + j];
        }
    }
  for (int a = 0; a < P; a++) {
    for (int i = 0; i < NCM; i++) {
      for (int j = 0; j < NCM; j++) {
        m->D[a][i][j] = (R)p->diffusivity[(a * NCM + i) * NCM + j];
        m->A[a][i][j] = (R)p->A[(a * NCM + i) * NCM + j];
        m->cmu[a][i][j] = (R)p->cmu[(a * NCM + i) * NCM + j];
      }
      m->B[a][i] = (R)p->B[a * NCM + i];
      m->Bmu[a][i] = (R)(p->Beq[a * NCM + i] + p->dBbdT[a * NCM + i] * (p->T - p->Teq));
      if (a < N && i < NC && a < L)
        m->dratio[a][i] = (R)(1.0 - p->diffusivity[(a * NCM + i) * NCM + i] / p->diffusivity[(L * NCM + i) * NCM + i]);
      for (int l = 0; l < NCM; l++)
        for (int q = 0; q < NCM; q++)
          m->DC[a][i][l][q] = (R)(p->diffusivity[(a * NCM + i) * NCM + q] * p->cmu[(a * NCM + q) * NCM + l]);
    }
    m->C[a] = (R)p->C[a];
  }
  /* mobility of a cell that is purely phase b (phi_b = 1, every other phase 0), formed with the same
   * operations, in the same order and precision, as the per-cell sum in mobility(): bit-identical to it */
  for (int b = 0; b < P; b++)
    for (int k = 0; k < NCM; k++)
      for (int l = 0; l < NCM; l++) {
        R s = R(0.0);
        if (b < N && k < NC && l < NC)
          for (int a = 0; a < N; a++)
            for (int q = 0; q < NC; q++) s += m->DC[a][k][l][q] * (a == b ? R(1.0) : R(0.0));
        m->Mpure[b][k][l] = s;
      }
  for (int a = 0; a < P; a++)
    for (int k = 0; k < NCM; k++) m->dcbdT[a][k] = (R)p->dcbdT_phase[a * NCM + k];
  m->tg_G = R(0.0); /* temperature_gradientY.GRADIENT, which the serial driver never assigns */
  m->tg_vel = (R)p->tg_velocity;
  m->lap_x = (R)(1.0 / (p->dx * p->dx));
  m->lap_y = (R)(1.0 / (p->dy * p->dy));
  for (int b = 0; b < N; b++)
    for (int k = 0; k < NC; k++)
      for (int l = 0; l < NC; l++) {
        double s = 0.0;
        for (int j = 0; j < NC; j++) {
          double mjl = 0.0;
          for (int q = 0; q < NC; q++)
            mjl += p->diffusivity[(b * NCM + j) * NCM + q] * p->cmu[(b * NCM + q) * NCM + l];
          const double hkj = k == j ? 2.0 * p->A[(b * NCM + k) * NCM + k] : p->A[(b * NCM + k) * NCM + j];
          s += hkj * p->dt * mjl;
        }
        m->Kbulk[b][k][l] = (R)s;
      }
  return 0;
}

}  // namespace gpk
