/*
 * gp_theta.h -- the adjoint's parameters theta and the dual-number model that carries their derivatives.
 *
 * theta (gp_hip.h): per unordered phase pair (a < b, pair order) dab, tau_ab and Gamma, each entering the model
 * symmetrically; then the liquid's diagonal diffusivities D[L][k][k]; then tau. Host- and device-compatible (the
 * device adjoint and the host test builds share it); not a public header.
 */
#pragma once

#include "gp_dual.h"
#include "gp_reverse.h"

namespace gpk {

/* Model<Dual<K>> whose theta entries carry the tangents w(t, k) (derivative of theta t along direction k), with
 * the coefficients derived from the diffusivities (D cmu, 1 - D_a/D_L, the pure-phase mobilities) re-formed in
 * dual arithmetic by the same expressions make_model uses. */
template <int K, class W>
inline void make_dual_model_w(const gp_params *p, W w, Model<Dual<K>> *m) {
  using D = Dual<K>;
  make_model(p, m);
  const int N = p->nphases, NC = p->ncomp - 1, L = N - 1;
  auto seed = [&](D &x, int t) {
    for (int k = 0; k < K; k++) x.d[k] = w(t, k);
  };
  int pi = 0;
  for (int a = 0; a < N; a++)
    for (int b = a + 1; b < N; b++, pi++) {
      seed(m->dab[a][b], 3 * pi);
      seed(m->dab[b][a], 3 * pi);
      seed(m->tau_ab[a][b], 3 * pi + 1);
      seed(m->tau_ab[b][a], 3 * pi + 1);
      seed(m->gamma[a][b], 3 * pi + 2);
      seed(m->gamma[b][a], 3 * pi + 2);
    }
  const int t0 = 3 * pi;
  for (int k = 0; k < NC; k++) seed(m->D[L][k][k], t0 + k);
  seed(m->tau, t0 + NC);
  for (int a = 0; a < N; a++)
    for (int i = 0; i < NC; i++) {
      if (a < L) m->dratio[a][i] = D(1.0) - m->D[a][i][i] / m->D[L][i][i];
      for (int l = 0; l < NC; l++)
        for (int q = 0; q < NC; q++) m->DC[a][i][l][q] = m->D[a][i][q] * m->cmu[a][q][l];
    }
  for (int b = 0; b < N; b++)
    for (int k = 0; k < NC; k++)
      for (int l = 0; l < NC; l++) {
        D s = D(0.0);
        for (int a = 0; a < N; a++)
          for (int q = 0; q < NC; q++) s += m->DC[a][k][l][q] * (a == b ? D(1.0) : D(0.0));
        m->Mpure[b][k][l] = s;
      }
}

}  // namespace gpk
