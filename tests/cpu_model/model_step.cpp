// Host build of the kernels' per-cell arithmetic (microsim_amd/csrc/gp_model.h): steps a halo-inclusive grid
// cell by cell in the same decomposition as the device kernels (delta-phi on [2, mesh+3]^2, currents and update
// on the interior, halo copies from the maps). Test infrastructure: lets the CPU suite compare the kernel source
// with the oracle before any GPU run.
#include <cstring>
#include <vector>

#include "gp_model.h"

using namespace gpk;

template <int N, int NC, bool P2>
static int run(const gp_params *p, double *phi, double *mu, double *dphi, int64_t nsteps, int smooth) {
  Model<double> m;
  if (make_model(p, &m)) return -1;
  if (P2 && !m.pow2) return -3;
  const int64_t rx = p->mesh_x + 6, ry = p->mesh_y + 6, n = rx * ry;
  std::vector<double> pn(phi, phi + N * n), mn(mu, mu + NC * n);
  std::vector<double> dp(N * n, 0.0);
  for (int64_t t = 0; t < nsteps; t++) {
    View<double, P2> v{phi, mu, dp.data(), n, ry, 0};
    for (int64_t x = 2; x <= p->mesh_x + 3; x++)
      for (int64_t y = 2; y <= p->mesh_y + 3; y++) {
        double d[N];
        phi_delta<N, NC>(m, v.shift((int)x, (int)y), smooth != 0, d);
        for (int a = 0; a < N; a++) dp[a * n + x * ry + y] = d[a];
      }
    for (int64_t x = 3; x <= p->mesh_x + 2; x++)
      for (int64_t y = 3; y <= p->mesh_y + 2; y++) {
        View<double, P2> c = v.shift((int)x, (int)y);
        double jx0[NC], jxb[NC], jy0[NC], jyl[NC], d[N], pnew[N], mnew[NC], dm[NC];
        if (!smooth) {
          jat_x<N, NC>(m, c, 0, 0, jx0);
          jat_x<N, NC>(m, c, -1, 0, jxb);
          jat_y<N, NC>(m, c, 0, 0, y, jy0);
          jat_y<N, NC>(m, c, 0, -1, y - 1, jyl);
        }
        for (int a = 0; a < N; a++) d[a] = dp[a * n + x * ry + y];
        mu_update<N, NC>(m, c, smooth != 0, d, jx0, jxb, jy0, jyl, pnew, mnew, dm);
        for (int a = 0; a < N; a++) pn[a * n + x * ry + y] = pnew[a];
        for (int k = 0; k < NC; k++) mn[k * n + x * ry + y] = mnew[k];
      }
    std::memcpy(phi, pn.data(), sizeof(double) * N * n);
    std::memcpy(mu, mn.data(), sizeof(double) * NC * n);
    const int64_t hx[6] = {0, 1, 2, rx - 3, rx - 2, rx - 1}, hy[6] = {0, 1, 2, ry - 3, ry - 2, ry - 1};
    for (int f = 0; f < 2; f++)
      for (int c = 0; c < (f ? NC : N); c++) {
        double *vv = (f ? mu : phi) + c * n;
        for (int s = 0; s < 6; s++)
          for (int64_t y = 0; y < ry; y++) vv[hx[s] * ry + y] = vv[(int64_t)p->halo_src_x[f][s] * ry + y];
        for (int s = 0; s < 6; s++)
          for (int64_t x = 0; x < rx; x++) vv[x * ry + hy[s]] = vv[x * ry + p->halo_src_y[f][s]];
      }
  }
  if (dphi) std::memcpy(dphi, dp.data(), sizeof(double) * N * n);
  return 0;
}

/* pow2: 1 forces the power-of-two spacing path (error -3 when dx, dy are not powers of two), 0 the division path */
extern "C" int model_step(const gp_params *p, double *phi, double *mu, double *dphi, int64_t nsteps, int smooth,
                          int pow2) {
  const int N = p->nphases, NC = p->ncomp - 1;
#define CASE(a, b) \
  if (N == a && NC == b) return pow2 ? run<a, b, true>(p, phi, mu, dphi, nsteps, smooth) : run<a, b, false>(p, phi, mu, dphi, nsteps, smooth);
  CASE(2, 1) CASE(3, 2) CASE(3, 1) CASE(2, 2) CASE(4, 2) CASE(4, 1) CASE(3, 3)
#undef CASE
  return -2;
}
