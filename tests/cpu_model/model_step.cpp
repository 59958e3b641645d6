// Host build of the kernels' per-cell arithmetic (microsim_amd/csrc/gp_model.h): steps a halo-inclusive grid
// cell by cell in the same decomposition as the device kernels (delta-phi on [2, mesh+3]^2, currents and update
// on the interior, halo copies from the maps). Test infrastructure: lets the CPU suite compare the kernel source
// with the oracle before any GPU run.
//
// model_jvp runs the same steps in forward-mode dual numbers (microsim_amd/csrc/gp_dual.h, one direction): the
// directional derivative of the trajectory along a tangent of the initial state and of theta
// (microsim_amd/csrc/gp_theta.h), exact to rounding and taking the branches the values take. The tests set it
// against finite differences (where the trajectory is smooth) and against the device adjoint (a dot-product
// test, which holds whatever the branches).
#include <cstring>
#include <vector>

#include "gp_theta.h"

using namespace gpk;

template <typename R>
static void apply_bc(const gp_params *p, R *phi, R *mu, int N, int NC) {
  const int64_t rx = p->mesh_x + 6, ry = p->mesh_y + 6, n = rx * ry;
  const int64_t hx[6] = {0, 1, 2, rx - 3, rx - 2, rx - 1}, hy[6] = {0, 1, 2, ry - 3, ry - 2, ry - 1};
  for (int f = 0; f < 2; f++)
    for (int c = 0; c < (f ? NC : N); c++) {
      R *vv = (f ? mu : phi) + c * n;
      for (int s = 0; s < 6; s++)
        for (int64_t y = 0; y < ry; y++) vv[hx[s] * ry + y] = vv[(int64_t)p->halo_src_x[f][s] * ry + y];
      for (int s = 0; s < 6; s++)
        for (int64_t x = 0; x < rx; x++) vv[x * ry + hy[s]] = vv[x * ry + p->halo_src_y[f][s]];
    }
}

template <int N, int NC, bool P2, typename R>
static void advance(const Model<R> &m, const gp_params *p, R *phi, R *mu, R *dphi, int64_t nsteps, int smooth) {
  const int64_t rx = p->mesh_x + 6, ry = p->mesh_y + 6, n = rx * ry;
  std::vector<R> pn(phi, phi + N * n), mn(mu, mu + NC * n);
  std::vector<R> dp(N * n, R(0.0));
  for (int64_t t = 0; t < nsteps; t++) {
    View<R, P2> v{phi, mu, dp.data(), n, ry, 0};
    for (int64_t x = 2; x <= p->mesh_x + 3; x++)
      for (int64_t y = 2; y <= p->mesh_y + 3; y++) {
        R d[N];
        phi_delta<N, NC>(m, v.shift((int)x, (int)y), smooth != 0, d);
        for (int a = 0; a < N; a++) dp[a * n + x * ry + y] = d[a];
      }
    for (int64_t x = 3; x <= p->mesh_x + 2; x++)
      for (int64_t y = 3; y <= p->mesh_y + 2; y++) {
        View<R, P2> c = v.shift((int)x, (int)y);
        R jx0[NC], jxb[NC], jy0[NC], jyl[NC], d[N], pnew[N], mnew[NC], dm[NC];
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
    std::copy(pn.begin(), pn.end(), phi);
    std::copy(mn.begin(), mn.end(), mu);
    apply_bc(p, phi, mu, N, NC);
  }
  if (dphi) std::copy(dp.begin(), dp.end(), dphi);
}

template <int N, int NC, bool P2>
static int run(const gp_params *p, double *phi, double *mu, double *dphi, int64_t nsteps, int smooth) {
  Model<double> m;
  if (make_model(p, &m)) return -1;
  if (P2 && !m.pow2) return -3;
  advance<N, NC, P2>(m, p, phi, mu, dphi, nsteps, smooth);
  return 0;
}

template <int N, int NC, bool P2>
static int run_jvp(const gp_params *p, double *phi, double *mu, const double *vphi, const double *vmu,
                   const double *vtheta, double *tphi, double *tmu, int64_t nsteps) {
  using D = Dual<1>;
  Model<D> m;
  if (make_model(p, &m)) return -1;
  if (P2 && !m.pow2) return -3;
  make_dual_model_w<1>(p, [=](int t, int) { return vtheta[t]; }, &m);
  const int64_t n = (p->mesh_x + 6) * (p->mesh_y + 6);
  std::vector<D> ph(N * n), mu_(NC * n);
  for (int64_t i = 0; i < N * n; i++) {
    ph[i].v = phi[i];
    ph[i].d[0] = vphi[i];
  }
  for (int64_t i = 0; i < NC * n; i++) {
    mu_[i].v = mu[i];
    mu_[i].d[0] = vmu[i];
  }
  apply_bc(p, ph.data(), mu_.data(), N, NC); /* the state is a function of the interior, as after an upload */
  advance<N, NC, P2>(m, p, ph.data(), mu_.data(), (D *)nullptr, nsteps, 0);
  for (int64_t i = 0; i < N * n; i++) {
    phi[i] = ph[i].v;
    tphi[i] = ph[i].d[0];
  }
  for (int64_t i = 0; i < NC * n; i++) {
    mu[i] = mu_[i].v;
    tmu[i] = mu_[i].d[0];
  }
  return 0;
}

#define GP_CASES(CASE) CASE(2, 1) CASE(3, 2) CASE(3, 1) CASE(2, 2) CASE(4, 2) CASE(4, 1) CASE(3, 3)

/* pow2: 1 forces the power-of-two spacing path (error -3 when dx, dy are not powers of two), 0 the division path */
extern "C" int model_step(const gp_params *p, double *phi, double *mu, double *dphi, int64_t nsteps, int smooth,
                          int pow2) {
  const int N = p->nphases, NC = p->ncomp - 1;
#define CASE(a, b) \
  if (N == a && NC == b) return pow2 ? run<a, b, true>(p, phi, mu, dphi, nsteps, smooth) : run<a, b, false>(p, phi, mu, dphi, nsteps, smooth);
  GP_CASES(CASE)
#undef CASE
  return -2;
}

/* nsteps steps from (phi, mu) (updated in place) with the tangent (vphi, vmu, vtheta); the tangent of the final
 * state goes to (tphi, tmu). Halo entries of vphi / vmu are replaced from the interior, as the state's are. */
extern "C" int model_jvp(const gp_params *p, double *phi, double *mu, const double *vphi, const double *vmu,
                         const double *vtheta, double *tphi, double *tmu, int64_t nsteps, int pow2) {
  const int N = p->nphases, NC = p->ncomp - 1;
#define CASE(a, b)                                                                        \
  if (N == a && NC == b)                                                                  \
    return pow2 ? run_jvp<a, b, true>(p, phi, mu, vphi, vmu, vtheta, tphi, tmu, nsteps) \
                : run_jvp<a, b, false>(p, phi, mu, vphi, vmu, vtheta, tphi, tmu, nsteps);
  GP_CASES(CASE)
#undef CASE
  return -2;
}
