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
//
// model_vjp runs the adjoint of the same steps with the reverse-mode per-cell functions of
// microsim_amd/csrc/gp_reverse.h (the ones the device adjoint runs), on every cell and without the device's bulk
// shortcuts: <J^T w, v> from it must equal <w, J v> from model_jvp.
#include <cmath>
#include <cstring>
#include <vector>

#include "gp_reverse.h"
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

/* stencil of the cell at index c (halo-inclusive grid, y fastest): phi 3x3, mu and delta-phi on the 5-point one */
template <int N, int NC>
static void gather(const double *phi, const double *mu, const double *dphi, int64_t n, int64_t ry, int64_t c,
                   RegStore<N, NC, double> &rs) {
  const int oi[5] = {0, 1, -1, 0, 0}, oj[5] = {0, 0, 0, 1, -1};
  for (int a = 0; a < N; a++)
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) rs.p[a][i][j] = phi[a * n + c + (i - 1) * ry + (j - 1)];
  for (int o = 0; o < 5; o++) {
    const int64_t co = c + oi[o] * ry + oj[o];
    for (int k = 0; k < NC; k++) rs.u[k][o] = mu[k * n + co];
    for (int a = 0; a < N; a++) rs.d[a][o] = dphi ? dphi[a * n + co] : 0.0;
  }
}

template <int N, int NC>
static void scatter(const RevAcc<N, NC> &acc, int64_t n, int64_t ry, int64_t c, double *gphi, double *gmu,
                    double *gdphi, double *gth) {
  const int oi[5] = {0, 1, -1, 0, 0}, oj[5] = {0, 0, 0, 1, -1};
  for (int a = 0; a < N; a++)
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) gphi[a * n + c + (i - 1) * ry + (j - 1)] += acc.p[a][i][j];
  for (int o = 0; o < 5; o++) {
    const int64_t co = c + oi[o] * ry + oj[o];
    for (int k = 0; k < NC; k++) gmu[k * n + co] += acc.u[k][o];
    if (gdphi)
      for (int a = 0; a < N; a++) gdphi[a * n + co] += acc.d[a][o];
  }
  for (int t = 0; t < kThetaN<N, NC>; t++) gth[t] += acc.th[t];
}

/* the boundary copies transposed: the y maps (applied last) first, then the x maps; halo adjoints are cleared */
static void bc_transpose(const gp_params *p, double *g, int ncomp, int f) {
  const int64_t rx = p->mesh_x + 6, ry = p->mesh_y + 6, n = rx * ry;
  const int64_t hx[6] = {0, 1, 2, rx - 3, rx - 2, rx - 1}, hy[6] = {0, 1, 2, ry - 3, ry - 2, ry - 1};
  for (int c = 0; c < ncomp; c++) {
    double *v = g + c * n;
    for (int s = 5; s >= 0; s--)
      for (int64_t x = rx - 1; x >= 0; x--) {
        v[x * ry + p->halo_src_y[f][s]] += v[x * ry + hy[s]];
        v[x * ry + hy[s]] = 0.0;
      }
    for (int s = 5; s >= 0; s--)
      for (int64_t y = ry - 1; y >= 0; y--) {
        v[(int64_t)p->halo_src_x[f][s] * ry + y] += v[hx[s] * ry + y];
        v[hx[s] * ry + y] = 0.0;
      }
  }
}

template <int N, int NC, bool P2>
static int run_vjp(const gp_params *p, const double *phi0, const double *mu0, const double *wphi, const double *wmu,
                   double *gphi0, double *gmu0, double *gtheta, int64_t nsteps) {
  Model<double> m;
  if (make_model(p, &m)) return -1;
  if (P2 && !m.pow2) return -3;
  const int64_t rx = p->mesh_x + 6, ry = p->mesh_y + 6, n = rx * ry;
  /* the trajectory: states s_0 .. s_{T-1} and the delta-phi of each step */
  std::vector<std::vector<double>> sp(nsteps), sm(nsteps), sd(nsteps);
  std::vector<double> ph(phi0, phi0 + N * n), mu(mu0, mu0 + NC * n);
  apply_bc(p, ph.data(), mu.data(), N, NC);
  for (int64_t t = 0; t < nsteps; t++) {
    sp[t] = ph;
    sm[t] = mu;
    sd[t].assign(N * n, 0.0);
    advance<N, NC, P2>(m, p, ph.data(), mu.data(), sd[t].data(), 1, 0);
  }
  std::vector<double> gp(wphi, wphi + N * n), gm(wmu, wmu + NC * n), gd(N * n);
  std::vector<double> gpn(N * n), gmn(NC * n);
  double th[kThetaN<N, NC>] = {0};
  for (int64_t t = nsteps - 1; t >= 0; t--) {
    bc_transpose(p, gp.data(), N, 0);
    bc_transpose(p, gm.data(), NC, 1);
    std::fill(gpn.begin(), gpn.end(), 0.0);
    std::fill(gmn.begin(), gmn.end(), 0.0);
    std::fill(gd.begin(), gd.end(), 0.0);
    const double *P = sp[t].data(), *M = sm[t].data(), *D = sd[t].data();
    /* stage Q (interior), then stage P ([2, mesh+3]^2) */
    for (int64_t x = 3; x <= p->mesh_x + 2; x++)
      for (int64_t y = 3; y <= p->mesh_y + 2; y++) {
        const int64_t c = x * ry + y;
        RegStore<N, NC, double> rs;
        gather<N, NC>(P, M, D, n, ry, c, rs);
        RevAcc<N, NC> acc;
        acc.clear();
        double go[N], gu[NC];
        for (int a = 0; a < N; a++) go[a] = gp[a * n + c];
        for (int k = 0; k < NC; k++) gu[k] = gm[k * n + c];
        cell_update_vjp<N, NC>(m, RegView<N, NC, double, P2>{&rs, 0, 0}, y, go, gu, acc);
        scatter<N, NC>(acc, n, ry, c, gpn.data(), gmn.data(), gd.data(), th);
      }
    for (int64_t x = 2; x <= p->mesh_x + 3; x++)
      for (int64_t y = 2; y <= p->mesh_y + 3; y++) {
        const int64_t c = x * ry + y;
        double go[N];
        bool nz = false;
        for (int a = 0; a < N; a++) {
          go[a] = gd[a * n + c];
          nz = nz || go[a] != 0.0;
        }
        if (!nz) continue;
        RegStore<N, NC, double> rs;
        gather<N, NC>(P, M, nullptr, n, ry, c, rs);
        RevAcc<N, NC> acc;
        acc.clear();
        phi_delta_vjp<N, NC>(m, RegView<N, NC, double, P2>{&rs, 0, 0}, go, acc);
        scatter<N, NC>(acc, n, ry, c, gpn.data(), gmn.data(), nullptr, th);
      }
    gp.swap(gpn);
    gm.swap(gmn);
  }
  bc_transpose(p, gp.data(), N, 0); /* the initial state's halos are copies of its interior */
  bc_transpose(p, gm.data(), NC, 1);
  std::copy(gp.begin(), gp.end(), gphi0);
  std::copy(gm.begin(), gm.end(), gmu0);
  for (int t = 0; t < kThetaN<N, NC>; t++) gtheta[t] = th[t];
  return 0;
}

/* The compensated FP32 step (GP_FP32C) in the device's decomposition: phi as pairs of FP32 values (df_encode /
 * df_decode of gp_model.h), mu as FP32 parts plus 8-bit extensions (c32_encode / c32_decode); phase codes from the
 * decoded phi; delta-phi in FP64 from the
 * decoded state at the candidates of [2, mesh+3]^2 (5-point codes not uniform) and 0 elsewhere; bulk cells (radius-2
 * code diamond purely phase b) take the FP32 closed form Kbulk[b] . lap(mu) from the FP32 parts, added to the
 * decoded mu and re-encoded, with phi = e_b; every other interior cell takes jat_x / jat_y / mu_update in FP64 from
 * the decoded state, encoded on store; halos are copies of their source cells (parts and extensions alike).
 * (phi, mu) in: the FP64 state (boundary maps applied first, as gp_upload does); out: the decoded final state. */
template <typename T>
static void bc_planes(const gp_params *p, T *v, int ncomp, int f) {
  const int64_t rx = p->mesh_x + 6, ry = p->mesh_y + 6, n = rx * ry;
  const int64_t hx[6] = {0, 1, 2, rx - 3, rx - 2, rx - 1}, hy[6] = {0, 1, 2, ry - 3, ry - 2, ry - 1};
  for (int c = 0; c < ncomp; c++) {
    T *w = v + c * n;
    for (int s = 0; s < 6; s++)
      for (int64_t y = 0; y < ry; y++) w[hx[s] * ry + y] = w[(int64_t)p->halo_src_x[f][s] * ry + y];
    for (int s = 0; s < 6; s++)
      for (int64_t x = 0; x < rx; x++) w[x * ry + hy[s]] = w[x * ry + p->halo_src_y[f][s]];
  }
}

template <int N, int NC, bool P2>
static int run_c32(const gp_params *p, double *phi, double *mu, int64_t nsteps) {
  Model<double> md;
  Model<float> mf;
  if (make_model(p, &md) || make_model(p, &mf)) return -1;
  if (P2 && !md.pow2) return -3;
  const int64_t mx = p->mesh_x, my = p->mesh_y, rx = mx + 6, ry = my + 6, n = rx * ry;
  apply_bc(p, phi, mu, N, NC);
  std::vector<float> ph(N * n), pq(N * n), mh(NC * n);
  std::vector<int8_t> mq(NC * n);
  for (int64_t i = 0; i < N * n; i++) df_encode(phi[i], ph[i], pq[i]);
  for (int64_t i = 0; i < NC * n; i++) c32_encode(mu[i], mh[i], mq[i]);
  std::vector<double> P(N * n), M(NC * n), D(N * n);
  std::vector<int8_t> code(n);
  for (int64_t t = 0; t < nsteps; t++) {
    for (int64_t i = 0; i < N * n; i++) P[i] = df_decode(ph[i], pq[i]);
    for (int64_t i = 0; i < NC * n; i++) M[i] = c32_decode(mh[i], mq[i]);
    for (int64_t i = 0; i < n; i++) {
      double v[N];
      for (int a = 0; a < N; a++) v[a] = P[a * n + i];
      int cd = -1, nz = 0;
      for (int a = 0; a < N; a++) {
        if (v[a] == 1.0) cd = a;
        nz += v[a] != 0.0;
      }
      code[i] = (int8_t)(nz == 1 ? cd : -1);
    }
    std::fill(D.begin(), D.end(), 0.0);
    View<double, P2, false, true> vw{P.data(), M.data(), D.data(), n, ry, 0};
    for (int64_t x = 2; x <= mx + 3; x++)
      for (int64_t y = 2; y <= my + 3; y++) {
        const int64_t c = x * ry + y;
        const int8_t k0 = code[c];
        const bool cand = k0 < 0 || code[c - ry] != k0 || code[c + ry] != k0 || code[c - 1] != k0 || code[c + 1] != k0;
        if (!cand) continue;
        double d[N];
        phi_delta<N, NC>(md, vw.shift((int)x, (int)y), false, d);
        for (int a = 0; a < N; a++) D[a * n + c] = d[a];
      }
    std::vector<float> nph(ph), npq(pq), nmh(mh);
    std::vector<int8_t> nmq(mq);
    for (int64_t x = 3; x <= mx + 2; x++)
      for (int64_t y = 3; y <= my + 2; y++) {
        const int64_t c = x * ry + y;
        int b = code[c];
        if (b >= 0) {
          const int o[12][2] = {{1, 0}, {-1, 0}, {2, 0}, {-2, 0}, {0, 1}, {0, -1},
                                {0, 2}, {0, -2}, {1, 1}, {1, -1}, {-1, 1}, {-1, -1}};
          for (auto &e : o)
            if (code[c + e[0] * ry + e[1]] != b) b = -1;
        }
        if (b >= 0) {
          float lap[NC];
          for (int l = 0; l < NC; l++) {
            const float *w = mh.data() + l * n;
            const float gx = (w[c + ry] - w[c]) - (w[c] - w[c - ry]);
            const float gy = (w[c + 1] - w[c]) - (w[c] - w[c - 1]);
            lap[l] = std::fmaf(gx, mf.lap_x, gy * mf.lap_y);
          }
          for (int k = 0; k < NC; k++) {
            float s = 0.f;
            for (int l = 0; l < NC; l++) s = l == 0 ? mf.Kbulk[b][k][l] * lap[0] : std::fmaf(mf.Kbulk[b][k][l], lap[l], s);
            c32_encode(c32_decode(mh[k * n + c], mq[k * n + c]) + (double)s, nmh[k * n + c], nmq[k * n + c]);
          }
          for (int a = 0; a < N; a++) {
            nph[a * n + c] = a == b ? 1.0f : 0.0f;
            npq[a * n + c] = 0.0f;
          }
          continue;
        }
        auto cv = vw.shift((int)x, (int)y);
        double jx0[NC], jxb[NC], jy0[NC], jyl[NC], d[N], pn[N], mn[NC], dm[NC];
        jat_x<N, NC>(md, cv, 0, 0, jx0);
        jat_x<N, NC>(md, cv, -1, 0, jxb);
        jat_y<N, NC>(md, cv, 0, 0, y, jy0);
        jat_y<N, NC>(md, cv, 0, -1, y - 1, jyl);
        for (int a = 0; a < N; a++) d[a] = D[a * n + c];
        mu_update<N, NC>(md, cv, false, d, jx0, jxb, jy0, jyl, pn, mn, dm);
        for (int a = 0; a < N; a++) df_encode(pn[a], nph[a * n + c], npq[a * n + c]);
        for (int k = 0; k < NC; k++) c32_encode(mn[k], nmh[k * n + c], nmq[k * n + c]);
      }
    bc_planes(p, nph.data(), N, 0);
    bc_planes(p, npq.data(), N, 0);
    bc_planes(p, nmh.data(), NC, 1);
    bc_planes(p, nmq.data(), NC, 1);
    ph.swap(nph);
    pq.swap(npq);
    mh.swap(nmh);
    mq.swap(nmq);
  }
  for (int64_t i = 0; i < N * n; i++) phi[i] = df_decode(ph[i], pq[i]);
  for (int64_t i = 0; i < NC * n; i++) mu[i] = c32_decode(mh[i], mq[i]);
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

/* Gradient of <wphi, phi_T> + <wmu, mu_T> after nsteps steps from (phi, mu) with respect to the initial state
 * (halo entries 0: they are copies of the interior) and theta, by the reverse-mode per-cell functions. */
extern "C" int model_vjp(const gp_params *p, const double *phi, const double *mu, const double *wphi,
                         const double *wmu, double *gphi, double *gmu, double *gtheta, int64_t nsteps, int pow2) {
  const int N = p->nphases, NC = p->ncomp - 1;
#define CASE(a, b)                                                                        \
  if (N == a && NC == b)                                                                  \
    return pow2 ? run_vjp<a, b, true>(p, phi, mu, wphi, wmu, gphi, gmu, gtheta, nsteps) \
                : run_vjp<a, b, false>(p, phi, mu, wphi, wmu, gphi, gmu, gtheta, nsteps);
  GP_CASES(CASE)
#undef CASE
  return -2;
}

/* Per-cell check of the reverse-mode functions against forward-mode duals: for the cell at index c of a state
 * (phi, mu and delta-phi halo-inclusive), stage 0 (P: phi_delta) or 1 (Q: currents and update), the transposed
 * local Jacobian applied to w (N outputs for P, N + NC for Q) by the reverse functions (vjp) and column by column
 * in dual numbers (jvp). Inputs in the order phi (a, i, j) 9N, mu (k, slot) 5NC, delta-phi (a, slot) 5N, theta. */
template <int N, int NC, bool P2>
static int cell_check_t(const gp_params *p, const double *phi, const double *mu, const double *dphi, int64_t c,
                        int stage, const double *w, double *vjp, double *jvp) {
  using D = Dual<1>;
  const int64_t ry = p->mesh_y + 6, n = (p->mesh_x + 6) * ry;
  const int64_t y = c % ry;
  Model<double> m;
  if (make_model(p, &m)) return -1;
  RegStore<N, NC, double> rs;
  gather<N, NC>(phi, mu, dphi, n, ry, c, rs);
  RevAcc<N, NC> acc;
  acc.clear();
  if (stage == 0) phi_delta_vjp<N, NC>(m, RegView<N, NC, double, P2>{&rs, 0, 0}, w, acc);
  else cell_update_vjp<N, NC>(m, RegView<N, NC, double, P2>{&rs, 0, 0}, y, w, w + N, acc);
  const int nth = kThetaN<N, NC>;
  int o = 0;
  for (int a = 0; a < N; a++)
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) vjp[o++] = acc.p[a][i][j];
  for (int k = 0; k < NC; k++)
    for (int s = 0; s < 5; s++) vjp[o++] = acc.u[k][s];
  for (int a = 0; a < N; a++)
    for (int s = 0; s < 5; s++) vjp[o++] = acc.d[a][s];
  for (int t = 0; t < nth; t++) vjp[o++] = acc.th[t];
  const int nin = o;
  for (int dir = 0; dir < nin; dir++) {
    Model<D> md;
    make_model(p, &md);
    const int tdir = dir - (9 * N + 5 * NC + 5 * N);
    make_dual_model_w<1>(p, [=](int t, int) { return t == tdir ? 1.0 : 0.0; }, &md);
    RegStore<N, NC, D> ds;
    for (int a = 0; a < N; a++)
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) ds.p[a][i][j] = D(rs.p[a][i][j]);
    for (int k = 0; k < NC; k++)
      for (int s = 0; s < 5; s++) ds.u[k][s] = D(rs.u[k][s]);
    for (int a = 0; a < N; a++)
      for (int s = 0; s < 5; s++) ds.d[a][s] = D(rs.d[a][s]);
    if (dir < 9 * N) ds.p[dir / 9][(dir % 9) / 3][dir % 3].d[0] = 1.0;
    else if (dir < 9 * N + 5 * NC) ds.u[(dir - 9 * N) / 5][(dir - 9 * N) % 5].d[0] = 1.0;
    else if (dir < 9 * N + 5 * NC + 5 * N) ds.d[(dir - 9 * N - 5 * NC) / 5][(dir - 9 * N - 5 * NC) % 5].d[0] = 1.0;
    const RegView<N, NC, D, P2> v{&ds, 0, 0};
    double s = 0.0;
    if (stage == 0) {
      D d[N];
      phi_delta<N, NC>(md, v, false, d);
      for (int a = 0; a < N; a++) s += w[a] * d[a].d[0];
    } else {
      D jx0[NC], jxb[NC], jy0[NC], jyl[NC], d[N], pn[N], mn[NC], dm[NC];
      jat_x<N, NC>(md, v, 0, 0, jx0);
      jat_x<N, NC>(md, v, -1, 0, jxb);
      jat_y<N, NC>(md, v, 0, 0, y, jy0);
      jat_y<N, NC>(md, v, 0, -1, y - 1, jyl);
      for (int a = 0; a < N; a++) d[a] = v.dp(a, 0, 0);
      mu_update<N, NC>(md, v, false, d, jx0, jxb, jy0, jyl, pn, mn, dm);
      for (int a = 0; a < N; a++) s += w[a] * pn[a].d[0];
      for (int k = 0; k < NC; k++) s += w[N + k] * mn[k].d[0];
    }
    jvp[dir] = s;
  }
  return nin;
}

extern "C" int cell_check(const gp_params *p, const double *phi, const double *mu, const double *dphi, int64_t c,
                          int stage, const double *w, double *vjp, double *jvp) {
  const int N = p->nphases, NC = p->ncomp - 1;
#define CASE(a, b) \
  if (N == a && NC == b) return cell_check_t<a, b, false>(p, phi, mu, dphi, c, stage, w, vjp, jvp);
  GP_CASES(CASE)
#undef CASE
  return -2;
}

/* nsteps compensated-FP32 steps (run_c32) from the FP64 state (phi, mu), replaced by the decoded final state */
extern "C" int model_step_c32(const gp_params *p, double *phi, double *mu, int64_t nsteps, int pow2) {
  const int N = p->nphases, NC = p->ncomp - 1;
#define CASE(a, b) \
  if (N == a && NC == b) return pow2 ? run_c32<a, b, true>(p, phi, mu, nsteps) : run_c32<a, b, false>(p, phi, mu, nsteps);
  GP_CASES(CASE)
#undef CASE
  return -2;
}

/* c32_encode / c32_decode of n values (tests of the encoding) */
extern "C" void c32_roundtrip(const double *v, int64_t n, float *hi, int8_t *q, double *out) {
  for (int64_t i = 0; i < n; i++) {
    c32_encode(v[i], hi[i], q[i]);
    out[i] = c32_decode(hi[i], q[i]);
  }
}

/* df_encode / df_decode of n values (tests of the two-part phi form) */
extern "C" void df_roundtrip(const double *v, int64_t n, float *hi, float *lo, double *out) {
  for (int64_t i = 0; i < n; i++) {
    df_encode(v[i], hi[i], lo[i]);
    out[i] = df_decode(hi[i], lo[i]);
  }
}
