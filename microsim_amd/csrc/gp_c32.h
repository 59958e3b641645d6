/*
 * gp_c32.h -- kernels of the compensated FP32 mode (GP_FP32C).
 *
 * State: every mu component is an FP32 plane (hi) plus an int8 plane (q), the value being
 * c32_decode(hi, q) = hi + q ulp32(hi) / 128; every phi component is a pair of FP32 planes (hi, lo), the value being
 * df_decode(hi, lo) = hi + lo (gp_model.h). The step is the cell-classified one of gp_cells.h:
 *   k_cbulk<..., C = true>  bulk cells: the FP32 closed-form increment of mu (from the hi parts), added to the
 *                           decoded value and re-encoded (the increments are not rounded away); phi stays e_b
 *                           with lo = 0;
 *   k_dlist_c               delta-phi of the listed candidates in FP64 from the decoded state (FP64 plane);
 *   k_ilist_c               the general update of the listed interface cells in FP64 from the decoded state,
 *                           encoded on store.
 * A cell's phase code is b when its decoded phi is exactly e_b (hi = e_b and lo = 0), so bulk cells carry lo = 0,
 * and a cell that k_cbulk keeps (pure two states back) already holds lo = 0.
 * Measured on the ternary seed workload (256^2, CPU build of the same arithmetic), the FP64 run started from the
 * FP32-rounded initial state and this mode agree to within 1.1x in max phi error through 2000 steps, where the
 * plain FP32 mode is 5-30x above it: the FP32 mode's error comes from storing each step's result in FP32 (bulk mu
 * increments of a few ulp) and from FP32 arithmetic at the interfaces, and this mode removes both.
 *
 * Also: the codes of an uploaded state, and the widening / narrowing of whole states to and from FP64 planes,
 * which the smoothing steps, delta-phi downloads, invariants and the adjoint use (FP64 kernels on the decoded
 * state).
 */
#pragma once

#include "gp_cells.h"

namespace gpk {

/* phase codes of every allocated cell from the decoded phi (b when it is exactly e_b, else -1) */
template <int N>
__global__ __launch_bounds__(256) void k_c32_codes(const float *__restrict__ phi, const float *__restrict__ plo,
                                                   int8_t *__restrict__ codes, int64_t fs, int64_t ps, int64_t rx,
                                                   int64_t ry) {
  const int64_t y = (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
  const int64_t x = (int64_t)blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x >= rx || y >= ry) return;
  const int64_t c = x * ps + y + kYOff;
  double ph[N];
#pragma unroll
  for (int a = 0; a < N; a++) ph[a] = df_decode(phi[a * fs + c], plo[a * fs + c]);
  codes[c] = (int8_t)pure_phase<N>(ph);
}

/* n elements of (hi, lo) FP32 planes -> FP64 planes (the same element offsets) */
static __global__ __launch_bounds__(256) void k_df_widen(const float *__restrict__ hi, const float *__restrict__ lo,
                                                         double *__restrict__ out, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = df_decode(hi[i], lo[i]);
}

/* n elements of FP64 planes -> (hi, lo) FP32 planes */
static __global__ __launch_bounds__(256) void k_df_narrow(const double *__restrict__ in, float *__restrict__ hi,
                                                          float *__restrict__ lo, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    float h, l;
    df_encode(in[i], h, l);
    hi[i] = h;
    lo[i] = l;
  }
}

/* n elements of (hi, q) planes -> FP64 planes (the same element offsets) */
static __global__ __launch_bounds__(256) void k_c32_widen(const float *__restrict__ hi, const int8_t *__restrict__ q,
                                                          double *__restrict__ out, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = c32_decode(hi[i], q[i]);
}

/* n elements of FP64 planes -> (hi, q) planes */
static __global__ __launch_bounds__(256) void k_c32_narrow(const double *__restrict__ in, float *__restrict__ hi,
                                                           int8_t *__restrict__ q, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    float h;
    int8_t e;
    c32_encode(in[i], h, e);
    hi[i] = h;
    q[i] = e;
  }
}

/* decoded phi of cell c's 3x3 neighbourhood and decoded mu of its 5-point one (slots as RegStore) */
template <int N, int NC>
__device__ __forceinline__ void c32_gather(const float *__restrict__ phi, const float *__restrict__ plo,
                                           const float *__restrict__ mu, const int8_t *__restrict__ mlo, int64_t c,
                                           int64_t fs, int64_t ps, bool mu5, RegStore<N, NC, double> &rs) {
  float ph[N][3][3], pl[N][3][3], mh[NC][5];
  int8_t mq[NC][5];
  const int oi[5] = {0, 1, -1, 0, 0}, oj[5] = {0, 0, 0, 1, -1};
  /* every load first, then the decoding */
#pragma unroll
  for (int a = 0; a < N; a++)
#pragma unroll
    for (int di = 0; di < 3; di++)
#pragma unroll
      for (int dj = 0; dj < 3; dj++) {
        const int64_t o = a * fs + c + (di - 1) * ps + (dj - 1);
        ph[a][di][dj] = phi[o];
        pl[a][di][dj] = plo[o];
      }
#pragma unroll
  for (int o = 0; o < 5; o++) {
    if (o > 0 && !mu5) break;
    const int64_t co = c + oi[o] * ps + oj[o];
#pragma unroll
    for (int k = 0; k < NC; k++) {
      mh[k][o] = mu[k * fs + co];
      mq[k][o] = mlo[k * fs + co];
    }
  }
#pragma unroll
  for (int a = 0; a < N; a++)
#pragma unroll
    for (int di = 0; di < 3; di++)
#pragma unroll
      for (int dj = 0; dj < 3; dj++) rs.p[a][di][dj] = df_decode(ph[a][di][dj], pl[a][di][dj]);
#pragma unroll
  for (int o = 0; o < 5; o++) {
    if (o > 0 && !mu5) break;
#pragma unroll
    for (int k = 0; k < NC; k++) rs.u[k][o] = c32_decode(mh[k][o], mq[k][o]);
  }
}

/* delta-phi (FP64) of the listed candidate cells from the decoded state: a grid-stride loop of full wavefronts */
template <int N, int NC, bool P2>
__global__ __launch_bounds__(256) void k_dlist_c(const Model<double> *__restrict__ mp, const float *__restrict__ phi,
                                                 const float *__restrict__ plo, const float *__restrict__ mu,
                                                 const int8_t *__restrict__ mlo, double *__restrict__ dphi, int64_t fs,
                                                 int64_t ps, const int32_t *__restrict__ clist,
                                                 const int *__restrict__ ccount) {
  const int n = *ccount;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int64_t c = clist[i];
    RegStore<N, NC, double> rs;
    c32_gather<N, NC>(phi, plo, mu, mlo, c, fs, ps, false, rs);
    const RegView<N, NC, double, P2, true> v{&rs, 0, 0};
    double d[N];
    phi_delta<N, NC>(*mp, v, false, d);
#pragma unroll
    for (int a = 0; a < N; a++) dphi[a * fs + c] = d[a];
  }
}

/* The general update of the listed interface cells in FP64 from the decoded state (k_ilist's gather and
 * arithmetic), the new values encoded on store and their codes taken from the decoded stored values. */
template <int N, int NC, bool P2, bool STATS>
__global__ __launch_bounds__(256) void k_ilist_c(const Model<double> *__restrict__ mp, const float *__restrict__ phi,
                                                 const float *__restrict__ plo, const float *__restrict__ mu,
                                                 const int8_t *__restrict__ mlo, const double *__restrict__ dphi,
                                                 const int8_t *__restrict__ codes_in, float *__restrict__ phi_o,
                                                 float *__restrict__ plo_o, float *__restrict__ mu_o,
                                                 int8_t *__restrict__ mlo_o, int8_t *__restrict__ codes_out,
                                                 double *__restrict__ partials, int64_t pstride, int64_t fs,
                                                 int64_t ps, const int32_t *__restrict__ ilist,
                                                 const int *__restrict__ icount) {
  constexpr int NS = 3 * N + 3 * NC + 1;
  constexpr int NSA = STATS ? NS : 1;
  const Model<double> &m = *mp;
  const int lane = threadIdx.x & 63;
  const int w = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int nw = gridDim.x * 4;
  double st[NSA];
  if (STATS) stats_init<N, NC>(st);
  const int n = *icount;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int64_t c = ilist[i];
    const int64_t y = (c - kYOff) % ps;
    int8_t cd[5][5];
#pragma unroll
    for (int di = -2; di <= 2; di++)
#pragma unroll
      for (int dj = -2; dj <= 2; dj++)
        if ((di < 0 ? -di : di) + (dj < 0 ? -dj : dj) <= 2) cd[di + 2][dj + 2] = codes_in[c + di * ps + dj];
    RegStore<N, NC, double> rs;
    c32_gather<N, NC>(phi, plo, mu, mlo, c, fs, ps, true, rs);
    const int oi[5] = {0, 1, -1, 0, 0}, oj[5] = {0, 0, 0, 1, -1};
#pragma unroll
    for (int o = 0; o < 5; o++) {
      const int64_t co = c + oi[o] * ps + oj[o];
      const int r = oi[o] + 2, q = oj[o] + 2;
      const int8_t k0 = cd[r][q];
      const bool cand = k0 < 0 || cd[r - 1][q] != k0 || cd[r + 1][q] != k0 || cd[r][q - 1] != k0 || cd[r][q + 1] != k0;
#pragma unroll
      for (int a = 0; a < N; a++) {
        const double v = dphi[a * fs + co];
        rs.d[a][o] = cand ? v : 0.0;
      }
    }
    const RegView<N, NC, double, P2, true> v{&rs, 0, 0};
    double jx0[NC], jxb[NC], jy0[NC], jyl[NC], d[N], pn[N], mn[NC], dm[NC];
    jat_x<N, NC>(m, v, 0, 0, jx0);
    jat_x<N, NC>(m, v, -1, 0, jxb);
    jat_y<N, NC>(m, v, 0, 0, y, jy0);
    jat_y<N, NC>(m, v, 0, -1, y - 1, jyl);
#pragma unroll
    for (int a = 0; a < N; a++) d[a] = v.dp(a, 0, 0);
    mu_update<N, NC>(m, v, false, d, jx0, jxb, jy0, jyl, pn, mn, dm);
    float ph[N], pl[N], mh[NC];
    int8_t mq[NC];
    double pd[N], md[NC];
#pragma unroll
    for (int a = 0; a < N; a++) {
      df_encode(pn[a], ph[a], pl[a]);
      pd[a] = df_decode(ph[a], pl[a]);
      phi_o[a * fs + c] = ph[a];
      plo_o[a * fs + c] = pl[a];
    }
#pragma unroll
    for (int k = 0; k < NC; k++) {
      c32_encode(mn[k], mh[k], mq[k]);
      md[k] = c32_decode(mh[k], mq[k]);
      mu_o[k * fs + c] = mh[k];
      mlo_o[k * fs + c] = mq[k];
    }
    codes_out[c] = (int8_t)pure_phase<N>(pd);
    if (STATS) stats_add<N, NC>(st, pd, d, md, dm);
  }
  if (!STATS) return;
  stats_partials<N, NC>(st, partials, pstride, w, nw, lane);
}

}  // namespace gpk
