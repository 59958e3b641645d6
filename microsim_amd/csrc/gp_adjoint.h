/*
 * gp_adjoint.h -- adjoint (reverse-mode gradient) of the FP64 solver step, with checkpoint and recompute.
 *
 * Included by gp_hip.hip after the handle and its helpers; not a public header.
 *
 * One forward step maps the state s_t (phi, mu, halo-inclusive) to s_{t+1} in three stages, each a per-cell
 * function of a fixed stencil:
 *   P  delta-phi at every cell of [2, mesh+3]^2 from phi on its 3x3 stencil and mu at the cell
 *      (phi_delta: calculate_divergence_phasefield_2D + simplex projection),
 *   Q  new phi and mu of every interior cell from phi on its 3x3 stencil, mu and delta-phi on its 5-point
 *      stencil (jat_x / jat_y / mu_update: calculate_fluxes_concentration_2D and
 *      calculate_divergence_concentration_2D),
 *   B  the boundary copies (each halo cell takes the value of one interior cell).
 * The adjoint step runs them backwards, transposed: for a cell, the block of its local Jacobian that a kernel
 * needs is evaluated exactly with forward-mode dual numbers (gp_dual.h) through the same per-cell functions the
 * forward step uses, K input directions per pass, and the products with the incoming adjoint are scatter-added
 * (FP64 atomics) onto the stencil's inputs and onto the parameter gradient. Cells whose incoming adjoint is zero
 * are skipped. B's transpose adds each halo cell's adjoint to its source cell.
 *
 * Parameters theta and the dual model that seeds them: gp_theta.h.
 */
#pragma once

#include <vector>

#include "gp_theta.h"

namespace gpk {

constexpr int kAdjK = 8; /* directions per dual pass */

/* Stencil accessor that seeds the inputs of one stage. MODE 0 (stage P): phi on the 3x3 stencil, mu at the
 * centre. MODE 1 (stage Q): phi on the 3x3 stencil, mu and delta-phi on the 5-point stencil. Directions:
 *   phi_a at (i,j)       a*9 + (i+1)*3 + (j+1)
 *   mu_k                 9N + k (P)  |  9N + k*5 + o(i,j) (Q)
 *   dphi_a (Q)           9N + 5NC + a*5 + o(i,j)
 * with o(0,0) = 0, o(1,0) = 1, o(-1,0) = 2, o(0,1) = 3, o(0,-1) = 4. */
__device__ __forceinline__ int five_point(int i, int j) {
  return (i == 0 && j == 0) ? 0 : (i == 1 && j == 0) ? 1 : (i == -1 && j == 0) ? 2 : (i == 0 && j == 1) ? 3 : 4;
}

template <int N, int NC, int K, bool P2, int MODE>
struct SeedView {
  static constexpr bool kPow2 = P2;
  const double *phi, *mu, *dphi;
  int64_t fs, ps, c;
  int si, sj, chunk0;
  __device__ Dual<K> seed(double x, int dir) const {
    Dual<K> r(x);
    const int k = dir - chunk0;
    if (k >= 0 && k < K) r.d[k] = 1.0;
    return r;
  }
  __device__ Dual<K> ph(int a, int i, int j) const {
    const int oi = si + i, oj = sj + j;
    return seed(phi[a * fs + c + oi * ps + oj], a * 9 + (oi + 1) * 3 + (oj + 1));
  }
  __device__ Dual<K> m(int k, int i, int j) const {
    const int oi = si + i, oj = sj + j;
    const int dir = MODE == 0 ? 9 * N + k : 9 * N + k * 5 + five_point(oi, oj);
    return seed(mu[k * fs + c + oi * ps + oj], dir);
  }
  __device__ Dual<K> dp(int a, int i, int j) const {
    const int oi = si + i, oj = sj + j;
    return seed(dphi[a * fs + c + oi * ps + oj], 9 * N + 5 * NC + a * 5 + five_point(oi, oj));
  }
  __device__ SeedView shift(int i, int j) const {
    SeedView v = *this;
    v.si += i;
    v.sj += j;
    return v;
  }
};

/* scatter one direction's adjoint contribution g onto the stencil input it belongs to */
template <int N, int NC, int MODE>
__device__ __forceinline__ void scatter_dir(int dir, double g, int64_t c, int64_t fs, int64_t ps, double *gphi,
                                           double *gmu, double *gdphi, double *th, int nth) {
  if (dir < 9 * N) {
    const int a = dir / 9, r = dir % 9, i = r / 3 - 1, j = r % 3 - 1;
    atomicAdd(&gphi[a * fs + c + i * ps + j], g);
    return;
  }
  dir -= 9 * N;
  const int nmu = MODE == 0 ? NC : 5 * NC;
  const int o2i[5] = {0, 1, -1, 0, 0}, o2j[5] = {0, 0, 0, 1, -1};
  if (dir < nmu) {
    if (MODE == 0) {
      atomicAdd(&gmu[dir * fs + c], g);
    } else {
      const int k = dir / 5, o = dir % 5;
      atomicAdd(&gmu[k * fs + c + o2i[o] * ps + o2j[o]], g);
    }
    return;
  }
  dir -= nmu;
  if (MODE == 1) {
    if (dir < 5 * N) {
      const int a = dir / 5, o = dir % 5;
      atomicAdd(&gdphi[a * fs + c + o2i[o] * ps + o2j[o]], g);
      return;
    }
    dir -= 5 * N;
  }
  if (dir < nth) th[dir] += g;
}

/* block sum of the per-thread parameter adjoints, one atomic per block and parameter */
template <int NTH>
__device__ __forceinline__ void theta_flush(const double *th, int nth, double *gtheta) {
  __shared__ double red[4][NTH];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int t = 0; t < nth; t++) {
    const double r = wave_sum(th[t]);
    if (lane == 0) red[wv][t] = r;
  }
  __syncthreads();
  if (threadIdx.x == 0)
    for (int t = 0; t < nth; t++) {
      const double s = red[0][t] + red[1][t] + red[2][t] + red[3][t];
      if (s != 0.0) atomicAdd(&gtheta[t], s);
    }
}

template <int N, int NC>
constexpr int kTheta = 3 * (N * (N - 1) / 2) + NC + 1;

/* Stage Q transposed for directions [chunk0, chunk0 + K): one thread per interior cell (blocks of 64 x 4). */
template <int N, int NC, bool P2>
__global__ __launch_bounds__(256) void k_adj_q(const Model<Dual<kAdjK>> *__restrict__ mp, const double *__restrict__ phi,
                                               const double *__restrict__ mu, const double *__restrict__ dphi,
                                               const double *__restrict__ gphi_n, const double *__restrict__ gmu_n,
                                               double *gphi, double *gmu, double *gdphi, double *gtheta, int64_t fs,
                                               int64_t ps, int64_t mx, int64_t my, int chunk0) {
  constexpr int K = kAdjK;
  const int64_t y = 3 + (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
  const int64_t x = 3 + (int64_t)blockIdx.y * 4 + (threadIdx.x >> 6);
  constexpr int NTH = kTheta<N, NC>;
  double th[NTH];
  for (int t = 0; t < NTH; t++) th[t] = 0.0;
  if (x <= mx + 2 && y <= my + 2) {
    const int64_t c = x * ps + y + kYOff;
    double go[N + NC];
    bool any = false;
    for (int a = 0; a < N; a++) go[a] = gphi_n[a * fs + c];
    for (int k = 0; k < NC; k++) go[N + k] = gmu_n[k * fs + c];
    for (int o = 0; o < N + NC; o++) any = any || go[o] != 0.0;
    if (any) {
      const Model<Dual<K>> &m = *mp;
      SeedView<N, NC, K, P2, 1> v{phi, mu, dphi, fs, ps, c, 0, 0, chunk0};
      Dual<K> jx0[NC], jxb[NC], jy0[NC], jyl[NC], d[N], pn[N], mn[NC], dm[NC];
      jat_x<N, NC>(m, v, 0, 0, jx0);
      jat_x<N, NC>(m, v, -1, 0, jxb);
      jat_y<N, NC>(m, v, 0, 0, y, jy0);
      jat_y<N, NC>(m, v, 0, -1, y - 1, jyl);
      for (int a = 0; a < N; a++) d[a] = v.dp(a, 0, 0);
      mu_update<N, NC>(m, v, false, d, jx0, jxb, jy0, jyl, pn, mn, dm);
      for (int k = 0; k < K; k++) {
        double g = 0.0;
        for (int a = 0; a < N; a++) g += go[a] * pn[a].d[k];
        for (int q = 0; q < NC; q++) g += go[N + q] * mn[q].d[k];
        if (g != 0.0) scatter_dir<N, NC, 1>(chunk0 + k, g, c, fs, ps, gphi, gmu, gdphi, th, NTH);
      }
    }
  }
  theta_flush<NTH>(th, NTH, gtheta);
}

/* Stage P transposed: one thread per cell of [2, mesh+3]^2 with a nonzero delta-phi adjoint. */
template <int N, int NC, bool P2>
__global__ __launch_bounds__(256) void k_adj_p(const Model<Dual<kAdjK>> *__restrict__ mp, const double *__restrict__ phi,
                                               const double *__restrict__ mu, const double *__restrict__ gdphi,
                                               double *gphi, double *gmu, double *gtheta, int64_t fs, int64_t ps,
                                               int64_t mx, int64_t my, int chunk0) {
  constexpr int K = kAdjK;
  const int64_t y = 2 + (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
  const int64_t x = 2 + (int64_t)blockIdx.y * 4 + (threadIdx.x >> 6);
  constexpr int NTH = kTheta<N, NC>;
  double th[NTH];
  for (int t = 0; t < NTH; t++) th[t] = 0.0;
  if (x <= mx + 3 && y <= my + 3) {
    const int64_t c = x * ps + y + kYOff;
    double go[N];
    bool any = false;
    for (int a = 0; a < N; a++) go[a] = gdphi[a * fs + c];
    for (int a = 0; a < N; a++) any = any || go[a] != 0.0;
    if (any) {
      const Model<Dual<K>> &m = *mp;
      SeedView<N, NC, K, P2, 0> v{phi, mu, nullptr, fs, ps, c, 0, 0, chunk0};
      Dual<K> d[N];
      phi_delta<N, NC>(m, v, false, d);
      for (int k = 0; k < K; k++) {
        double g = 0.0;
        for (int a = 0; a < N; a++) g += go[a] * d[a].d[k];
        if (g != 0.0) scatter_dir<N, NC, 0>(chunk0 + k, g, c, fs, ps, gphi, gmu, nullptr, th, NTH);
      }
    }
  }
  theta_flush<NTH>(th, NTH, gtheta);
}

/* Stage B transposed for one field: every halo cell's adjoint goes to the interior cell it copies (x maps of
 * the field for halo rows, y maps for halo columns, both for corners), and the halo adjoint is cleared. */
__global__ __launch_bounds__(256) void k_adj_bc(double *g, int ncomp, Maps maps, int f, int64_t fs, int64_t ps,
                                                int64_t rx, int64_t ry) {
  const int64_t y = (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
  const int64_t x = (int64_t)blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x >= rx || y >= ry) return;
  int64_t sxv = x, syv = y;
  for (int k = 0; k < 6; k++) {
    if (maps.hx[k] == x) sxv = maps.sx[f][k];
    if (maps.hy[k] == y) syv = maps.sy[f][k];
  }
  if (sxv == x && syv == y) return; /* interior cell */
  const int64_t c = x * ps + y + kYOff, s = sxv * ps + syv + kYOff;
  for (int a = 0; a < ncomp; a++) {
    const double v = g[a * fs + c];
    if (v != 0.0) atomicAdd(&g[a * fs + s], v);
    g[a * fs + c] = 0.0;
  }
}

}  // namespace gpk
