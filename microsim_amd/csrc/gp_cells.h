/*
 * gp_cells.h -- the update part of the step as a pair of cell-classified kernels (the default path).
 *
 * Every interior cell is either a bulk cell or an interface cell. A bulk cell's radius-2 diamond of phase codes
 * (the 13 cells within Manhattan distance 2) is purely one phase b: then it and its four neighbours are not
 * delta-phi candidates (delta-phi exactly 0), every anti-trapping current on its faces returns +0, every mobility
 * it reads is Mpure[b], and its update is the pure-phase arithmetic of phase b. Every other interior cell is an
 * interface cell.
 *
 * k_cbulk updates the bulk cells of a tile (60-cell y-strip x seg_len x-rows, one wavefront, as the march tiles):
 * it streams the mu rows with a deep prefetch and forms, per lane, the same operations in the same order as the
 * pure-phase step and k_bulk, so the results are bit-identical. phi of a bulk cell is e_b; when the output buffer
 * already holds e_b there (the cell was purely b two states back) it is not rewritten.
 * k_ciface collects the interface cells of a tile into an LDS list and runs the general per-cell update
 * (gp_model.h: jat_x, jat_y, mu_update -- bit-identical to the reference by construction) on 64 of them at a
 * time, so that every lane works on an interface cell, not one in five as in a march row.
 * Both kernels stage the tile's phase codes (rows xa-2 .. xb+2, the strip's 64 lanes) in LDS and apply the same
 * predicate, so each interior cell is written by exactly one of them. They read the codes of the current state
 * and write those of the new one (a second buffer).
 */
#pragma once

#include "gp_march.h"

namespace gpk {

constexpr int kCellSegMax = 16;                /* longest x-segment the kernels stage (LDS sized for it) */
constexpr int kCellRows = kCellSegMax + 4;     /* staged code rows per tile */

/* b when the radius-2 diamond of lane l at staged row r is purely phase b, else -1 (l in 2..61) */
__device__ __forceinline__ int bulk_of(const int8_t *cs, int r, int l) {
  const int b = cs[r * 64 + l];
  if (b < 0) return -1;
  const bool ok = cs[(r - 1) * 64 + l] == b && cs[(r + 1) * 64 + l] == b && cs[(r - 2) * 64 + l] == b &&
                  cs[(r + 2) * 64 + l] == b && cs[r * 64 + l - 1] == b && cs[r * 64 + l + 1] == b &&
                  cs[r * 64 + l - 2] == b && cs[r * 64 + l + 2] == b && cs[(r - 1) * 64 + l - 1] == b &&
                  cs[(r - 1) * 64 + l + 1] == b && cs[(r + 1) * 64 + l - 1] == b && cs[(r + 1) * 64 + l + 1] == b;
  return ok ? b : -1;
}

/* per-wavefront tile geometry, shared by the two kernels */
struct CellTile {
  int64_t y, yl, xa, xb;
  bool out_lane, img_y;
  __device__ CellTile(int w, int lane, int64_t mx, int64_t my, int seg_len, int nstrips) {
    const int strip = w % nstrips, seg = w / nstrips;
    const int64_t ry = my + 6;
    y = 3 + (int64_t)strip * kMarchOut - 2 + lane;
    yl = y < ry - 1 ? y : ry - 1;
    xa = 3 + (int64_t)seg * seg_len;
    xb = (xa + seg_len - 1 < mx + 2) ? xa + seg_len - 1 : mx + 2;
    out_lane = lane >= 2 && lane <= 61 && y <= my + 2;
    img_y = out_lane && (y < 7 || y > my - 2);
  }
};

/* the tile's phase codes, rows xa-2 .. xb+2, into LDS (each lane its own column) */
__device__ __forceinline__ void stage_codes(int8_t *cs, const int8_t *__restrict__ codes, const CellTile &t,
                                            int64_t ps, int lane) {
  const int nr = (int)(t.xb - t.xa) + 5;
  for (int r = 0; r < nr; r++) cs[r * 64 + lane] = codes[(t.xa - 2 + r) * ps + t.yl + kYOff];
}

/* View of the general per-cell functions whose delta-phi reads k_dphi's values on candidate cells (codes of the
 * 5-point neighbourhood not uniform) and exactly 0 elsewhere, as the step defines them */
template <typename R, bool POW2>
struct CodesView {
  static constexpr bool kPow2 = POW2;
  const R *phi, *mu, *dphi;
  int64_t fs, ps, c;
  const int8_t *cs;
  int r, l;
  __device__ R ph(int a, int i, int j) const { return phi[a * fs + c + i * ps + j]; }
  __device__ R m(int k, int i, int j) const { return mu[k * fs + c + i * ps + j]; }
  __device__ R dp(int a, int i, int j) const {
    const int rr = r + i, ll = l + j;
    const int8_t k0 = cs[rr * 64 + ll];
    const bool cand = k0 < 0 || cs[(rr - 1) * 64 + ll] != k0 || cs[(rr + 1) * 64 + ll] != k0 ||
                      cs[rr * 64 + ll - 1] != k0 || cs[rr * 64 + ll + 1] != k0;
    return cand ? dphi[a * fs + c + i * ps + j] : R(0.0);
  }
  __device__ CodesView shift(int i, int j) const {
    CodesView v = *this;
    v.c = c + i * ps + j;
    v.r = r + i;
    v.l = l + j;
    return v;
  }
};

/* the bulk update of one cell in phase b (UNIFORM: b is the same on every lane, read as a scalar) */
template <int N, int NC, typename R, bool P2, bool STATS, bool UNIFORM, int NSA>
__device__ __forceinline__ void cbulk_cell(const Model<R> &m, int b, int64_t x, const CellTile &t, const R *m0,
                                           const R *m1, const R *m2, const R *mud, const R *muu, R divjat0,
                                           R *__restrict__ phi_o, R *__restrict__ mu_o,
                                           int8_t *__restrict__ codes_out, const Maps &maps, int64_t fs, int64_t ps,
                                           int64_t mx, int keep, R (&st)[NSA]) {
  using Vw = View<R, P2>;
  R Mb[NC][NC], pe[N];
#pragma unroll
  for (int k = 0; k < NC; k++)
#pragma unroll
    for (int l = 0; l < NC; l++) {
      if (UNIFORM) {
        Mb[k][l] = m.Mpure[b][k][l];
      } else {
        Mb[k][l] = m.Mpure[0][k][l];
#pragma unroll
        for (int a = 1; a < N; a++)
          if (b == a) Mb[k][l] = m.Mpure[a][k][l];
      }
    }
#pragma unroll
  for (int a = 0; a < N; a++) pe[a] = a == b ? R(1.0) : R(0.0);
  R divflux[NC], deltamu[NC], munew[NC];
#pragma unroll
  for (int k = 0; k < NC; k++) {
    divflux[k] = R(0.0);
#pragma unroll
    for (int l = 0; l < NC; l++) {
      const R Fb = Mb[k][l] * divx<Vw>(m, (m1[l] - m0[l]));
      const R Ff = Mb[k][l] * divx<Vw>(m, (m2[l] - m1[l]));
      const R Fy = Mb[k][l] * divy<Vw>(m, (mud[l] - m1[l]));
      const R Fyl = Mb[k][l] * divy<Vw>(m, (m1[l] - muu[l])); /* = Fy of the lane below, bit for bit */
      divflux[k] += divx<Vw>(m, (Ff - Fb));
      divflux[k] += divy<Vw>(m, (Fy - Fyl));
    }
  }
  if (kDirectBulkMu<R>) {
    R dc[NC];
#pragma unroll
    for (int k = 0; k < NC; k++) dc[k] = m.dt * (divflux[k]) - divjat0;
#pragma unroll
    for (int k = 0; k < NC; k++) {
      R s = R(0);
      if (UNIFORM) {
        s = Mu_lin<NC>(m, dc, b, k);
      } else {
#pragma unroll
        for (int a = 0; a < N; a++)
          if (a == b) s = Mu_lin<NC>(m, dc, a, k);
      }
      deltamu[k] = s;
      munew[k] = m1[k] + deltamu[k];
    }
  } else {
    R c_new[NC];
#pragma unroll
    for (int k = 0; k < NC; k++) {
      R s = R(0);
      if (UNIFORM) {
        s = c_mu<NC>(m, m1, b, k);
      } else {
#pragma unroll
        for (int a = 0; a < N; a++)
          if (a == b) s = c_mu<NC>(m, m1, a, k);
      }
      c_new[k] = s + m.dt * (divflux[k]) - divjat0;
    }
#pragma unroll
    for (int k = 0; k < NC; k++) {
      R s = R(0);
      if (UNIFORM) {
        s = Mu<NC>(m, c_new, b, k);
      } else {
#pragma unroll
        for (int a = 0; a < N; a++)
          if (a == b) s = Mu<NC>(m, c_new, a, k);
      }
      munew[k] = s;
      deltamu[k] = munew[k] - m1[k];
    }
  }
  const int64_t c = x * ps + t.y + kYOff;
  const int8_t code = (int8_t)b;
  /* keep: the output buffers hold the state before the current one and its codes; a cell that was purely
   * phase b there already holds phi = e_b, so only mu is written */
  if (!(keep && codes_out[c] == code)) {
#pragma unroll
    for (int a = 0; a < N; a++) phi_o[a * fs + c] = pe[a];
    codes_out[c] = code;
  }
#pragma unroll
  for (int k = 0; k < NC; k++) mu_o[k * fs + c] = munew[k];
  if (t.img_y || x < 7 || x > mx - 2) {
    write_images(codes_out, 1, &code, 0, maps, x, t.y, 0, ps);
    write_images(phi_o, N, pe, 0, maps, x, t.y, fs, ps);
    write_images(mu_o, NC, munew, 1, maps, x, t.y, fs, ps);
  }
  if (STATS) {
    const R z[N] = {};
    stats_add<N, NC>(st, pe, z, munew, deltamu);
  }
}

constexpr int kCellBulkDepth = 4; /* mu rows in flight ahead of the current one */

template <int N, int NC, typename R, bool P2, bool STATS>
__global__ __launch_bounds__(256) void k_cbulk(const Model<R> *__restrict__ mp, const R *__restrict__ mu,
                                               const int8_t *__restrict__ codes_in, R *__restrict__ phi_o,
                                               R *__restrict__ mu_o, int8_t *__restrict__ codes_out, Maps maps,
                                               double *__restrict__ partials, int64_t pstride, int64_t fs, int64_t ps,
                                               int64_t mx, int64_t my, int seg_len, int nstrips, int nwaves,
                                               int keep) {
  using Vw = View<R, P2>;
  constexpr int NS = 3 * N + 3 * NC + 1;
  constexpr int NSA = STATS ? NS : 1;
  __shared__ int8_t cs_all[4][kCellRows * 64];
  const Model<R> &m = *mp;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int w = blockIdx.x * 4 + wv;
  int8_t *cs = cs_all[wv];
  R st[NSA];
  if (STATS) stats_init<N, NC>(st);
  const CellTile t(w, lane, mx, my, seg_len, nstrips);
  if (w < nwaves) stage_codes(cs, codes_in, t, ps, lane);
  __syncthreads();
  if (w < nwaves) {
    const R zj = R(0.0) * m.jat_fac; /* (j - j_prev) * jat_fac with both currents +0 */
    R divjat0 = divx<Vw>(m, zj);
    divjat0 += divy<Vw>(m, zj);
    const int64_t c0 = t.yl + kYOff;
    R m0[NC], m1[NC], q[kCellBulkDepth][NC];
#pragma unroll
    for (int k = 0; k < NC; k++) {
      m0[k] = mu[k * fs + (t.xa - 1) * ps + c0];
      m1[k] = mu[k * fs + t.xa * ps + c0];
    }
#pragma unroll
    for (int d = 0; d < kCellBulkDepth; d++)
      if (t.xa + 1 + d <= t.xb + 1) {
#pragma unroll
        for (int k = 0; k < NC; k++) q[d][k] = mu[k * fs + (t.xa + 1 + d) * ps + c0];
      }
    for (int64_t x = t.xa; x <= t.xb; x++) {
      R m2[NC];
#pragma unroll
      for (int k = 0; k < NC; k++) m2[k] = q[0][k];
#pragma unroll
      for (int d = 0; d + 1 < kCellBulkDepth; d++)
#pragma unroll
        for (int k = 0; k < NC; k++) q[d][k] = q[d + 1][k];
      if (x + 1 + kCellBulkDepth <= t.xb + 1) {
#pragma unroll
        for (int k = 0; k < NC; k++) q[kCellBulkDepth - 1][k] = mu[k * fs + (x + 1 + kCellBulkDepth) * ps + c0];
      }
      R mud[NC], muu[NC];
#pragma unroll
      for (int k = 0; k < NC; k++) {
        mud[k] = lane_dn(m1[k]);
        muu[k] = lane_up(m1[k]);
      }
      const int bl = t.out_lane ? bulk_of(cs, (int)(x - t.xa) + 2, lane) : -1;
      /* a row whose output lanes are all bulk in one phase (the common case) takes the phase as a scalar */
      const int b0 = __builtin_amdgcn_readlane(bl, 2);
      if (b0 >= 0 && __all(!t.out_lane || bl == b0)) {
        if (t.out_lane)
          cbulk_cell<N, NC, R, P2, STATS, true>(m, b0, x, t, m0, m1, m2, mud, muu, divjat0, phi_o, mu_o, codes_out,
                                                maps, fs, ps, mx, keep, st);
      } else if (bl >= 0)
        cbulk_cell<N, NC, R, P2, STATS, false>(m, bl, x, t, m0, m1, m2, mud, muu, divjat0, phi_o, mu_o, codes_out,
                                               maps, fs, ps, mx, keep, st);
#pragma unroll
      for (int k = 0; k < NC; k++) {
        m0[k] = m1[k];
        m1[k] = m2[k];
      }
    }
  }
  if (!STATS) return;
  stats_partials<N, NC>(st, partials, pstride, w, nwaves, lane);
}

template <int N, int NC, typename R, bool P2, bool STATS>
__global__ __launch_bounds__(256) void k_ciface(const Model<R> *__restrict__ mp, const R *__restrict__ phi,
                                                const R *__restrict__ mu, const R *__restrict__ dphi,
                                                const int8_t *__restrict__ codes_in, R *__restrict__ phi_o,
                                                R *__restrict__ mu_o, int8_t *__restrict__ codes_out, Maps maps,
                                                double *__restrict__ partials, int64_t pstride, int64_t fs,
                                                int64_t ps, int64_t mx, int64_t my, int seg_len, int nstrips,
                                                int nwaves) {
  constexpr int NS = 3 * N + 3 * NC + 1;
  constexpr int NSA = STATS ? NS : 1;
  __shared__ int8_t cs_all[4][kCellRows * 64];
  __shared__ uint16_t list_all[4][kCellSegMax * 64];
  const Model<R> &m = *mp;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int w = blockIdx.x * 4 + wv;
  int8_t *cs = cs_all[wv];
  uint16_t *list = list_all[wv];
  R st[NSA];
  if (STATS) stats_init<N, NC>(st);
  const CellTile t(w, lane, mx, my, seg_len, nstrips);
  if (w < nwaves) stage_codes(cs, codes_in, t, ps, lane);
  __syncthreads();
  int count = 0;
  if (w < nwaves) {
    for (int64_t x = t.xa; x <= t.xb; x++) {
      const bool iface = t.out_lane && bulk_of(cs, (int)(x - t.xa) + 2, lane) < 0;
      const uint64_t mask = __ballot(iface);
      const int pre = __popcll(mask & ((1ull << lane) - 1ull));
      if (iface) list[count + pre] = (uint16_t)((x - t.xa) * 64 + lane);
      count += __popcll(mask);
    }
  }
  __syncthreads();
  for (int i = lane; i < count; i += 64) {
    const int e = list[i];
    const int xr = e >> 6, l = e & 63;
    const int64_t x = t.xa + xr, y = t.y - lane + l;
    const int64_t c = x * ps + y + kYOff;
    /* gather the cell's stencil (delta-phi gated by the phase codes), then compute from registers */
    RegStore<N, NC, R> rs;
    {
      const CodesView<R, P2> g{phi, mu, dphi, fs, ps, c, cs, xr + 2, l};
      const int oi[5] = {0, 1, -1, 0, 0}, oj[5] = {0, 0, 0, 1, -1};
#pragma unroll
      for (int a = 0; a < N; a++)
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
          for (int j = 0; j < 3; j++) rs.p[a][i][j] = g.ph(a, i - 1, j - 1);
#pragma unroll
      for (int o = 0; o < 5; o++) {
#pragma unroll
        for (int k = 0; k < NC; k++) rs.u[k][o] = g.m(k, oi[o], oj[o]);
#pragma unroll
        for (int a = 0; a < N; a++) rs.d[a][o] = g.dp(a, oi[o], oj[o]);
      }
    }
    const RegView<N, NC, R, P2> v{&rs, 0, 0};
    R jx0[NC], jxb[NC], jy0[NC], jyl[NC], d[N], pn[N], mn[NC], dm[NC];
    jat_x<N, NC>(m, v, 0, 0, jx0);
    jat_x<N, NC>(m, v, -1, 0, jxb);
    jat_y<N, NC>(m, v, 0, 0, y, jy0);
    jat_y<N, NC>(m, v, 0, -1, y - 1, jyl);
#pragma unroll
    for (int a = 0; a < N; a++) d[a] = v.dp(a, 0, 0);
    mu_update<N, NC>(m, v, false, d, jx0, jxb, jy0, jyl, pn, mn, dm);
#pragma unroll
    for (int a = 0; a < N; a++) phi_o[a * fs + c] = pn[a];
#pragma unroll
    for (int k = 0; k < NC; k++) mu_o[k * fs + c] = mn[k];
    const int8_t code = (int8_t)pure_phase<N>(pn);
    codes_out[c] = code;
    if (y < 7 || y > my - 2 || x < 7 || x > mx - 2) {
      write_images(codes_out, 1, &code, 0, maps, x, y, 0, ps);
      write_images(phi_o, N, pn, 0, maps, x, y, fs, ps);
      write_images(mu_o, NC, mn, 1, maps, x, y, fs, ps);
    }
    if (STATS) stats_add<N, NC>(st, pn, d, mn, dm);
  }
  if (!STATS) return;
  stats_partials<N, NC>(st, partials, pstride, w, nwaves, lane);
}

}  // namespace gpk
