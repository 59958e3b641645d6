/*
 * gp_march.h -- single-pass step kernel for CDNA4: one wavefront per 60-cell y-strip marching along x.
 *
 * Lane l of a wavefront holds cell y = y0 - 2 + l of the current rows; y-neighbours come from lane shuffles,
 * x-neighbours from the rows kept in registers between march steps. At march step x the wavefront holds the
 * raw rows x-1, x, x+1 (phi, mu and their y+-1 neighbours) and
 *   1. forms the face gradients and averages of rows x-1..x+1,
 *   2. computes delta-phi of row x -- the anisotropic face terms (q_ab, a_c, da_c/dq) once per face and per
 *      unordered phase pair: the back x-face is the previous step's front face and the left y-face is the
 *      right y-face of the lane below; q_ba = -q_ab and a_c(q_ba) = a_c(q_ab) hold bit for bit, so the ordered
 *      pairs reuse them with an exact sign change,
 *   3. computes the anti-trapping currents on the right y-face of row x and the x-face between x-1 and x,
 *   4. updates mu and phi of row x-1 and writes them (plus the halo images of boundary cells).
 * Every value keeps the operation order of the per-cell functions in gp_model.h (which follow the serial
 * reference), so the FP64 results are identical to those of gp_model.h and of the reference.
 * A wavefront whose rows are all bulk (every Laplacian of phi zero, phi >= 0) skips the phase-field work of the
 * row: delta-phi is exactly zero there; one whose three rows are purely one phase takes the pure-phase step.
 * Lanes 2..61 produce output; lanes 0, 1, 62, 63 only feed neighbours.
 *
 * The step is split over tiles (strip x segment, one wavefront each): k_march takes the tiles near an interface,
 * k_bulk the "deep" tiles, whose own cells and whose neighbour tiles' cells were all purely one phase at the
 * start of the step. k_bulk streams mu only, with few registers and a deep row prefetch, so that the bulk of
 * the grid runs at memory speed while the register-heavy interface work stays in k_march.
 */
#pragma once

#include "gp_common.h"
#include "gp_model.h"

namespace gpk {

constexpr int kMarchLanes = 64;
constexpr int kMarchOut = 60;
/* rows the march keeps in flight ahead of the current one (registers: (N + NC) per row) */
template <typename R>
constexpr int kMarchDepth = sizeof(R) == sizeof(float) ? 3 : 2;

/* Neighbour lanes through DPP wavefront shifts (GFX9 wave_shr:1 / wave_shl:1): a VALU move that takes the value
 * of lane-1 (lane_up) or lane+1 (lane_dn) without the LDS round trip of a bpermute; the lane without a source
 * (0, resp. 63) keeps its own value, like __shfl_up / __shfl_down. */
constexpr int kDppWaveShr1 = 0x138, kDppWaveShl1 = 0x130;
__device__ __forceinline__ int lane_up(int v) { return __builtin_amdgcn_update_dpp(v, v, kDppWaveShr1, 0xf, 0xf, false); }
__device__ __forceinline__ int lane_dn(int v) { return __builtin_amdgcn_update_dpp(v, v, kDppWaveShl1, 0xf, 0xf, false); }
__device__ __forceinline__ float lane_up(float v) { return __int_as_float(lane_up(__float_as_int(v))); }
__device__ __forceinline__ float lane_dn(float v) { return __int_as_float(lane_dn(__float_as_int(v))); }
__device__ __forceinline__ double lane_up(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = lane_up((int)(b & 0xffffffffll)), hi = lane_up((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double lane_dn(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = lane_dn((int)(b & 0xffffffffll)), hi = lane_dn((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

template <int N, int NC, typename R>
struct RowRaw {
  R ph[N], phu[N], phd[N];
  R mu[NC], mud[NC];
  int pure; /* b when the cell is purely phase b (phi_b == 1, every other phi == 0), else -1 */
};

template <int N, typename R>
__device__ __forceinline__ int pure_phase(const R *ph) {
  int code = -1, nz = 0;
#pragma unroll
  for (int a = 0; a < N; a++) {
    if (ph[a] == R(1.0)) code = a;
    nz += (ph[a] != R(0.0));
  }
  return nz == 1 ? code : -1;
}

template <int N, int NC, typename R>
__device__ __forceinline__ void row_load(RowRaw<N, NC, R> &r, const R *__restrict__ phi, const R *__restrict__ mu,
                                         int64_t fs, int64_t idx) {
#pragma unroll
  for (int a = 0; a < N; a++) r.ph[a] = phi[a * fs + idx];
#pragma unroll
  for (int k = 0; k < NC; k++) r.mu[k] = mu[k * fs + idx];
}

template <int N, int NC, typename R>
__device__ __forceinline__ void row_halo(RowRaw<N, NC, R> &r) {
  r.pure = pure_phase<N>(r.ph);
#pragma unroll
  for (int a = 0; a < N; a++) {
    r.phu[a] = lane_up(r.ph[a]);
    r.phd[a] = lane_dn(r.ph[a]);
  }
#pragma unroll
  for (int k = 0; k < NC; k++) {
    r.mud[k] = lane_dn(r.mu[k]);
  }
}

/* anti-trapping current on one face, from the two face cells c (centre) and o (other: front or right).
 * gL: normal derivative of the liquid across the face; tL: tangential liquid average; ga, ta per phase. */
template <int N, int NC, typename R>
__device__ __forceinline__ void jat_face(const Model<R> &m, const R *phc, const R *pho, const R (*ccc)[NC],
                                         const R (*cco)[NC], const R *dc, const R *dov, const R *gnorm,
                                         const R *gtan, R *j, unsigned phmask) {
  const int L = N - 1;
#pragma unroll
  for (int k = 0; k < NC; k++) j[k] = R(0.0);
  bool any = false;
#pragma unroll
  for (int a = 0; a < N - 1; a++) any = any || (dc[a] != R(0.0)) || (dov[a] != R(0.0));
  if (!any) return;
  const R nl = qsqrt(gtan[L] * gtan[L] + gnorm[L] * gnorm[L]);
#pragma unroll
  for (int a = 0; a < N - 1; a++) {
    if (!((phmask >> a) & 1u)) continue; /* phase absent: its gradient is zero on every lane */
    const R ng = qsqrt(gtan[a] * gtan[a] + gnorm[a] * gnorm[a]);
    if (!(ng != R(0.0))) continue;
    R s_c = R(0), s_o = R(0);
    if (m.fn_w == 1) {
      s_c = s_obst<N>(phc, a);
      s_o = s_obst<N>(pho, a);
    }
#pragma unroll
    for (int k = 0; k < NC; k++) {
      const R cl_o = cco[L][k], cl_c = ccc[L][k], cs_o = cco[a][k], cs_c = ccc[a][k];
      R sp = gnorm[a] * gnorm[L];
      sp += gtan[a] * gtan[L];
      if (nl > R(0.0)) sp = qdiv(sp, (nl * ng));
      if (m.fn_w == 2)
        j[k] += qdiv(m.dratio[a][k] * fabs(sp) * m.jat_well * ((cl_o - cs_o) * dov[a] + (cl_c - cs_c) * dc[a]) * gnorm[a], ng);
      else
        j[k] += qdiv(m.dratio[a][k] * fabs(sp) * R(0.5) * (s_o * (cl_o - cs_o) * dov[a] + s_c * (cl_c - cs_c) * dc[a]) * gnorm[a], ng);
    }
  }
}

template <int N, int NC, typename R>
__device__ __forceinline__ void cmu_all(const Model<R> &m, const R *mu, R (*cc)[NC]) {
#pragma unroll
  for (int a = 0; a < N; a++)
#pragma unroll
    for (int k = 0; k < NC; k++) cc[a][k] = c_mu<NC>(m, mu, a, k);
}

template <int N, int NC, typename R>
__device__ __forceinline__ void mobility_cell(const Model<R> &m, const R *phi, R (*M)[NC]) {
#pragma unroll
  for (int k = 0; k < NC; k++)
#pragma unroll
    for (int l = 0; l < NC; l++) {
      R s = R(0.0);
#pragma unroll
      for (int b = 0; b < N; b++)
#pragma unroll
        for (int q = 0; q < NC; q++) s += m.DC[b][k][l][q] * phi[b];
      M[k][l] = s;
    }
}

/* Per-lane statistics in the state's precision R: max/min of each phi and mu (exact in R), the sums of squared
 * increments over the lane's cells of one tile (at most a few dozen terms) and the non-finite count; the wavefront
 * and grid reductions are FP64. In FP64 every term is as the reference forms it; in FP32 the per-lane partial sums
 * carry single-precision rounding (relative 1e-6), which the FP32 mode accepts for these diagnostics, in exchange
 * for half the registers. */
template <int N, int NC, int NSA, typename R>
__device__ __forceinline__ void stats_init(R (&st)[NSA]) {
#pragma unroll
  for (int a = 0; a < N; a++) {
    st[a] = -INFINITY;
    st[N + a] = INFINITY;
    st[2 * N + 2 * NC + a] = R(0.0);
  }
#pragma unroll
  for (int k = 0; k < NC; k++) {
    st[2 * N + k] = -INFINITY;
    st[2 * N + NC + k] = INFINITY;
    st[3 * N + 2 * NC + k] = R(0.0);
  }
  st[NSA - 1] = R(0.0);
}

/* accumulate one updated cell: new phi / mu and their increments */
template <int N, int NC, int NSA, typename R>
__device__ __forceinline__ void stats_add(R (&st)[NSA], const R *pn, const R *dph, const R *munew, const R *dmu) {
#pragma unroll
  for (int a = 0; a < N; a++) {
    st[a] = fmax(st[a], pn[a]);
    st[N + a] = fmin(st[N + a], pn[a]);
    st[2 * N + 2 * NC + a] += dph[a] * dph[a];
    if (!isfinite(pn[a])) st[NSA - 1] += R(1.0);
  }
#pragma unroll
  for (int k = 0; k < NC; k++) {
    st[2 * N + k] = fmax(st[2 * N + k], munew[k]);
    st[2 * N + NC + k] = fmin(st[2 * N + NC + k], munew[k]);
    st[3 * N + 2 * NC + k] += dmu[k] * dmu[k];
    if (!isfinite(munew[k])) st[NSA - 1] += R(1.0);
  }
}

/* wavefront reduction of the per-lane statistics into partials[i * pstride + w] (statistic-major, FP64) */
template <int N, int NC, bool PHI_SUMS_ZERO = false, int NSA, typename R>
__device__ __forceinline__ void stats_partials(R (&st)[NSA], double *__restrict__ partials, int64_t pstride,
                                               int w, int nwaves, int lane) {
  /* extrema reduce in the statistics' own precision (exact), sums in FP64. PHI_SUMS_ZERO: the kernel updates
   * bulk cells only, whose squared phi increments are +0 on every lane, so those sums are +0 */
#pragma unroll
  for (int i = 0; i < NSA; i++) {
    double r;
    if (i < N || (i >= 2 * N && i < 2 * N + NC)) r = (double)wave_max(st[i]);
    else if (i < 2 * N + 2 * NC) r = (double)wave_min(st[i]);
    else if (PHI_SUMS_ZERO && i < 3 * N + 2 * NC) r = 0.0;
    else r = wave_sum((double)st[i]);
    if (lane == 0 && w < nwaves) partials[(int64_t)i * pstride + w] = r;
  }
}

/* Tiles: wavefront w of a step kernel handles tile w = (strip w % nstrips, segment w / nstrips), i.e. the 60
 * output cells of a y-strip over the seg_len rows of an x-segment. After each step, flags[tile] holds b when
 * every output cell of the tile is purely phase b (phi_b == 1, every other phi == 0), else -1.
 * deep_phase: b when the tile and its eight neighbours all held purely phase b at the start of the step, else
 * -1. Every cell the tile's update reads (up to three cells outside it) is then purely phase b, delta-phi and the
 * anti-trapping currents vanish in the whole stencil, and the bulk kernel below updates the tile. Tiles on the
 * edge of the tile grid (which write halo images) are never deep. */
__device__ __forceinline__ int deep_phase(const int *__restrict__ flags, int strip, int seg, int nstrips, int nsegs) {
  if (!flags || strip == 0 || seg == 0 || strip >= nstrips - 1 || seg >= nsegs - 1) return -1;
  const int b = flags[seg * nstrips + strip];
  if (b < 0) return -1;
  bool all = true;
#pragma unroll
  for (int dx = -1; dx <= 1; dx++)
#pragma unroll
    for (int ds = -1; ds <= 1; ds++) all = all && flags[(seg + dx) * nstrips + strip + ds] == b;
  return all ? b : -1;
}

/* STATS: accumulate the step statistics (one partial per wavefront at partials[i * pstride + w]); off for the
 * steps that do not report. dphi_in: delta-phi of the candidate cells (k_dphi). codes_out: phase codes of the new
 * state (pure_phase of every written cell, halo images included). flags_in (may be NULL: no tile is deep) and
 * flags_out: tile flags of the previous and of this step; deep tiles are left to k_bulk. */
template <int N, int NC, typename R, bool P2, bool STATS>
__global__ __launch_bounds__(256) void k_march(const Model<R> *__restrict__ mp, const R *__restrict__ phi,
                                               const R *__restrict__ mu, const R *__restrict__ dphi_in,
                                               R *__restrict__ phi_o, R *__restrict__ mu_o,
                                               int8_t *__restrict__ codes_out, Maps maps,
                                               double *__restrict__ partials, int64_t pstride,
                                               const int *__restrict__ flags_in, int *__restrict__ flags_out,
                                               int64_t fs, int64_t ps, int64_t mx, int64_t my, int seg_len,
                                               int nstrips, int nsegs, int nwaves) {
  using Vw = View<R, P2, false, kFastOf<R>>;
  constexpr int NS = 3 * N + 3 * NC + 1;
  const Model<R> &m = *mp;
  const int lane = threadIdx.x & 63;
  const int w = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int64_t ry = my + 6;

  constexpr int NSA = STATS ? NS : 1;
  R st[NSA];
  if (STATS) stats_init<N, NC>(st);

  const int strip = w % nstrips, seg = w / nstrips;
  if (w < nwaves && deep_phase(flags_in, strip, seg, nstrips, nsegs) < 0) {
    const int64_t y0 = 3 + (int64_t)strip * kMarchOut;
    const int64_t y = y0 - 2 + lane;
    const int64_t yl = y < ry - 1 ? y : ry - 1;
    const int64_t xa = 3 + (int64_t)seg * seg_len;
    const int64_t xb = (xa + seg_len - 1 < mx + 2) ? xa + seg_len - 1 : mx + 2;
    const bool out_lane = lane >= 2 && lane <= 61 && y <= my + 2;
    const bool img_y = out_lane && (y < 7 || y > my - 2);

    RowRaw<N, NC, R> A0, A1, A2;
    row_load(A0, phi, mu, fs, (xa - 2) * ps + yl + kYOff);
    row_halo(A0);
    row_load(A1, phi, mu, fs, (xa - 1) * ps + yl + kYOff);
    row_halo(A1);
    /* software pipeline: rows x+1 .. x+kMarchDepth are in flight while step x computes (the row loads are the
     * only dependence of a step on memory, and one step takes less time than an HBM round trip) */
    constexpr int D = kMarchDepth<R>;
    R qph[D][N], qmu[D][NC];
#pragma unroll
    for (int d = 0; d < D; d++)
      if (xa + d <= xb + 2) {
#pragma unroll
        for (int a = 0; a < N; a++) qph[d][a] = phi[a * fs + (xa + d) * ps + yl + kYOff];
#pragma unroll
        for (int k = 0; k < NC; k++) qmu[d][k] = mu[k * fs + (xa + d) * ps + yl + kYOff];
      }

    R dph_prev[N], cc_prev[N][NC], M_prev[NC][NC], jx_prev[NC], jy_prev[NC], Fx_prev[NC][NC];
#pragma unroll
    for (int a = 0; a < N; a++) dph_prev[a] = R(0);
#pragma unroll
    for (int k = 0; k < NC; k++) jx_prev[k] = jy_prev[k] = R(0);
    bool cc_valid = false; /* cc_prev holds c_mu of row x-1 (the pure-phase steps do not form it) */
    int oc = -2; /* purity of this lane's output cells: -2 none yet, b all purely phase b, -1 otherwise */

    /* writes the updated row x-1 (and its halo images) and accumulates the statistics and the purity */
    auto emit = [&](int64_t xo, const R *pn, const R *munew, const R *dph_o, const R *deltamu) {
      if (!out_lane) return;
      const int pc = pure_phase<N>(pn);
      oc = (oc == -2 || oc == pc) ? pc : -1;
      const int64_t c = xo * ps + y + kYOff;
#pragma unroll
      for (int a = 0; a < N; a++) phi_o[a * fs + c] = pn[a];
#pragma unroll
      for (int k = 0; k < NC; k++) mu_o[k * fs + c] = munew[k];
      const int8_t code = (int8_t)pc;
      codes_out[c] = code;
      if (img_y || xo < 7 || xo > mx - 2) {
        write_images(codes_out, 1, &code, 0, maps, xo, y, 0, ps);
        write_images(phi_o, N, pn, 0, maps, xo, y, fs, ps);
        write_images(mu_o, NC, munew, 1, maps, xo, y, fs, ps);
      }
      if (STATS) stats_add<N, NC>(st, pn, dph_o, munew, deltamu);
    };

    for (int64_t x = xa - 1; x <= xb + 1; x++) {
#pragma unroll
      for (int a = 0; a < N; a++) A2.ph[a] = qph[0][a];
#pragma unroll
      for (int k = 0; k < NC; k++) A2.mu[k] = qmu[0][k];
#pragma unroll
      for (int d = 0; d + 1 < D; d++) {
#pragma unroll
        for (int a = 0; a < N; a++) qph[d][a] = qph[d + 1][a];
#pragma unroll
        for (int k = 0; k < NC; k++) qmu[d][k] = qmu[d + 1][k];
      }
      if (x + 1 + D <= xb + 2) {
        const int64_t ni = (x + 1 + D) * ps + yl + kYOff;
#pragma unroll
        for (int a = 0; a < N; a++) qph[D - 1][a] = phi[a * fs + ni];
#pragma unroll
        for (int k = 0; k < NC; k++) qmu[D - 1][k] = mu[k * fs + ni];
      }
      row_halo(A2);

      /* Pure-phase step: every lane of rows x-1, x, x+1 is purely the same phase b, and row x-1 carries no
       * delta-phi and no anti-trapping current. Then delta-phi of row x and every anti-trapping current of the
       * step are exactly zero, each mobility is Mpure[b] and each face average (M + M) * 0.5 equals M, so the
       * general step below reduces to the mu fluxes and the bulk update of phase b; this path forms exactly
       * those values, with the same operations in the same order. */
      {
        bool lp = A0.pure >= 0 && A0.pure == A1.pure && A1.pure == A2.pure;
#pragma unroll
        for (int a = 0; a < N; a++) lp = lp && (dph_prev[a] == R(0.0));
#pragma unroll
        for (int k = 0; k < NC; k++) lp = lp && (jx_prev[k] == R(0.0)) && (jy_prev[k] == R(0.0));
        const int b = __builtin_amdgcn_readfirstlane(A1.pure); /* every lane is active here */
        if (__all(lp && A1.pure == b)) {
          R Fx0[NC][NC];
          if (x >= xa) {
#pragma unroll
            for (int k = 0; k < NC; k++)
#pragma unroll
              for (int l = 0; l < NC; l++) Fx0[k][l] = m.Mpure[b][k][l] * divx<Vw>(m, (A1.mu[l] - A0.mu[l]));
          }
          if (x >= xa + 1) {
            R Fy[NC][NC], Fyl[NC][NC];
#pragma unroll
            for (int k = 0; k < NC; k++)
#pragma unroll
              for (int l = 0; l < NC; l++) Fy[k][l] = m.Mpure[b][k][l] * divy<Vw>(m, (A0.mud[l] - A0.mu[l]));
#pragma unroll
            for (int k = 0; k < NC; k++)
#pragma unroll
              for (int l = 0; l < NC; l++) Fyl[k][l] = lane_up(Fy[k][l]);
            const R zj = R(0.0) * m.jat_fac; /* (j - j_prev) * jat_fac with both currents +0 */
            R divflux[NC], divjat[NC], deltamu[NC], munew[NC];
#pragma unroll
            for (int k = 0; k < NC; k++) {
              divflux[k] = R(0.0);
#pragma unroll
              for (int l = 0; l < NC; l++) {
                divflux[k] += divx<Vw>(m, (Fx0[k][l] - Fx_prev[k][l]));
                divflux[k] += divy<Vw>(m, (Fy[k][l] - Fyl[k][l]));
              }
              divjat[k] = divx<Vw>(m, zj);
              divjat[k] += divy<Vw>(m, zj);
            }
            if (kDirectBulkMu<R>) {
              R dc[NC];
#pragma unroll
              for (int k = 0; k < NC; k++) dc[k] = m.dt * (divflux[k]) - divjat[k];
#pragma unroll
              for (int k = 0; k < NC; k++) {
                deltamu[k] = Mu_lin<NC>(m, dc, b, k);
                munew[k] = A0.mu[k] + deltamu[k];
              }
            } else {
              R c_new[NC];
#pragma unroll
              for (int k = 0; k < NC; k++) c_new[k] = c_mu<NC>(m, A0.mu, b, k) + m.dt * (divflux[k]) - divjat[k];
#pragma unroll
              for (int k = 0; k < NC; k++) {
                munew[k] = Mu<NC>(m, c_new, b, k);
                deltamu[k] = munew[k] - A0.mu[k];
              }
            }
            R pn[N];
#pragma unroll
            for (int a = 0; a < N; a++) pn[a] = A0.ph[a] + dph_prev[a];
            emit(x - 1, pn, munew, dph_prev, deltamu);
          }
          /* carry: row x has delta-phi 0, no currents, mobility Mpure[b] */
#pragma unroll
          for (int a = 0; a < N; a++) dph_prev[a] = R(0.0);
#pragma unroll
          for (int k = 0; k < NC; k++) {
            jy_prev[k] = R(0.0);
            if (x >= xa) jx_prev[k] = R(0.0);
#pragma unroll
            for (int l = 0; l < NC; l++) {
              M_prev[k][l] = m.Mpure[b][k][l];
              if (x >= xa) Fx_prev[k][l] = Fx0[k][l];
            }
          }
          cc_valid = false;
          A0 = A1;
          A1 = A2;
          continue;
        }
      }
      if (!cc_valid) cmu_all<N, NC>(m, A0.mu, cc_prev);
      cc_valid = true;
      /* phases present (nonzero on some lane) in rows x-1, x, x+1: wavefront-uniform */
      unsigned phmask = 0;
#pragma unroll
      for (int a = 0; a < N; a++)
        if (__any(A0.ph[a] != R(0.0) || A1.ph[a] != R(0.0) || A2.ph[a] != R(0.0))) phmask |= 1u << a;

      /* face gradients of row x used by the currents (calculate_gradients_phasefield_2D) */
      R gx0[N], gx1[N], gy1[N];
#pragma unroll
      for (int a = 0; a < N; a++) {
        gx0[a] = divx<Vw>(m, A1.ph[a] - A0.ph[a]);
        gx1[a] = divx<Vw>(m, A2.ph[a] - A1.ph[a]);
        gy1[a] = divy<Vw>(m, A1.phd[a] - A1.ph[a]);
      }
      R cc1[N][NC];
      cmu_all<N, NC>(m, A1.mu, cc1);
      /* delta-phi of row x: k_dphi wrote it for every candidate cell (a cell whose 5-point neighbourhood is not
       * purely one phase); every other cell has delta-phi exactly 0. Lanes 1..62 are the ones whose delta-phi
       * the output lanes read. */
      R dph[N];
#pragma unroll
      for (int a = 0; a < N; a++) dph[a] = R(0.0);
      {
        const int up = lane_up(A1.pure), dn = lane_dn(A1.pure);
        const bool cand = lane >= 1 && lane <= 62 && y <= my + 3 &&
                          (A1.pure < 0 || A0.pure != A1.pure || A2.pure != A1.pure || up != A1.pure || dn != A1.pure);
        if (cand) {
          const int64_t c = x * ps + y + kYOff;
#pragma unroll
          for (int a = 0; a < N; a++) dph[a] = dphi_in[a * fs + c];
        }
      }

      /* anti-trapping currents: right y-face of row x, x-face between x-1 and x */
      R jy1[NC], jx0[NC];
#pragma unroll
      for (int k = 0; k < NC; k++) jy1[k] = jx0[k] = R(0.0);
      bool jbusy = false;
#pragma unroll
      for (int a = 0; a < N - 1; a++) jbusy = jbusy || (dph[a] != R(0.0)) || (dph_prev[a] != R(0.0));
      R dph_dn[N];
#pragma unroll
      for (int a = 0; a < N; a++) dph_dn[a] = lane_dn(dph[a]);
#pragma unroll
      for (int a = 0; a < N - 1; a++) jbusy = jbusy || (dph_dn[a] != R(0.0));
      if (__any(jbusy)) {
        R cyA0[N], cyA1[N], cxA1[N], cxA1d[N], gyA1[N];
#pragma unroll
        for (int a = 0; a < N; a++) {
          cyA0[a] = div2y<Vw>(m, A0.phd[a] - A0.phu[a]);
          cyA1[a] = div2y<Vw>(m, A1.phd[a] - A1.phu[a]);
          cxA1[a] = R(0.5) * (gx1[a] + gx0[a]);
          cxA1d[a] = R(0.5) * (divx<Vw>(m, A2.phd[a] - A1.phd[a]) + divx<Vw>(m, A1.phd[a] - A0.phd[a]));
          gyA1[a] = gy1[a];
        }
        /* y-face (x, y+1/2): normal gradient gy, tangential average of cx over the two cells */
        if (y != 0 && y != m.jat_skip_row) {
          R cco[N][NC], gt[N];
          cmu_all<N, NC>(m, A1.mud, cco);
#pragma unroll
          for (int a = 0; a < N; a++) gt[a] = R(0.5) * (cxA1[a] + cxA1d[a]);
          jat_face<N, NC>(m, A1.ph, A1.phd, cc1, cco, dph, dph_dn, gyA1, gt, jy1, phmask);
        }
        /* x-face (x-1 .. x): normal gradient gx0, tangential average of cy over rows x-1, x */
        if (x >= xa) {
          R gt[N];
#pragma unroll
          for (int a = 0; a < N; a++) gt[a] = R(0.5) * (cyA0[a] + cyA1[a]);
          jat_face<N, NC>(m, A0.ph, A1.ph, cc_prev, cc1, dph_prev, dph, gx0, gt, jx0, phmask);
        }
      }

      /* mobility fluxes (calculate_gradients_concentration_2D) */
      R M1[NC][NC], Fx0[NC][NC];
      mobility_cell<N, NC>(m, A1.ph, M1);
      if (x >= xa) {
#pragma unroll
        for (int k = 0; k < NC; k++)
#pragma unroll
          for (int l = 0; l < NC; l++)
            Fx0[k][l] = ((M1[k][l] + M_prev[k][l]) * R(0.5)) * divx<Vw>(m, (A1.mu[l] - A0.mu[l]));
      }

      /* update of row x-1 (calculate_divergence_concentration_2D body) */
      if (x >= xa + 1) {
        R Fy[NC][NC], Fyl[NC][NC], jyl[NC];
#pragma unroll
        for (int k = 0; k < NC; k++) {
          jyl[k] = lane_up(jy_prev[k]);
#pragma unroll
          for (int l = 0; l < NC; l++) {
            const R Md = lane_dn(M_prev[k][l]);
            Fy[k][l] = ((Md + M_prev[k][l]) * R(0.5)) * divy<Vw>(m, (A0.mud[l] - A0.mu[l]));
          }
        }
#pragma unroll
        for (int k = 0; k < NC; k++)
#pragma unroll
          for (int l = 0; l < NC; l++) Fyl[k][l] = lane_up(Fy[k][l]);

        const R *phc = A0.ph, *muc = A0.mu;
        R sum[NC], divflux[NC], divjat[NC], deltamu[NC], munew[NC];
#pragma unroll
        for (int k = 0; k < NC; k++) {
          sum[k] = R(0.0);
          deltamu[k] = R(0.0);
          munew[k] = muc[k];
        }
#pragma unroll
        for (int a = 0; a < N; a++) {
          R sdh = R(0.0);
#pragma unroll
          for (int b = 0; b < N; b++) sdh += dhphi<N>(phc, a, b) * dph_prev[b];
#pragma unroll
          for (int k = 0; k < NC; k++) sum[k] += cc_prev[a][k] * sdh;
        }
#pragma unroll
        for (int k = 0; k < NC; k++) {
          divflux[k] = R(0.0);
#pragma unroll
          for (int l = 0; l < NC; l++) {
            divflux[k] += divx<Vw>(m, (Fx0[k][l] - Fx_prev[k][l]));
            divflux[k] += divy<Vw>(m, (Fy[k][l] - Fyl[k][l]));
          }
          divjat[k] = divx<Vw>(m, (jx0[k] - jx_prev[k]) * m.jat_fac);
          divjat[k] += divy<Vw>(m, (jy_prev[k] - jyl[k]) * m.jat_fac);
        }
        int bulk = -1;
#pragma unroll
        for (int a = 0; a < N; a++)
          if (bulk < 0 && phc[a] == R(1.0)) bulk = a;
        if (bulk >= 0 && kDirectBulkMu<R>) {
          R dc[NC];
#pragma unroll
          for (int k = 0; k < NC; k++) dc[k] = m.dt * (divflux[k]) - divjat[k];
#pragma unroll
          for (int k = 0; k < NC; k++) {
            deltamu[k] = Mu_lin<NC>(m, dc, bulk, k);
            munew[k] += deltamu[k];
          }
        } else if (bulk >= 0) {
          R c_new[NC];
#pragma unroll
          for (int k = 0; k < NC; k++) {
            R c_old = R(0);
#pragma unroll
            for (int a = 0; a < N; a++)
              if (a == bulk) c_old = cc_prev[a][k];
            c_new[k] = c_old + m.dt * (divflux[k]) - divjat[k];
          }
#pragma unroll
          for (int k = 0; k < NC; k++) {
            R nm = R(0);
#pragma unroll
            for (int a = 0; a < N; a++)
              if (a == bulk) nm = Mu<NC>(m, c_new, a, k);
            deltamu[k] = nm - munew[k];
            munew[k] = nm;
          }
        } else {
          R deltac[NC], dcdmu[NC][NC], h[N];
#pragma unroll
          for (int a = 0; a < N; a++) h[a] = hphi<N>(phc, a);
#pragma unroll
          for (int k = 0; k < NC; k++) {
            deltac[k] = m.dt * (divflux[k]) - divjat[k] - sum[k];
#pragma unroll
            for (int l = 0; l < NC; l++) {
              dcdmu[k][l] = R(0.0);
#pragma unroll
              for (int a = 0; a < N; a++) dcdmu[k][l] += m.cmu[a][k][l] * h[a];
            }
          }
          if (m.solve == GP_SOLVE_BINARY && NC == 1) {
            R inv = qdiv(R(1.0), dcdmu[0][0]);
            deltamu[0] = deltac[0] * inv;
            munew[0] += deltamu[0];
          } else if (m.solve == GP_SOLVE_TERNARY && NC == 2) {
            R DET = dcdmu[0][0] * dcdmu[1 % NC][1 % NC] - dcdmu[0][1 % NC] * dcdmu[1 % NC][0];
            R inv[NC][NC];
            inv[0][0] = qdiv(dcdmu[1 % NC][1 % NC], DET);
            inv[1 % NC][1 % NC] = qdiv(dcdmu[0][0], DET);
            inv[0][1 % NC] = qdiv(-dcdmu[0][1 % NC], DET);
            inv[1 % NC][0] = qdiv(-dcdmu[1 % NC][0], DET);
#pragma unroll
            for (int k = 0; k < NC; k++) {
              deltamu[k] = R(0.0);
#pragma unroll
              for (int l = 0; l < NC; l++) deltamu[k] += inv[k][l] * deltac[l];
            }
#pragma unroll
            for (int k = 0; k < NC; k++) munew[k] = deltamu[k] + munew[k];
          } else if (m.solve == GP_SOLVE_DILUTE) {
#pragma unroll
            for (int k = 0; k < NC; k++) {
              R inv = qdiv(R(1.0), dcdmu[k][k]);
              deltamu[k] = deltac[k] * inv;
              munew[k] += deltamu[k];
            }
          } else {
            R inv[NC][NC];
            lu_inverse<NC>(dcdmu, inv);
#pragma unroll
            for (int i = 0; i < NC; i++) {
              R s = R(0);
#pragma unroll
              for (int j = 0; j < NC; j++) s = s + inv[i][j] * deltac[j];
              deltamu[i] = s;
            }
#pragma unroll
            for (int k = 0; k < NC; k++) munew[k] = deltamu[k] + munew[k];
          }
        }
        R pn[N];
#pragma unroll
        for (int b = 0; b < N; b++) pn[b] = phc[b] + dph_prev[b];
        emit(x - 1, pn, munew, dph_prev, deltamu);
      }

      /* carry row x into the next step */
#pragma unroll
      for (int a = 0; a < N; a++) dph_prev[a] = dph[a];
#pragma unroll
      for (int a = 0; a < N; a++)
#pragma unroll
        for (int k = 0; k < NC; k++) cc_prev[a][k] = cc1[a][k];
#pragma unroll
      for (int k = 0; k < NC; k++) {
        jy_prev[k] = jy1[k];
        if (x >= xa) jx_prev[k] = jx0[k];
#pragma unroll
        for (int l = 0; l < NC; l++) {
          M_prev[k][l] = M1[k][l];
          if (x >= xa) Fx_prev[k][l] = Fx0[k][l];
        }
      }
      A0 = A1;
      A1 = A2;
    }
    /* tile flag: lane 2 is an output lane of every tile */
    const int ref = __shfl(oc, 2, 64);
    const bool same = __all(!out_lane || oc == ref);
    if (lane == 0) flags_out[w] = (same && ref >= 0) ? ref : -1;
  }

  /* statistics: wavefront reduction, one partial per wavefront (statistic-major); a deep tile's wavefront
   * writes the neutral values */
  if (!STATS) return;
  stats_partials<N, NC>(st, partials, pstride, w, nwaves, lane);
}



/* Bulk step of the deep tiles (deep_phase >= 0): every cell the update of such a tile reads is purely phase b, so
 * delta-phi and the anti-trapping currents vanish, phi is unchanged and mu follows the bulk update of phase b with
 * the mobility Mpure[b]. The values, operations and order are those of the pure-phase step of k_march (and so of
 * the general step), so the results are bit-identical. The kernel reads mu only and writes phi = e_b (phi + 0 of
 * the general step), mu and the tile flag. Its few registers keep many wavefronts resident, and the mu rows are
 * prefetched kBulkDepth march steps ahead: the work is pure streaming, bound by HBM latency and bandwidth. */
constexpr int kBulkDepth = 4;

template <int N, int NC, typename R, bool P2, bool STATS>
__global__ __launch_bounds__(256) void k_bulk(const Model<R> *__restrict__ mp, const R *__restrict__ mu,
                                              R *__restrict__ phi_o, R *__restrict__ mu_o,
                                              int8_t *__restrict__ codes_out, double *__restrict__ partials,
                                              int64_t pstride,
                                              const int *__restrict__ flags_in, int *__restrict__ flags_out,
                                              int64_t fs, int64_t ps, int64_t mx, int seg_len, int nstrips, int nsegs,
                                              int nwaves) {
  using Vw = View<R, P2, false, kFastOf<R>>;
  constexpr int NS = 3 * N + 3 * NC + 1;
  constexpr int NSA = STATS ? NS : 1;
  const Model<R> &m = *mp;
  const int lane = threadIdx.x & 63;
  const int w = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  R st[NSA];
  if (STATS) stats_init<N, NC>(st);
  const int strip = w % nstrips, seg = w / nstrips;
  const int b = w < nwaves ? deep_phase(flags_in, strip, seg, nstrips, nsegs) : -1;
  if (b >= 0) {
    /* a deep tile is not on the edge of the tile grid: all its lanes are inside the grid, no halo images */
    const int64_t y = 3 + (int64_t)strip * kMarchOut - 2 + lane;
    const int64_t xa = 3 + (int64_t)seg * seg_len;
    const int64_t xb = (xa + seg_len - 1 < mx + 2) ? xa + seg_len - 1 : mx + 2;
    const bool out_lane = lane >= 2 && lane <= 61;
    R Mb[NC][NC];
#pragma unroll
    for (int k = 0; k < NC; k++)
#pragma unroll
      for (int l = 0; l < NC; l++) Mb[k][l] = m.Mpure[b][k][l];
    const R zj = R(0.0) * m.jat_fac; /* (j - j_prev) * jat_fac with both currents +0 */
    R divjat0 = divx<Vw>(m, zj);
    divjat0 += divy<Vw>(m, zj);
    R pe[N];
#pragma unroll
    for (int a = 0; a < N; a++) pe[a] = a == b ? R(1.0) : R(0.0);

    R m0[NC], m1[NC], q[kBulkDepth][NC];
    const int64_t c0 = y + kYOff;
#pragma unroll
    for (int k = 0; k < NC; k++) {
      m0[k] = mu[k * fs + (xa - 1) * ps + c0];
      m1[k] = mu[k * fs + xa * ps + c0];
    }
#pragma unroll
    for (int d = 0; d < kBulkDepth; d++)
      if (xa + 1 + d <= xb + 1) {
#pragma unroll
        for (int k = 0; k < NC; k++) q[d][k] = mu[k * fs + (xa + 1 + d) * ps + c0];
      }
    R Fb[NC][NC];
#pragma unroll
    for (int k = 0; k < NC; k++)
#pragma unroll
      for (int l = 0; l < NC; l++) Fb[k][l] = Mb[k][l] * divx<Vw>(m, (m1[l] - m0[l]));

    for (int64_t x = xa; x <= xb; x++) {
      R m2[NC];
#pragma unroll
      for (int k = 0; k < NC; k++) m2[k] = q[0][k];
#pragma unroll
      for (int d = 0; d + 1 < kBulkDepth; d++)
#pragma unroll
        for (int k = 0; k < NC; k++) q[d][k] = q[d + 1][k];
      if (x + 1 + kBulkDepth <= xb + 1) {
#pragma unroll
        for (int k = 0; k < NC; k++) q[kBulkDepth - 1][k] = mu[k * fs + (x + 1 + kBulkDepth) * ps + c0];
      }
      R Ff[NC][NC], Fy[NC][NC], Fyl[NC][NC], mud[NC], muu[NC];
#pragma unroll
      for (int k = 0; k < NC; k++) {
        mud[k] = lane_dn(m1[k]);
        muu[k] = lane_up(m1[k]);
      }
#pragma unroll
      for (int k = 0; k < NC; k++)
#pragma unroll
        for (int l = 0; l < NC; l++) {
          Ff[k][l] = Mb[k][l] * divx<Vw>(m, (m2[l] - m1[l]));
          Fy[k][l] = Mb[k][l] * divy<Vw>(m, (mud[l] - m1[l]));
          Fyl[k][l] = Mb[k][l] * divy<Vw>(m, (m1[l] - muu[l])); /* = Fy of the lane below, bit for bit */
        }
      R divflux[NC], deltamu[NC], munew[NC];
#pragma unroll
      for (int k = 0; k < NC; k++) {
        divflux[k] = R(0.0);
#pragma unroll
        for (int l = 0; l < NC; l++) {
          divflux[k] += divx<Vw>(m, (Ff[k][l] - Fb[k][l]));
          divflux[k] += divy<Vw>(m, (Fy[k][l] - Fyl[k][l]));
        }
      }
      if (kDirectBulkMu<R>) {
        R dc[NC];
#pragma unroll
        for (int k = 0; k < NC; k++) dc[k] = m.dt * (divflux[k]) - divjat0;
#pragma unroll
        for (int k = 0; k < NC; k++) {
          deltamu[k] = Mu_lin<NC>(m, dc, b, k);
          munew[k] = m1[k] + deltamu[k];
        }
      } else {
        R c_new[NC];
#pragma unroll
        for (int k = 0; k < NC; k++) c_new[k] = c_mu<NC>(m, m1, b, k) + m.dt * (divflux[k]) - divjat0;
#pragma unroll
        for (int k = 0; k < NC; k++) {
          munew[k] = Mu<NC>(m, c_new, b, k);
          deltamu[k] = munew[k] - m1[k];
        }
      }
      if (out_lane) {
        const int64_t c = x * ps + c0;
#pragma unroll
        for (int a = 0; a < N; a++) phi_o[a * fs + c] = pe[a];
#pragma unroll
        for (int k = 0; k < NC; k++) mu_o[k * fs + c] = munew[k];
        codes_out[c] = (int8_t)b;
        if (STATS) {
          const R z[N] = {};
          stats_add<N, NC>(st, pe, z, munew, deltamu);
        }
      }
#pragma unroll
      for (int k = 0; k < NC; k++) {
        m0[k] = m1[k];
        m1[k] = m2[k];
#pragma unroll
        for (int l = 0; l < NC; l++) Fb[k][l] = Ff[k][l];
      }
    }
    if (lane == 0) flags_out[w] = b; /* phi is unchanged: the tile stays purely phase b */
  }
  if (!STATS) return;
  stats_partials<N, NC>(st, partials, pstride, w, nwaves, lane);
}


/* Phase codes of a state: codes[cell] = pure_phase of its phi (b when purely phase b, -1 otherwise), over the
 * whole allocated grid (halo images included); one thread per cell. The step kernels keep the codes of the state
 * they write; this kernel forms them after an upload or a step of the two-pass path. */
template <int N, typename R>
__global__ __launch_bounds__(256) void k_codes(const R *__restrict__ phi, int8_t *__restrict__ codes, int64_t fs,
                                               int64_t ps, int64_t rx, int64_t ry) {
  const int64_t y = (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
  const int64_t x = (int64_t)blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x >= rx || y >= ry) return;
  const int64_t c = x * ps + y + kYOff;
  R ph[N];
#pragma unroll
  for (int a = 0; a < N; a++) ph[a] = phi[a * fs + c];
  codes[c] = (int8_t)pure_phase<N>(ph);
}

/* Delta-phi of the candidate cells. A cell whose 5-point neighbourhood is purely one phase has every Laplacian
 * zero and phi in {0, 1}, so its delta-phi is exactly 0 (no active phase, nothing for the projection to do);
 * every other cell of x, y in [2, mesh+3] is a candidate. One wavefront per tile of 64 y-cells by kDTileRows
 * x-rows: it finds the candidates from the phase codes, compacts them in LDS (ballot and prefix count, in row
 * order), and evaluates phi_delta (the per-cell form of calculate_divergence_phasefield_2D and the simplex
 * projection) for 64 candidates at a time, so every lane works on an interface cell. */
constexpr int kDTileRows = 16;

/* phi on the 3x3 stencil and mu at the centre gathered once into registers, then phi_delta */
template <int N, int NC, typename R, bool P2>
__device__ __forceinline__ void dphi_cell(const Model<R> &m, const R *__restrict__ phi, const R *__restrict__ mu,
                                          R *__restrict__ dphi, int64_t fs, int64_t ps, int64_t c) {
  RegStore<N, NC, R> rs;
#pragma unroll
  for (int a = 0; a < N; a++)
#pragma unroll
    for (int di = 0; di < 3; di++)
#pragma unroll
      for (int dj = 0; dj < 3; dj++) rs.p[a][di][dj] = phi[a * fs + c + (di - 1) * ps + (dj - 1)];
#pragma unroll
  for (int k = 0; k < NC; k++) rs.u[k][0] = mu[k * fs + c];
  const RegView<N, NC, R, P2, kFastOf<R>> v{&rs, 0, 0};
  R d[N];
  phi_delta<N, NC>(m, v, false, d);
#pragma unroll
  for (int a = 0; a < N; a++) dphi[a * fs + c] = d[a];
}

template <int N, int NC, typename R, bool P2>
__global__ __launch_bounds__(256) void k_dphi(const Model<R> *__restrict__ mp, const R *__restrict__ phi,
                                              const R *__restrict__ mu, const int8_t *__restrict__ codes,
                                              R *__restrict__ dphi, int64_t fs, int64_t ps, int64_t mx, int64_t my,
                                              int nty, int ntiles) {
  __shared__ uint16_t list[4][kDTileRows * 64];
  /* the tile's codes, rows x0-1 .. x1+1, columns y-1 .. y+64 (slot j+1 holds y0+j), staged with one batch of
   * loads so that the candidate scan does not wait on memory row by row */
  __shared__ int8_t cs_all[4][(kDTileRows + 2) * 66];
  const Model<R> &m = *mp;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int w = blockIdx.x * 4 + wv;
  int8_t *cs = cs_all[wv];
  int count = 0;
  const int ty = w % nty, tx = w / nty;
  const int64_t y = 2 + (int64_t)ty * 64 + lane;
  const int64_t x0 = 2 + (int64_t)tx * kDTileRows;
  const int64_t x1 = (x0 + kDTileRows - 1 < mx + 3) ? x0 + kDTileRows - 1 : mx + 3;
  if (w < ntiles) {
    const int64_t ry = my + 6;
    const int64_t yc = y < ry - 1 ? y : ry - 1;
    const int nr = (int)(x1 - x0) + 3;
    for (int r = 0; r < nr; r++) {
      const int64_t row = (x0 - 1 + r) * ps + kYOff;
      cs[r * 66 + lane + 1] = codes[row + yc];
      if (lane == 0) cs[r * 66] = codes[row + y - 1];
      if (lane == 63) cs[r * 66 + 65] = codes[row + (y + 1 < ry - 1 ? y + 1 : ry - 1)];
    }
  }
  __syncthreads();
  if (w < ntiles) {
    const bool ylive = y <= my + 3;
    for (int64_t x = x0; x <= x1; x++) {
      bool cand = false;
      if (ylive) {
        const int r = (int)(x - x0) + 1, j = lane + 1;
        const int8_t k0 = cs[r * 66 + j];
        cand = k0 < 0 || cs[(r - 1) * 66 + j] != k0 || cs[(r + 1) * 66 + j] != k0 || cs[r * 66 + j - 1] != k0 ||
               cs[r * 66 + j + 1] != k0;
      }
      const uint64_t mask = __ballot(cand);
      const int pre = __popcll(mask & ((1ull << lane) - 1ull));
      if (cand) list[wv][count + pre] = (uint16_t)((x - x0) * 64 + lane);
      count += __popcll(mask);
    }
  }
  __syncthreads();
  for (int i = lane; i < count; i += 64) {
    const int r = list[wv][i];
    const int64_t c = (x0 + r / 64) * ps + (2 + (int64_t)ty * 64 + r % 64) + kYOff;
    dphi_cell<N, NC, R, P2>(m, phi, mu, dphi, fs, ps, c);
  }
}

/* delta-phi of the listed candidate cells: a grid-stride loop of full wavefronts */
template <int N, int NC, typename R, bool P2>
__global__ __launch_bounds__(256) void k_dlist(const Model<R> *__restrict__ mp, const R *__restrict__ phi,
                                               const R *__restrict__ mu, R *__restrict__ dphi, int64_t fs, int64_t ps,
                                               const int32_t *__restrict__ clist, const int *__restrict__ ccount) {
  const int n = *ccount;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    dphi_cell<N, NC, R, P2>(*mp, phi, mu, dphi, fs, ps, clist[i]);
}
}  // namespace gpk
