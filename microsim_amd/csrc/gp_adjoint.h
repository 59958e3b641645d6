/*
 * gp_adjoint.h -- adjoint (reverse-mode gradient) of the FP64 solver step, with checkpoint and recompute.
 *
 * Included by the kernel translation units after the handle and its helpers; not a public header.
 *
 * One forward step maps the state s_t (phi, mu, halo-inclusive) to s_{t+1} in three stages, each a per-cell
 * function of a fixed stencil:
 *   P  delta-phi at every cell of [2, mesh+3]^2 from phi on its 3x3 stencil and mu at the cell
 *      (phi_delta: calculate_divergence_phasefield_2D + simplex projection),
 *   Q  new phi and mu of every interior cell from phi on its 3x3 stencil, mu and delta-phi on its 5-point
 *      stencil (jat_x / jat_y / mu_update: calculate_fluxes_concentration_2D and
 *      calculate_divergence_concentration_2D),
 *   B  the boundary copies (each halo cell takes the value of one interior cell).
 * The adjoint step runs them backwards, transposed. A cell of stage P or Q is transposed in one pass by the
 * reverse-mode per-cell functions of gp_reverse.h (its forward values recomputed, then the incoming adjoint carried
 * back through every operation), and the result is scatter-added (FP64 atomics) onto the stencil's inputs and onto
 * the parameter gradient. Bulk cells of stage Q are transposed in closed form by a gather (k_adj_bulk). Cells whose
 * incoming adjoint is zero are skipped. B's transpose adds each halo cell's adjoint to its source cell.
 */
#pragma once

#include <vector>

#include "gp_reverse.h"

namespace gpk {

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

/* ---------------------------------------------------------------- stage Q of bulk cells, in closed form
 *
 * A "bulk" cell c of stage Q has phi_b(c) exactly 1 for some phase (mu_update takes its bulk branch, with the
 * first such b) and no solid-phase delta-phi on itself or its four neighbours (each anti-trapping current of
 * its faces returns +0 before any arithmetic, so carries no derivative). Its stage Q is then
 *   phi_new = phi + dphi,   mu_new = Mu_b(c_b(mu) + dt divflux),
 *   divflux_k = sum over faces f (neighbour e, spacing d_f) of sum_l M_f[k][l] (mu_l(e) - mu_l(c)) / d_f^2,
 *   M_f = (M(c) + M(e)) / 2,   M(x)[k][l] = sum_a sum_q DC[a][k][l][q] phi_a(x),
 * and its transpose is formed here in closed form (exact derivatives, rounded differently from the dual
 * evaluation). With cbar_k(c) = sum_j gmu'_j dMu_j/dc_k and Fbar(c) = dt cbar(c) (zero for other cells), the
 * adjoint of a cell n gathers, over its faces f with neighbour e:
 *   gmu_l(n)  += sum_k (Fbar_k(e) - Fbar_k(n)) M_f[k][l] / d_f^2
 *   gphi_a(n) += sum_k (Fbar_k(n) - Fbar_k(e)) / 2 sum_l DCs[a][k][l] (mu_l(e) - mu_l(n)) / d_f^2
 * plus, when n itself is bulk, gmu_l(n) += sum_k cbar_k(n) cmu_b[k][l], gphi(n) += gphi'(n), gdphi(n) = gphi'(n),
 * and the liquid diffusivities' gradient sum_k Fbar_k(n) sum_f (phi_L(n) + phi_L(e)) / 2 sum_l cmu_L[k][l]
 * (mu_l(e) - mu_l(n)) / d_f^2 (DCs[a][k][l] = sum_q DC[a][k][l][q] = dM/dphi_a). */
/* Append the cells of this wavefront with `flag` set to list (count: entries so far): one atomic per wavefront,
 * issued by the first flagged lane. */
__device__ __forceinline__ void list_append(bool flag, int32_t c, int32_t *list, int *count) {
  const unsigned long long mask = __ballot(flag);
  if (!mask) return;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((long long)mask) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(count, __popcll(mask));
  base = __shfl(base, leader);
  if (flag) list[base + __popcll(mask & ((1ull << lane) - 1ull))] = c;
}

template <int N, int NC>
__device__ __forceinline__ int bulk_q(const double *phi, const double *dphi, int64_t c, int64_t fs, int64_t ps) {
  int b = -1;
  for (int a = 0; a < N; a++)
    if (b < 0 && phi[a * fs + c] == 1.0) b = a;
  if (b < 0) return -1;
  for (int a = 0; a < N - 1; a++) {
    const double *d = dphi + a * fs + c;
    if (d[0] != 0.0 || d[ps] != 0.0 || d[-ps] != 0.0 || d[1] != 0.0 || d[-1] != 0.0) return -1;
  }
  return b;
}

/* bulk_q from the byte planes of the evaluation state: the pure-phase code (a cell that is not pure falls back to
 * phi) and the candidate flags (delta-phi is read only at candidates; it is +0 at every other cell). Same result. */
template <int N, int NC>
__device__ __forceinline__ int bulk_q_codes(const double *phi, const double *dphi, const int8_t *pcode,
                                            const int8_t *cand, int64_t c, int64_t fs, int64_t ps) {
  int b = pcode[c];
  if (b < 0) {
    for (int a = 0; a < N; a++)
      if (b < 0 && phi[a * fs + c] == 1.0) b = a;
    if (b < 0) return -1;
  }
  const int64_t o[5] = {0, ps, -ps, 1, -1};
  for (int i = 0; i < 5; i++)
    if (cand[c + o[i]])
      for (int a = 0; a < N - 1; a++)
        if (dphi[a * fs + c + o[i]] != 0.0) return -1;
  return b;
}

/* cbar of every cell of the allocated grid (zero outside the interior and for cells that are not bulk), and the
 * list of the interior cells that are not bulk and have a nonzero incoming adjoint (stage Q by reverse mode) */
template <int N, int NC>
__global__ __launch_bounds__(256) void k_adj_cbar(const Model<double> *__restrict__ mp, const double *__restrict__ phi,
                                                  const double *__restrict__ dphi, const double *__restrict__ gphi_n,
                                                  const double *__restrict__ gmu_n, double *__restrict__ cbar,
                                                  int8_t *__restrict__ bq, const int8_t *__restrict__ pcode,
                                                  const int8_t *__restrict__ cand, int32_t *__restrict__ qlist,
                                                  int *qcount, int64_t fs, int64_t ps, int64_t mx, int64_t my) {
  const int64_t y = (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
  const int64_t x = (int64_t)blockIdx.y * 4 + (threadIdx.x >> 6);
  const bool valid = x < mx + 6 && y < my + 6;
  const int64_t c = x * ps + y + kYOff;
  const bool inner = valid && x >= 3 && x <= mx + 2 && y >= 3 && y <= my + 2;
  const int b = inner ? bulk_q_codes<N, NC>(phi, dphi, pcode, cand, c, fs, ps) : -1;
  bool dual = false;
  if (valid) {
    bq[c] = (int8_t)b;
    const Model<double> &m = *mp;
    for (int k = 0; k < NC; k++) {
      double s = 0.0;
      if (b >= 0)
        for (int j = 0; j < NC; j++) s += gmu_n[j * fs + c] * (j == k ? 2.0 * m.A[b][j][j] : m.A[b][j][k]);
      cbar[k * fs + c] = s;
    }
    if (inner && b < 0) {
      for (int a = 0; a < N; a++) dual = dual || gphi_n[a * fs + c] != 0.0;
      for (int k = 0; k < NC; k++) dual = dual || gmu_n[k * fs + c] != 0.0;
    }
  }
  list_append(dual, (int32_t)c, qlist, qcount);
}

/* the gather above for every cell n of [2, mesh+3]^2: plain stores into gphi / gmu and into gdphi (its value for a
 * bulk n, else 0); every other allocated cell gets zeros in all three, so no clearing pass is needed. The
 * reverse-mode kernels then add the other cells' contributions. */
template <int N, int NC>
__global__ __launch_bounds__(256) void k_adj_bulk(const Model<double> *__restrict__ mp, const double *__restrict__ phi,
                                                  const double *__restrict__ mu, const int8_t *__restrict__ bq,
                                                  const int8_t *__restrict__ pcode, const double *__restrict__ cbar,
                                                  const double *__restrict__ gphi_n, double *__restrict__ gphi,
                                                  double *__restrict__ gmu, double *__restrict__ gdphi, double *gtheta,
                                                  int th0, int64_t fs, int64_t ps, int64_t mx, int64_t my) {
  const Model<double> &m = *mp;
  const int64_t y = (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
  const int64_t x = (int64_t)blockIdx.y * 4 + (threadIdx.x >> 6);
  constexpr int L = N - 1;
  double th[NC];
  for (int k = 0; k < NC; k++) th[k] = 0.0;
  const bool alloc = x < mx + 6 && y < my + 6;
  if (alloc && !(x >= 2 && x <= mx + 3 && y >= 2 && y <= my + 3)) {
    const int64_t c = x * ps + y + kYOff;
    for (int a = 0; a < N; a++) {
      gphi[a * fs + c] = 0.0;
      gdphi[a * fs + c] = 0.0;
    }
    for (int k = 0; k < NC; k++) gmu[k * fs + c] = 0.0;
  } else if (alloc) {
    const int64_t c = x * ps + y + kYOff;
    const int b = bq[c]; /* -1 outside the inner cells */
    const int pc = pcode[c];
    double gph[N], gm[NC], cb[NC], mun[NC], phn[N], Mn[NC][NC];
    for (int k = 0; k < NC; k++) {
      cb[k] = cbar[k * fs + c];
      mun[k] = mu[k * fs + c];
      gm[k] = 0.0;
    }
    for (int a = 0; a < N; a++) {
      /* a pure cell's phi is e_pc exactly */
      phn[a] = pc >= 0 ? (a == pc ? 1.0 : 0.0) : phi[a * fs + c];
      gph[a] = 0.0;
    }
    if (b >= 0) {
      for (int a = 0; a < N; a++) {
        gph[a] = gphi_n[a * fs + c];
        gdphi[a * fs + c] = gph[a];
      }
      for (int l = 0; l < NC; l++)
        for (int k = 0; k < NC; k++) gm[l] += cb[k] * m.cmu[b][k][l];
    } else {
      for (int a = 0; a < N; a++) gdphi[a * fs + c] = 0.0;
    }
    /* M(n); a pure cell's is Mpure, formed with the same operations (bit-identical) */
    for (int k = 0; k < NC; k++)
      for (int l = 0; l < NC; l++) {
        double s = 0.0;
        if (pc >= 0) {
          s = m.Mpure[pc][k][l];
        } else {
          for (int a = 0; a < N; a++)
            for (int q = 0; q < NC; q++) s += m.DC[a][k][l][q] * phn[a];
        }
        Mn[k][l] = s;
      }
    const double wx = m.inv_dx * m.inv_dx, wy = m.inv_dy * m.inv_dy;
    const int64_t offs[4] = {ps, -ps, 1, -1};
    for (int f = 0; f < 4; f++) {
      const int64_t e = c + offs[f];
      const double w = f < 2 ? wx : wy;
      double dF[NC], due[NC];
      bool nz = false;
      for (int k = 0; k < NC; k++) {
        dF[k] = m.dt * (cbar[k * fs + e] - cb[k]);
        nz = nz || dF[k] != 0.0;
      }
      if (!nz) continue;
      const int pe = pcode[e];
      double phe[N];
      for (int a = 0; a < N; a++) phe[a] = pe >= 0 ? (a == pe ? 1.0 : 0.0) : phi[a * fs + e];
      for (int l = 0; l < NC; l++) due[l] = mu[l * fs + e] - mun[l];
      for (int l = 0; l < NC; l++) {
        double s = 0.0;
        for (int k = 0; k < NC; k++) {
          double Me = 0.0;
          if (pe >= 0) {
            Me = m.Mpure[pe][k][l];
          } else {
            for (int a = 0; a < N; a++)
              for (int q = 0; q < NC; q++) Me += m.DC[a][k][l][q] * phe[a];
          }
          s += dF[k] * 0.5 * (Mn[k][l] + Me);
        }
        gm[l] += w * s;
      }
      for (int a = 0; a < N; a++) {
        double s = 0.0;
        for (int k = 0; k < NC; k++) {
          double t = 0.0;
          for (int l = 0; l < NC; l++) {
            double dcs = 0.0;
            for (int q = 0; q < NC; q++) dcs += m.DC[a][k][l][q];
            t += dcs * due[l];
          }
          s -= dF[k] * t;
        }
        gph[a] += 0.5 * w * s;
      }
      if (b >= 0)
        for (int k = 0; k < NC; k++) {
          double t = 0.0;
          for (int l = 0; l < NC; l++) t += m.cmu[L][k][l] * due[l];
          th[k] += m.dt * cb[k] * 0.5 * w * (phn[L] + phe[L]) * t;
        }
    }
    for (int a = 0; a < N; a++) gphi[a * fs + c] = gph[a];
    for (int k = 0; k < NC; k++) gmu[k * fs + c] = gm[k];
  }
  __shared__ double red[4][NC];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int k = 0; k < NC; k++) {
    const double r = wave_sum(th[k]);
    if (lane == 0) red[wv][k] = r;
  }
  __syncthreads();
  if (threadIdx.x == 0)
    for (int k = 0; k < NC; k++) {
      const double s = red[0][k] + red[1][k] + red[2][k] + red[3][k];
      if (s != 0.0) atomicAdd(&gtheta[th0 + k], s);
    }
}

/* scatter one cell's stencil adjoint (RevAcc) onto the state adjoints with atomics (zero entries skipped) */
template <int N, int NC>
__device__ __forceinline__ void scatter_acc(const RevAcc<N, NC> &acc, int64_t c, int64_t fs, int64_t ps,
                                            double *gphi, double *gmu, double *gdphi) {
  const int oi[5] = {0, 1, -1, 0, 0}, oj[5] = {0, 0, 0, 1, -1};
#pragma unroll
  for (int a = 0; a < N; a++)
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++)
        if (acc.p[a][i][j] != 0.0) atomicAdd(&gphi[a * fs + c + (i - 1) * ps + (j - 1)], acc.p[a][i][j]);
#pragma unroll
  for (int o = 0; o < 5; o++) {
    const int64_t co = c + oi[o] * ps + oj[o];
#pragma unroll
    for (int k = 0; k < NC; k++)
      if (acc.u[k][o] != 0.0) atomicAdd(&gmu[k * fs + co], acc.u[k][o]);
    if (gdphi)
#pragma unroll
      for (int a = 0; a < N; a++)
        if (acc.d[a][o] != 0.0) atomicAdd(&gdphi[a * fs + co], acc.d[a][o]);
  }
}

/* the stencil of cell c: phi on the 3x3 neighbourhood, mu and (when dphi is given) delta-phi on the 5-point one */
template <int N, int NC>
__device__ __forceinline__ void gather_stencil(const double *__restrict__ phi, const double *__restrict__ mu,
                                               const double *__restrict__ dphi, int64_t c, int64_t fs, int64_t ps,
                                               RegStore<N, NC, double> &rs) {
  const int oi[5] = {0, 1, -1, 0, 0}, oj[5] = {0, 0, 0, 1, -1};
#pragma unroll
  for (int a = 0; a < N; a++)
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) rs.p[a][i][j] = phi[a * fs + c + (i - 1) * ps + (j - 1)];
#pragma unroll
  for (int o = 0; o < 5; o++) {
    const int64_t co = c + oi[o] * ps + oj[o];
#pragma unroll
    for (int k = 0; k < NC; k++) rs.u[k][o] = mu[k * fs + co];
#pragma unroll
    for (int a = 0; a < N; a++) rs.d[a][o] = dphi ? dphi[a * fs + co] : 0.0;
  }
}

/* Stage Q transposed by reverse mode (gp_reverse.h: cell_update_vjp), one pass per listed cell (interior, not bulk,
 * nonzero incoming adjoint; k_adj_bulk covers the bulk cells): a grid-stride loop over the list. */
template <int N, int NC, bool P2>
__global__ __launch_bounds__(256) void k_adj_qrev(const Model<double> *__restrict__ mp, const double *__restrict__ phi,
                                                  const double *__restrict__ mu, const double *__restrict__ dphi,
                                                  const double *__restrict__ gphi_n, const double *__restrict__ gmu_n,
                                                  double *gphi, double *gmu, double *gdphi, double *gtheta,
                                                  const int32_t *__restrict__ list, const int *__restrict__ count,
                                                  int64_t fs, int64_t ps) {
  constexpr int NTH = kThetaN<N, NC>;
  double th[NTH];
  for (int t = 0; t < NTH; t++) th[t] = 0.0;
  const Model<double> &m = *mp;
  const int n = *count;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int64_t c = list[i];
    const int64_t y = (c - kYOff) % ps;
    double go[N], gu[NC];
    for (int a = 0; a < N; a++) go[a] = gphi_n[a * fs + c];
    for (int k = 0; k < NC; k++) gu[k] = gmu_n[k * fs + c];
    RegStore<N, NC, double> rs;
    gather_stencil<N, NC>(phi, mu, dphi, c, fs, ps, rs);
    RevAcc<N, NC> acc;
    acc.clear();
    cell_update_vjp<N, NC>(m, RegView<N, NC, double, P2>{&rs, 0, 0}, y, go, gu, acc);
    scatter_acc<N, NC>(acc, c, fs, ps, gphi, gmu, gdphi);
    for (int t = 0; t < NTH; t++) th[t] += acc.th[t];
  }
  theta_flush<NTH>(th, NTH, gtheta);
}

/* Stage P transposed runs on the cells of [2, mesh+3]^2 with a nonzero delta-phi adjoint whose 5-point
 * neighbourhood is not purely one phase. (Where it is, every Laplacian of phi is exactly 0 and no phi is
 * negative: phi_delta returns zeros untouched by the projection, with zero derivative.) */
template <int N>
__device__ __forceinline__ bool p_candidate(const double *phi, int64_t c, int64_t fs, int64_t ps) {
  const int64_t o[5] = {0, ps, -ps, 1, -1};
  int b0 = -2;
  for (int i = 0; i < 5; i++) {
    int code = -1, nz = 0;
    for (int a = 0; a < N; a++) {
      const double v = phi[a * fs + c + o[i]];
      if (v == 1.0) code = a;
      nz += (v != 0.0);
    }
    if (nz != 1 || code < 0) return true;
    if (b0 == -2) b0 = code;
    else if (code != b0) return true;
  }
  return false;
}

/* Stage B transposed at cell (x, y) for field group f (the ncomp planes at g): a halo cell's adjoint goes to the
 * interior cell it copies (x maps of the field for halo rows, y maps for halo columns, both for corners), and the
 * halo adjoint is cleared. */
__device__ __forceinline__ void adj_bc_cell(double *g, int ncomp, const Maps &maps, int f, int64_t x, int64_t y,
                                            int64_t fs, int64_t ps) {
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

/* the pure-phase code of every allocated cell (pcode) and the list of the P candidates of [2, mesh+3]^2: a cell whose
 * five-point phi is not one pure phase throughout. Every other cell's delta-phi is +0 (no divergence, no clamp):
 * this pass stores +0 at every cell of [2, mesh+3]^2, and k_dlist then forms the candidates' delta-phi
 * (phi_delta, FP64, as the step does) over the list in full wavefronts. */
template <int N, int NC>
__global__ __launch_bounds__(256) void k_adj_codes(const double *__restrict__ phi, double *__restrict__ dphi,
                                                   int8_t *__restrict__ pcode, int8_t *__restrict__ cflag,
                                                   int32_t *__restrict__ plist, int *pcount, int *pclear,
                                                   double *gfold, Maps maps, int *qclear, int64_t fs, int64_t ps,
                                                   int64_t mx, int64_t my) {
  /* the other P slot (read by the previous transposed step, which ran before this pass) starts the next one at 0,
   * and so does the stage-Q list that the coming transposed step fills */
  if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
    *pclear = 0;
    *qclear = 0;
  }
  const int64_t y = (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
  const int64_t x = (int64_t)blockIdx.y * 4 + (threadIdx.x >> 6);
  const int64_t c = x * ps + y + kYOff;
  bool cand = false;
  if (x < mx + 6 && y < my + 6) {
    int code = -1, nz = 0;
    for (int a = 0; a < N; a++) {
      const double v = phi[a * fs + c];
      if (v == 1.0) code = a;
      nz += (v != 0.0);
    }
    pcode[c] = (int8_t)(nz == 1 ? code : -1);
  }
  if (x >= 2 && x <= mx + 3 && y >= 2 && y <= my + 3) {
    cand = p_candidate<N>(phi, c, fs, ps);
    for (int a = 0; a < N; a++) dphi[a * fs + c] = 0.0;
  }
  if (x < mx + 6 && y < my + 6) {
    cflag[c] = cand ? 1 : 0;
    /* stage B of the coming transposed step (its adjoint input gfold), which nothing here reads */
    adj_bc_cell(gfold, N, maps, 0, x, y, fs, ps);
    adj_bc_cell(gfold + (int64_t)N * fs, NC, maps, 1, x, y, fs, ps);
  }
  list_append(cand, (int32_t)c, plist, pcount);
}

/* Stage P transposed by reverse mode (gp_reverse.h: phi_delta_vjp), one pass per listed candidate (k_adj_codes'
 * list) with a nonzero delta-phi adjoint */
template <int N, int NC, bool P2>
__global__ __launch_bounds__(256) void k_adj_prev(const Model<double> *__restrict__ mp, const double *__restrict__ phi,
                                                  const double *__restrict__ mu, const double *__restrict__ gdphi,
                                                  double *gphi, double *gmu, double *gtheta,
                                                  const int32_t *__restrict__ list, const int *__restrict__ count,
                                                  int64_t fs, int64_t ps) {
  constexpr int NTH = kThetaN<N, NC>;
  double th[NTH];
  for (int t = 0; t < NTH; t++) th[t] = 0.0;
  const Model<double> &m = *mp;
  const int n = *count;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int64_t c = list[i];
    double go[N];
    bool nz = false;
    for (int a = 0; a < N; a++) {
      go[a] = gdphi[a * fs + c];
      nz = nz || go[a] != 0.0;
    }
    if (!nz) continue; /* a candidate with no incoming delta-phi adjoint */
    RegStore<N, NC, double> rs;
    gather_stencil<N, NC>(phi, mu, nullptr, c, fs, ps, rs);
    RevAcc<N, NC> acc;
    acc.clear();
    phi_delta_vjp<N, NC>(m, RegView<N, NC, double, P2>{&rs, 0, 0}, go, acc);
    scatter_acc<N, NC>(acc, c, fs, ps, gphi, gmu, nullptr);
    for (int t = 0; t < NTH; t++) th[t] += acc.th[t];
  }
  theta_flush<NTH>(th, NTH, gtheta);
}

/* Stage B transposed over the allocated grid; blockIdx.z is the field group (0: the N phi planes at g, 1: the NC mu
 * planes after them) */
static __global__ __launch_bounds__(256) void k_adj_bc(double *g, int N, int NC, Maps maps, int64_t fs, int64_t ps,
                                                int64_t rx, int64_t ry) {
  const int f = (int)blockIdx.z;
  const int64_t y = (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
  const int64_t x = (int64_t)blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x >= rx || y >= ry) return;
  adj_bc_cell(f ? g + (int64_t)N * fs : g, f ? NC : N, maps, f, x, y, fs, ps);
}

}  // namespace gpk
