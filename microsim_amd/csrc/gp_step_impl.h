/*
 * gp_step_impl.h -- the kernels of the step, the invariants and the transposed step, and their launch sequences,
 * as templates over (N, NC). Included by gp_ops.hip, which instantiates them for one pair; not a public header.
 */
#pragma once

#include "gp_internal.h"
#include "gp_cells.h"
#include "gp_march.h"
#include "gp_adjoint.h"
#include "gp_c32.h"
#include "gp_frame.h"

namespace gpk {

/* ------------------------------------------------------------------------------------------------ kernels */

/* TG: non-isothermal, th holds the coefficients of every allocated row (gp_frame.h) */
/* Grid of a list-driven kernel (a grid-stride loop over a list, 256-thread blocks): the blocks the device holds at
 * once, its occupancy times the compute units, at most kListBlocks (the bound the statistics partials are sized
 * for). One batch of resident wavefronts: none waits for a slot while others finish their share of the list. */
template <class K>
static int list_grid(const gp_handle *h, K kernel) {
  int nb = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, 256, 0) != hipSuccess || nb < 1) return kListBlocks;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->device) != hipSuccess || cus < 1)
    return kListBlocks;
  const int64_t g = (int64_t)nb * cus;
  return (int)(g < kListBlocks ? g : kListBlocks);
}

template <int N, int NC, typename R, bool SMOOTH, bool P2, bool TG = false>
__global__ __launch_bounds__(kBX *kBY) void k_phi(const Model<R> *__restrict__ mp, const R *__restrict__ phi,
                                                  const R *__restrict__ mu, R *__restrict__ dphi, int64_t fs,
                                                  int64_t ps, int64_t mx, int64_t my, const RowThermo<R> *th) {
  const int64_t y = 2 + (int64_t)blockIdx.x * kBX + threadIdx.x;
  const int64_t x = 2 + (int64_t)blockIdx.y * kBY + threadIdx.y;
  if (x > mx + 3 || y > my + 3) return;
  const Model<R> &m = *mp;
  const int64_t c = x * ps + y + kYOff;
  View<R, P2, TG, kFastOf<R>> v{phi, mu, nullptr, fs, ps, c, th, y};
  R d[N];
  phi_delta<N, NC>(m, v, SMOOTH, d);
  for (int a = 0; a < N; a++) dphi[a * fs + c] = d[a];
}

template <int N, int NC, typename R, bool SMOOTH, bool P2, bool TG = false>
__global__ __launch_bounds__(kBX *kBY) void k_mu(const Model<R> *__restrict__ mp, const R *__restrict__ phi,
                                                 const R *__restrict__ mu, const R *__restrict__ dphi,
                                                 R *__restrict__ phi_o, R *__restrict__ mu_o, Maps maps,
                                                 double *__restrict__ partials, int64_t fs, int64_t ps, int64_t mx,
                                                 int64_t my, const RowThermo<R> *th) {
  const int64_t y = 3 + (int64_t)blockIdx.x * kBX + threadIdx.x;
  const int64_t x = 3 + (int64_t)blockIdx.y * kBY + threadIdx.y;
  const bool live = (x <= mx + 2 && y <= my + 2);
  constexpr int NS = 3 * N + 3 * NC + 1;
  double st[NS];
  for (int a = 0; a < N; a++) {
    st[a] = -INFINITY;
    st[N + a] = INFINITY;
    st[2 * N + 2 * NC + a] = 0.0;
  }
  for (int k = 0; k < NC; k++) {
    st[2 * N + k] = -INFINITY;
    st[2 * N + NC + k] = INFINITY;
    st[3 * N + 2 * NC + k] = 0.0;
  }
  st[NS - 1] = 0.0;
  if (live) {
    const Model<R> &m = *mp;
    const int64_t c = x * ps + y + kYOff;
    View<R, P2, TG, kFastOf<R>> v{phi, mu, dphi, fs, ps, c, th, y};
    R jx0[NC], jxb[NC], jy0[NC], jyl[NC], d[N], pn[N], mn[NC], dm[NC];
    if (!SMOOTH) {
      jat_x<N, NC>(m, v, 0, 0, jx0);
      jat_x<N, NC>(m, v, -1, 0, jxb);
      jat_y<N, NC>(m, v, 0, 0, y, jy0);
      jat_y<N, NC>(m, v, 0, -1, y - 1, jyl);
    } else {
      for (int k = 0; k < NC; k++) jx0[k] = jxb[k] = jy0[k] = jyl[k] = R(0);
    }
    for (int a = 0; a < N; a++) d[a] = dphi[a * fs + c];
    mu_update<N, NC>(m, v, SMOOTH, d, jx0, jxb, jy0, jyl, pn, mn, dm);
    for (int a = 0; a < N; a++) phi_o[a * fs + c] = pn[a];
    for (int k = 0; k < NC; k++) mu_o[k * fs + c] = mn[k];
    if (x < 7 || x > mx - 2 || y < 7 || y > my - 2) {
      write_images(phi_o, N, pn, 0, maps, x, y, fs, ps);
      write_images(mu_o, NC, mn, 1, maps, x, y, fs, ps);
    }
    for (int a = 0; a < N; a++) {
      double pv = (double)pn[a], dv = (double)d[a];
      st[a] = pv;
      st[N + a] = pv;
      st[2 * N + 2 * NC + a] = dv * dv;
      if (!isfinite(pv)) st[NS - 1] += 1.0;
    }
    for (int k = 0; k < NC; k++) {
      double mv = (double)mn[k], dv = (double)dm[k];
      st[2 * N + k] = mv;
      st[2 * N + NC + k] = mv;
      st[3 * N + 2 * NC + k] = dv * dv;
      if (!isfinite(mv)) st[NS - 1] += 1.0;
    }
  }
  /* block reduction: wave shuffles, then the kBY waves through LDS */
  __shared__ double red[kBY][NS];
  for (int i = 0; i < NS; i++) {
    double r;
    if (i < N || (i >= 2 * N && i < 2 * N + NC)) r = wave_max(st[i]);
    else if (i < 2 * N + 2 * NC) r = wave_min(st[i]);
    else r = wave_sum(st[i]);
    if (threadIdx.x == 0) red[threadIdx.y][i] = r;
  }
  __syncthreads();
  if (threadIdx.y == 0 && threadIdx.x < NS) {
    const int i = threadIdx.x;
    double r = red[0][i];
    for (int w = 1; w < kBY; w++) {
      if (i < N || (i >= 2 * N && i < 2 * N + NC)) r = fmax(r, red[w][i]);
      else if (i < 2 * N + 2 * NC) r = fmin(r, red[w][i]);
      else r += red[w][i];
    }
    const int64_t bid = (int64_t)blockIdx.y * gridDim.x + blockIdx.x;
    const int64_t nparts = (int64_t)gridDim.x * gridDim.y;
    partials[i * nparts + bid] = r;  /* statistic-major: the fold reads each statistic contiguously */
  }
}

/* Fold the per-block partials into the cumulative statistics: one block per statistic, each reading its
 * contiguous row of partials (statistic-major layout) in a fixed order, so the result is deterministic. */
constexpr int kStatThreads = 1024;

template <int N, int NC>
__device__ __forceinline__ void stats_fold_block(const double *__restrict__ partials, int64_t stride, int64_t nparts,
                                                 double *__restrict__ cum, int i) {
  constexpr int NS = 3 * N + 3 * NC + 1;
  __shared__ double red[kStatThreads];
  if (i >= NS) return;
  const bool mx = (i < N || (i >= 2 * N && i < 2 * N + NC));
  const bool mn = !mx && i < 2 * N + 2 * NC;
  double r = mx ? -INFINITY : mn ? INFINITY : 0.0;
  const double *row = partials + (int64_t)i * stride;
  /* a thread folds its partials in index order; the loads of eight of them are issued before the first is used */
  int64_t b = threadIdx.x;
  for (; b + 7 * kStatThreads < nparts; b += 8 * kStatThreads) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) v[u] = row[b + u * kStatThreads];
#pragma unroll
    for (int u = 0; u < 8; u++) r = mx ? fmax(r, v[u]) : mn ? fmin(r, v[u]) : r + v[u];
  }
  for (; b < nparts; b += kStatThreads) {
    const double v = row[b];
    r = mx ? fmax(r, v) : mn ? fmin(r, v) : r + v;
  }
  red[threadIdx.x] = r;
  __syncthreads();
  for (int s = kStatThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      double o = red[threadIdx.x + s];
      red[threadIdx.x] = mx ? fmax(red[threadIdx.x], o) : mn ? fmin(red[threadIdx.x], o) : red[threadIdx.x] + o;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double v = red[0];
    if (mx) cum[i] = fmax(cum[i], v);
    else if (mn) cum[i] = fmin(cum[i], v);
    else if (i < NS - 1) cum[i] = v;
    else cum[i] += v;
  }
}

template <int N, int NC>
__global__ __launch_bounds__(kStatThreads) void k_stats(const double *__restrict__ partials, int64_t stride,
                                                        int64_t nparts, double *__restrict__ cum) {
  stats_fold_block<N, NC>(partials, stride, nparts, cum, (int)blockIdx.x);
}

/* The statistics fold and the halo pass of the cell-classified step in one launch (both small and latency-bound,
 * and independent: the fold reads the partials, the halo pass the new state): blocks below the number of
 * statistics fold one statistic each, the others copy kStatThreads halo cells of one plane each (k_halo). */
template <int N, int NC, typename R, bool C>
__global__ __launch_bounds__(kStatThreads) void k_stats_halo(const double *__restrict__ partials, int64_t stride,
                                                             int64_t nparts, double *__restrict__ cum,
                                                             R *__restrict__ phi, R *__restrict__ mu,
                                                             int8_t *__restrict__ codes, float *__restrict__ plo,
                                                             int8_t *__restrict__ mlo, Maps maps, int64_t fs,
                                                             int64_t ps, int64_t mx, int64_t my) {
  constexpr int NS = 3 * N + 3 * NC + 1;
  if ((int)blockIdx.x < NS) {
    stats_fold_block<N, NC>(partials, stride, nparts, cum, (int)blockIdx.x);
    return;
  }
  const int64_t nchunk = (halo_cells(mx, my) + kStatThreads - 1) / kStatThreads, q = blockIdx.x - NS;
  halo_cell<N, NC, R, C>((q % nchunk) * kStatThreads + threadIdx.x, (int)(q / nchunk), phi, mu, codes, plo, mlo, maps,
                         fs, ps, mx, my);
}

/* Global solute and grand-potential energy of the current state (no reference counterpart; the FP64 oracle's
 * oracle_invariants defines them): per interior cell
 *   c_k    = sum_a h_a(phi) c_a^k(mu) dx dy,
 *   energy = [ sum_a h_a psi_a(mu) + eps sum_{a<b} Gamma_ab ac(q_ab)^2 |q_ab|^2 + w(phi) / eps ] dx dy,
 * with q_ab from the centred gradients. Cells are evaluated in the state's precision and accumulated in FP64:
 * a block sum per quantity into partials[i * nblocks + block], folded by k_fold_sum. */
template <int N, int NC, typename R>
__global__ __launch_bounds__(kBX *kBY) void k_invariants(const Model<R> *__restrict__ mp, const R *__restrict__ phi,
                                                         const R *__restrict__ mu, double *__restrict__ partials,
                                                         int64_t fs, int64_t ps, int64_t yoff, int64_t mx,
                                                         int64_t my) {
  const int64_t y = 3 + (int64_t)blockIdx.x * kBX + threadIdx.x;
  const int64_t x = 3 + (int64_t)blockIdx.y * kBY + threadIdx.y;
  double v[NC + 1];
  for (int i = 0; i <= NC; i++) v[i] = 0.0;
  if (x <= mx + 2 && y <= my + 2) {
    const Model<R> &m = *mp;
    const int64_t c = x * ps + y + yoff;
    const double area = (double)m.dx * (double)m.dy;
    R ph[N], mv[NC], gxc[N], gyc[N], h[N];
    for (int a = 0; a < N; a++) {
      ph[a] = phi[a * fs + c];
      gxc[a] = (phi[a * fs + c + ps] - phi[a * fs + c - ps]) / (R(2.0) * m.dx);
      gyc[a] = (phi[a * fs + c + 1] - phi[a * fs + c - 1]) / m.two_dy;
    }
    for (int k = 0; k < NC; k++) mv[k] = mu[k * fs + c];
    for (int a = 0; a < N; a++) h[a] = hphi<N>(ph, a);
    R e = R(0.0);
    for (int a = 0; a < N; a++) {
      R cc[NC], psi = R(0.0);
      for (int k = 0; k < NC; k++) {
        cc[k] = c_mu<NC>(m, mv, a, k);
        v[k] += (double)(h[a] * cc[k]) * area;
        psi -= mv[k] * cc[k];
      }
      psi += free_energy<NC>(m, cc, a);
      e += h[a] * psi;
    }
    R grad = R(0.0), w = R(0.0);
    for (int a = 0; a < N; a++)
      for (int b = a + 1; b < N; b++) {
        R r[2] = {ph[a] * gxc[b] - ph[b] * gxc[a], ph[a] * gyc[b] - ph[b] * gyc[a]}, q[2];
        rot2(m.rot[a][b], r, q);
        const R ac = m.fn_aniso ? aniso_ac(m.dab[a][b], q) : R(1.0);
        grad += m.gamma[a][b] * ac * ac * qnorm2(q);
        if (m.fn_w == 1) w += m.w_obst * m.gamma[a][b] * ph[a] * ph[b];
        else w += R(9.0) * m.gamma[a][b] * ph[a] * ph[a] * ph[b] * ph[b];
        for (int cc = b + 1; cc < N; cc++) w += m.gamma_abc[a][b][cc] * ph[a] * ph[b] * ph[cc];
      }
    e += m.eps * grad + w / m.eps;
    v[NC] = (double)e * area;
  }
  __shared__ double red[NC + 1][kBY];
  const int lane = threadIdx.x & 63;
  for (int i = 0; i <= NC; i++) {
    const double r = wave_sum(v[i]);
    if (lane == 0) red[i][threadIdx.y] = r;
  }
  __syncthreads();
  if (threadIdx.x == 0 && threadIdx.y == 0) {
    const int64_t nb = (int64_t)gridDim.x * gridDim.y, bid = (int64_t)blockIdx.y * gridDim.x + blockIdx.x;
    for (int i = 0; i <= NC; i++) {
      double t = 0.0;
      for (int j = 0; j < kBY; j++) t += red[i][j];
      partials[i * nb + bid] = t;
    }
  }
}

/* out[i] = sum_b partials[i * count + b]: one block per quantity, fixed order (deterministic) */
static __global__ __launch_bounds__(kStatThreads) void k_fold_sum(const double *__restrict__ partials, int64_t count,
                                                          double *__restrict__ out) {
  __shared__ double red[kStatThreads];
  const double *row = partials + (int64_t)blockIdx.x * count;
  double r = 0.0;
  for (int64_t b = threadIdx.x; b < count; b += kStatThreads) r += row[b];
  red[threadIdx.x] = r;
  __syncthreads();
  for (int st = kStatThreads / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

/* solute[NC] and energy of the current state into h->inv_dev */
/* solute[NC] and energy of a state into out (device, NC + 1 doubles) on `stream`, with the block sums in partials
 * (at least (NC + 1) * invariant_blocks(mesh) doubles): the handle's own state (its plane layout) or a caller's in the
 * reference's plain halo-inclusive layout (fs = rows_x * rows_y, ps = rows_y, yoff = 0) */
template <int N, int NC, typename R>
static int launch_invariants_model(gp_handle *h, const void *model, const void *phi, const void *mu, int64_t fs,
                                   int64_t ps, int64_t yoff, double *partials, double *out, hipStream_t stream) {
  const int64_t mx = h->p.mesh_x, my = h->p.mesh_y;
  dim3 blk(kBX, kBY);
  dim3 g2((unsigned)((my + kBX - 1) / kBX), (unsigned)((mx + kBY - 1) / kBY));
  hipLaunchKernelGGL((k_invariants<N, NC, R>), g2, blk, 0, stream, (const Model<R> *)model, (const R *)phi,
                     (const R *)mu, partials, fs, ps, yoff, mx, my);
  hipLaunchKernelGGL(k_fold_sum, dim3(NC + 1), dim3(kStatThreads), 0, stream, (const double *)partials,
                     (int64_t)g2.x * g2.y, out);
  HIPCHK(h, hipGetLastError());
  return GP_OK;
}

template <int N, int NC, typename R>
static int launch_invariants_on(gp_handle *h, const void *phi, const void *mu, int64_t fs, int64_t ps, int64_t yoff,
                                double *partials, double *out, hipStream_t stream) {
  return launch_invariants_model<N, NC, R>(h, h->model, phi, mu, fs, ps, yoff, partials, out, stream);
}

/* solute[NC] and energy of the handle's current state into h->inv_dev */
template <int N, int NC, typename R>
static int launch_invariants(gp_handle *h) {
  return launch_invariants_on<N, NC, R>(h, h->phi[h->cur], h->mu[h->cur], h->fs, h->ps, kYOff, h->partials,
                                        h->inv_dev, h->stream);
}

/* ------------------------------------------------------------------------------------------------ dispatch */

template <int N, int NC, typename R, bool P2>
static int launch_step_p(gp_handle *h, int mode) {
  const R *phi = (const R *)h->phi[h->cur], *mu = (const R *)h->mu[h->cur];
  R *phi_o = (R *)h->phi[1 - h->cur], *mu_o = (R *)h->mu[1 - h->cur];
  R *dphi = (R *)h->dphi;
  const Model<R> *m = (const Model<R> *)h->model;
  const int64_t mx = h->p.mesh_x, my = h->p.mesh_y;
  dim3 blk(kBX, kBY);
  dim3 g1((unsigned)((my + 2 + kBX - 1) / kBX), (unsigned)((mx + 2 + kBY - 1) / kBY));
  dim3 g2((unsigned)((my + kBX - 1) / kBX), (unsigned)((mx + kBY - 1) / kBY));
  const bool smooth = mode == MODE_SMOOTH;
  const bool tg = h->p.tempgrady != 0;
  const RowThermo<R> *thp = (const RowThermo<R> *)h->thp, *thc = (const RowThermo<R> *)h->thc;
  if (mode == MODE_DPHI_PREV) {
    /* delta-phi of the last step, recomputed from the previous state (still in the other buffer; non-isothermal
     * runs: with the last step's row coefficients) */
    const R *pp = (const R *)h->phi[1 - h->cur], *mp = (const R *)h->mu[1 - h->cur];
    if (tg)
      hipLaunchKernelGGL((k_phi<N, NC, R, false, P2, true>), g1, blk, 0, h->stream, m, pp, mp, dphi, h->fs, h->ps,
                         mx, my, thp);
    else
      hipLaunchKernelGGL((k_phi<N, NC, R, false, P2>), g1, blk, 0, h->stream, m, pp, mp, dphi, h->fs, h->ps, mx, my,
                         (const RowThermo<R> *)nullptr);
    HIPCHK(h, hipGetLastError());
    return GP_OK;
  }
  if (mode == MODE_MARCH || mode == MODE_MARCH_STATS) {
    /* delta-phi of the candidate cells, then the update: by cell class (k_cbulk + k_ilist) or by march tiles
     * (k_march near an interface, k_bulk on deep tiles). The update kernels read the current state and its phase
     * codes and write disjoint cells of the other buffers (state and codes). */
    const int nt = (int)h->march_nwaves;
    const unsigned nblk = (unsigned)((nt + 3) / 4);
    const int *fin = h->flags_valid ? h->flags[h->fcur] : nullptr;
    int *fout = h->flags[1 - h->fcur];
    const int64_t pstride = 2 * (int64_t)nt;
    const int8_t *codes = h->codes[h->ccur];
    int8_t *codes_o = h->codes[1 - h->ccur];
    const bool stats = mode == MODE_MARCH_STATS;
    if (!h->codes_valid) {
      hipLaunchKernelGGL((k_codes<N, R>), dim3((unsigned)((h->ry + 63) / 64), (unsigned)((h->rx + 3) / 4)), dim3(256),
                         0, h->stream, phi, h->codes[h->ccur], h->fs, h->ps, h->rx, h->ry);
    }
    if (!h->use_cells) {
      /* delta-phi of the candidate cells, tiles of 64 y-cells by kDTileRows rows over x, y in [2, mesh+3] */
      const int nty = (int)((my + 2 + 63) / 64), ntx = (int)((mx + 2 + kDTileRows - 1) / kDTileRows);
      ProfScope ps(h, GP_KERNEL_DPHI);
      hipLaunchKernelGGL((k_dphi<N, NC, R, P2>), dim3((unsigned)((nty * ntx + 3) / 4)), dim3(256), 0, h->stream, m,
                         phi, mu, codes, dphi, h->fs, h->ps, mx, my, nty, nty * ntx);
    }
    if (h->use_cells) {
      /* statistics partials: k_cbulk's wavefronts, then k_ilist's */
      const int ct = (int)h->cell_nwaves;
      const unsigned cblk = (unsigned)((ct + 3) / 4);
      if (!h->lgrid_d) {
        h->lgrid_d = list_grid(h, k_dlist<N, NC, R, P2>);
        h->lgrid_i = list_grid(h, k_ilist<N, NC, R, P2, true>);
      }
      const int64_t cstride = (int64_t)ct + 4 * h->lgrid_i;
      /* this step's list lengths are zero (the previous step's k_cbulk, or the set-up, cleared them); this
       * step's k_cbulk clears the other pair for the next step, which the previous step's kernels read */
      int *lc = h->lcount + 2 * h->lsel, *lnext = h->lcount + 2 * (1 - h->lsel);
      h->lsel = 1 - h->lsel;
      /* k_cbulk's blocks past the tiles list the candidates of the boundary ring, one thread per ring cell */
      const int64_t nring = 2 * (my + 2) + 2 * mx;
      const unsigned rblk = (unsigned)((nring + 255) / 256);
      {
        ProfScope ps(h, GP_KERNEL_CBULK);
        if (stats)
          hipLaunchKernelGGL((k_cbulk<N, NC, R, P2, true>), dim3(cblk + rblk), dim3(256), 0, h->stream, m, mu, codes,
                             phi_o, mu_o, codes_o, h->partials, cstride, h->fs, h->ps, mx, my, h->cell_seg,
                             h->march_nstrips, ct, (int)h->prev_valid, h->ilist, h->clist, lc, lnext, (int)cblk,
                             (const int8_t *)nullptr, (float *)nullptr, (int8_t *)nullptr);
        else
          hipLaunchKernelGGL((k_cbulk<N, NC, R, P2, false>), dim3(cblk + rblk), dim3(256), 0, h->stream, m, mu, codes,
                             phi_o, mu_o, codes_o, h->partials, cstride, h->fs, h->ps, mx, my, h->cell_seg,
                             h->march_nstrips, ct, (int)h->prev_valid, h->ilist, h->clist, lc, lnext, (int)cblk,
                             (const int8_t *)nullptr, (float *)nullptr, (int8_t *)nullptr);
      }
      {
        /* delta-phi of the listed candidates, full wavefronts */
        ProfScope ps(h, GP_KERNEL_DPHI);
        hipLaunchKernelGGL((k_dlist<N, NC, R, P2>), dim3(h->lgrid_d), dim3(256), 0, h->stream, m, phi, mu, dphi,
                           h->fs, h->ps, (const int32_t *)h->clist, (const int *)lc + 1);
      }
      {
        ProfScope ps(h, GP_KERNEL_CIFACE);
        if (stats)
          hipLaunchKernelGGL((k_ilist<N, NC, R, P2, true>), dim3(h->lgrid_i), dim3(256), 0, h->stream, m, phi, mu,
                             dphi, codes, phi_o, mu_o, codes_o, h->partials + ct, cstride, h->fs, h->ps,
                             (const int32_t *)h->ilist, (const int *)lc);
        else
          hipLaunchKernelGGL((k_ilist<N, NC, R, P2, false>), dim3(h->lgrid_i), dim3(256), 0, h->stream, m, phi, mu,
                             dphi, codes, phi_o, mu_o, codes_o, h->partials + ct, cstride, h->fs, h->ps,
                             (const int32_t *)h->ilist, (const int *)lc);
      }
      /* the new state's halo, from the interior cells both kernels wrote; with the statistics fold when asked */
      const int64_t nh = halo_cells(mx, my);
      if (stats) {
        ProfScope ps(h, GP_KERNEL_STATS);
        hipLaunchKernelGGL((k_stats_halo<N, NC, R, false>),
                           dim3((unsigned)(3 * N + 3 * NC + 1 +
                                           kHaloPlanes<N, NC, false> * ((nh + kStatThreads - 1) / kStatThreads))),
                           dim3(kStatThreads), 0, h->stream, h->partials, cstride, cstride, h->stats_dev, phi_o, mu_o,
                           codes_o, (float *)nullptr, (int8_t *)nullptr, h->maps, h->fs, h->ps, mx, my);
      } else {
        ProfScope ps(h, GP_KERNEL_HALO);
        hipLaunchKernelGGL((k_halo<N, NC, R, false>), dim3((unsigned)((nh + 255) / 256), kHaloPlanes<N, NC, false>),
                           dim3(256), 0, h->stream,
                           phi_o, mu_o, codes_o, (float *)nullptr, (int8_t *)nullptr, h->maps, h->fs, h->ps, mx, my);
      }
      HIPCHK(h, hipGetLastError());
      h->cur = 1 - h->cur;
      h->ccur = 1 - h->ccur;
      h->flags_valid = false; /* this path forms no tile flags */
      h->codes_valid = true;
      return GP_OK;
    }
    if (mode == MODE_MARCH) {
      {
        ProfScope ps(h, GP_KERNEL_FUSED);
        hipLaunchKernelGGL((k_march<N, NC, R, P2, false>), dim3(nblk), dim3(256), 0, h->stream, m, phi, mu, dphi,
                           phi_o, mu_o, codes_o, h->maps, h->partials, pstride, fin, fout, h->fs, h->ps, mx, my,
                           h->march_seg, h->march_nstrips, h->march_nsegs, nt);
      }
      if (fin) {
        ProfScope ps(h, GP_KERNEL_BULK);
        hipLaunchKernelGGL((k_bulk<N, NC, R, P2, false>), dim3(nblk), dim3(256), 0, h->stream, m, mu, phi_o, mu_o,
                           codes_o, h->partials + nt, pstride, fin, fout, h->fs, h->ps, mx, h->march_seg,
                           h->march_nstrips, h->march_nsegs, nt);
      }
    } else {
      {
        ProfScope ps(h, GP_KERNEL_FUSED_STATS);
        hipLaunchKernelGGL((k_march<N, NC, R, P2, true>), dim3(nblk), dim3(256), 0, h->stream, m, phi, mu, dphi,
                           phi_o, mu_o, codes_o, h->maps, h->partials, pstride, fin, fout, h->fs, h->ps, mx, my,
                           h->march_seg, h->march_nstrips, h->march_nsegs, nt);
      }
      if (fin) {
        ProfScope ps(h, GP_KERNEL_BULK);
        hipLaunchKernelGGL((k_bulk<N, NC, R, P2, true>), dim3(nblk), dim3(256), 0, h->stream, m, mu, phi_o, mu_o,
                           codes_o, h->partials + nt, pstride, fin, fout, h->fs, h->ps, mx, h->march_seg,
                           h->march_nstrips, h->march_nsegs, nt);
      }
      ProfScope ps(h, GP_KERNEL_STATS);
      hipLaunchKernelGGL((k_stats<N, NC>), dim3(3 * N + 3 * NC + 1), dim3(kStatThreads), 0, h->stream,
                         h->partials, pstride, fin ? pstride : (int64_t)nt, h->stats_dev);
    }
    HIPCHK(h, hipGetLastError());
    h->cur = 1 - h->cur;
    h->ccur = 1 - h->ccur;
    h->fcur = 1 - h->fcur;
    h->flags_valid = true;
    h->codes_valid = true;
    return GP_OK;
  }
  h->flags_valid = false; /* the two-pass kernels form neither tile flags nor phase codes */
  h->codes_valid = false;
  if (tg) {
    /* non-isothermal: this step's row coefficients, then the two sweeps with them */
    hipLaunchKernelGGL((k_thermo<R>), dim3((unsigned)((h->ry + 255) / 256)), dim3(256), 0, h->stream, h->tp, h->trow,
                       (RowThermo<R> *)h->thp, (RowThermo<R> *)h->thc, h->ry);
    {
      ProfScope ps(h, GP_KERNEL_PHI);
      if (smooth)
        hipLaunchKernelGGL((k_phi<N, NC, R, true, P2, true>), g1, blk, 0, h->stream, m, phi, mu, dphi, h->fs, h->ps,
                           mx, my, thp);
      else
        hipLaunchKernelGGL((k_phi<N, NC, R, false, P2, true>), g1, blk, 0, h->stream, m, phi, mu, dphi, h->fs, h->ps,
                           mx, my, thp);
    }
    {
      ProfScope ps(h, GP_KERNEL_MU);
      if (smooth)
        hipLaunchKernelGGL((k_mu<N, NC, R, true, P2, true>), g2, blk, 0, h->stream, m, phi, mu, dphi, phi_o, mu_o,
                           h->maps, h->partials, h->fs, h->ps, mx, my, thc);
      else
        hipLaunchKernelGGL((k_mu<N, NC, R, false, P2, true>), g2, blk, 0, h->stream, m, phi, mu, dphi, phi_o, mu_o,
                           h->maps, h->partials, h->fs, h->ps, mx, my, thc);
    }
    if (!smooth) {
      ProfScope ps(h, GP_KERNEL_STATS);
      hipLaunchKernelGGL((k_stats<N, NC>), dim3(3 * N + 3 * NC + 1), dim3(kStatThreads), 0, h->stream,
                         h->partials, h->nblocks_mu, h->nblocks_mu, h->stats_dev);
    }
    HIPCHK(h, hipGetLastError());
    h->cur = 1 - h->cur;
    return GP_OK;
  }
  if (smooth) {
    {
      ProfScope ps(h, GP_KERNEL_PHI);
      hipLaunchKernelGGL((k_phi<N, NC, R, true, P2>), g1, blk, 0, h->stream, m, phi, mu, dphi, h->fs, h->ps, mx, my, (const RowThermo<R> *)nullptr);
    }
    {
      ProfScope ps(h, GP_KERNEL_MU);
      hipLaunchKernelGGL((k_mu<N, NC, R, true, P2>), g2, blk, 0, h->stream, m, phi, mu, dphi, phi_o, mu_o, h->maps,
                         h->partials, h->fs, h->ps, mx, my, (const RowThermo<R> *)nullptr);
    }
  } else {
    {
      ProfScope ps(h, GP_KERNEL_PHI);
      hipLaunchKernelGGL((k_phi<N, NC, R, false, P2>), g1, blk, 0, h->stream, m, phi, mu, dphi, h->fs, h->ps, mx, my, (const RowThermo<R> *)nullptr);
    }
    {
      ProfScope ps(h, GP_KERNEL_MU);
      hipLaunchKernelGGL((k_mu<N, NC, R, false, P2>), g2, blk, 0, h->stream, m, phi, mu, dphi, phi_o, mu_o, h->maps,
                         h->partials, h->fs, h->ps, mx, my, (const RowThermo<R> *)nullptr);
    }
    {
      ProfScope ps(h, GP_KERNEL_STATS);
      hipLaunchKernelGGL((k_stats<N, NC>), dim3(3 * N + 3 * NC + 1), dim3(kStatThreads), 0, h->stream,
                         h->partials, h->nblocks_mu, h->nblocks_mu, h->stats_dev);
    }
  }
  HIPCHK(h, hipGetLastError());
  h->cur = 1 - h->cur;
  return GP_OK;
}

template <int N, int NC, typename R>
static int launch_step(gp_handle *h, int mode) {
  return h->pow2 ? launch_step_p<N, NC, R, true>(h, mode) : launch_step_p<N, NC, R, false>(h, mode);
}

/* ------------------------------------------------------------------------------------------------ adjoint */

/* one adjoint step: the current state is s_t (delta-phi of s_t in h->dphi); A->g[1] holds the adjoint of
 * s_{t+1} and becomes, in A->g[0] (then swapped into g[1]), the adjoint of s_t */
template <int N, int NC, bool P2>
static int adj_step_t(gp_handle *h, const double *phi, const double *mu, const double *dphi, const void *model) {
  Adjoint *A = h->adj;
  const int64_t mx = h->p.mesh_x, my = h->p.mesh_y, fs = h->fs;
  double *gn = A->g[1], *gc = A->g[0];
  const dim3 blk(256);
  const dim3 gall((unsigned)((h->ry + 63) / 64), (unsigned)((h->rx + 3) / 4));
  /* stage B ran in adj_dphi_t's k_adj_codes pass, which also reset the stage-Q list length. The bulk cells of
   * stage Q in closed form (plain stores over the whole allocated grid: no clearing pass), then the reverse-mode
   * kernels add the rest; counts[pslot] is the P candidate list of the last adj_dphi_t */
  const Model<double> *md = (const Model<double> *)model;
  hipLaunchKernelGGL((k_adj_cbar<N, NC>), gall, blk, 0, h->stream, md, phi, dphi, gn, gn + (size_t)N * fs, A->cbar,
                     A->bq, (const int8_t *)A->pcode, (const int8_t *)A->cand, A->qlist, A->counts, fs, h->ps, mx,
                     my);
  hipLaunchKernelGGL((k_adj_bulk<N, NC>), gall, blk, 0, h->stream, md, phi, mu, (const int8_t *)A->bq,
                     (const int8_t *)A->pcode, (const double *)A->cbar, gn, gc, gc + (size_t)N * fs, A->gdphi, A->gth,
                     3 * (N * (N - 1) / 2), fs, h->ps, mx, my);
  /* the other cells by reverse mode, one pass per listed cell (grid-stride loops) */
  const dim3 gl(kAdjBlocks);
  hipLaunchKernelGGL((k_adj_qrev<N, NC, P2>), gl, blk, 0, h->stream, md, phi, mu, dphi, gn, gn + (size_t)N * fs, gc,
                     gc + (size_t)N * fs, A->gdphi, A->gth, (const int32_t *)A->qlist, (const int *)A->counts, fs,
                     h->ps);
  hipLaunchKernelGGL((k_adj_prev<N, NC, P2>), gl, blk, 0, h->stream, md, phi, mu, (const double *)A->gdphi, gc,
                     gc + (size_t)N * fs, A->gth, (const int32_t *)A->plist, (const int *)A->counts + A->pslot, fs,
                     h->ps);
  HIPCHK(h, hipGetLastError());
  std::swap(A->g[0], A->g[1]);
  return GP_OK;
}

/* delta-phi (FP64 model) of the state an adjoint step is evaluated at, and the list of its P candidates (into
 * A->plist, its length into counts[1 + psel]; the slots alternate, so that no clearing call is needed) */
template <int N, int NC, bool P2>
static int adj_dphi_t(gp_handle *h, const double *phi, const double *mu, double *dphi, const void *model) {
  Adjoint *A = h->adj;
  const int64_t mx = h->p.mesh_x, my = h->p.mesh_y;
  const int slot = 1 + A->psel, other = 1 + (1 - A->psel);
  A->psel = 1 - A->psel;
  A->pslot = slot;
  const dim3 gall((unsigned)((h->ry + 63) / 64), (unsigned)((h->rx + 3) / 4));
  /* with the codes, stage B of the transposed step that follows (its adjoint input A->g[1]) */
  hipLaunchKernelGGL((k_adj_codes<N, NC>), gall, dim3(256), 0, h->stream, phi, dphi, A->pcode, A->cand, A->plist,
                     A->counts + slot, A->counts + other, A->g[1], h->maps, A->counts, h->fs, h->ps, mx, my);
  static const int grid = list_grid(h, k_dlist<N, NC, double, P2>);
  hipLaunchKernelGGL((k_dlist<N, NC, double, P2>), dim3(grid), dim3(256), 0, h->stream,
                     (const Model<double> *)model, phi, mu, dphi, h->fs, h->ps, (const int32_t *)A->plist,
                     (const int *)A->counts + slot);
  HIPCHK(h, hipGetLastError());
  return GP_OK;
}
template <int N, int NC>
static int adj_dphi_any(gp_handle *h, const double *phi, const double *mu, double *dphi, const void *model) {
  return h->pow2 ? adj_dphi_t<N, NC, true>(h, phi, mu, dphi, model) : adj_dphi_t<N, NC, false>(h, phi, mu, dphi, model);
}

/* FP32 handles: the current state widened into A->st64 (its delta-phi then comes from adj_dphi_t) */
static __global__ __launch_bounds__(256) void k_widen(const float *__restrict__ src, double *__restrict__ dst, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = (double)src[i];
}
template <int N, int NC, bool P2>
static int adj_widen_t(gp_handle *h) {
  Adjoint *A = h->adj;
  const int64_t np = (int64_t)N * h->fs, nm = (int64_t)NC * h->fs;
  hipLaunchKernelGGL(k_widen, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, h->stream,
                     (const float *)h->phi[h->cur], A->st64, np);
  hipLaunchKernelGGL(k_widen, dim3((unsigned)((nm + 255) / 256)), dim3(256), 0, h->stream,
                     (const float *)h->mu[h->cur], A->st64 + np, nm);
  HIPCHK(h, hipGetLastError());
  return GP_OK;
}

template <int N, int NC>
static int adj_step_any(gp_handle *h, const double *phi, const double *mu, const double *dphi, const void *model) {
  return h->pow2 ? adj_step_t<N, NC, true>(h, phi, mu, dphi, model) : adj_step_t<N, NC, false>(h, phi, mu, dphi, model);
}
template <int N, int NC>
static int adj_widen_any(gp_handle *h) {
  return h->pow2 ? adj_widen_t<N, NC, true>(h) : adj_widen_t<N, NC, false>(h);
}

/* ------------------------------------------------------------------------------------------------ GP_FP32C */

/* h->wide[i] (FP64 copy of a decoded state), allocated on first use */
static int c32_wide_buffer(gp_handle *h, int i) {
  if (!h->wide[i] && hipMalloc((void **)&h->wide[i], state_elems(h) * sizeof(double)) != hipSuccess) {
    h->wide[i] = nullptr;
    return gp_fail(h, GP_E_OOM, "FP64 state copy does not fit in device memory");
  }
  return GP_OK;
}

/* decode state buffer b (phi planes, then mu planes) into dst (FP64 planes in the same layout) */
static int c32_widen_state(gp_handle *h, int b, double *dst) {
  const int64_t np = (int64_t)h->N * h->fs, nm = (int64_t)h->NC * h->fs;
  hipLaunchKernelGGL(k_df_widen, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, h->stream,
                     (const float *)h->phi[b], c32_plo(h, b), dst, np);
  hipLaunchKernelGGL(k_c32_widen, dim3((unsigned)((nm + 255) / 256)), dim3(256), 0, h->stream,
                     (const float *)h->mu[b], c32_mlo(h, b), dst + np, nm);
  HIPCHK(h, hipGetLastError());
  return GP_OK;
}

/* encode FP64 planes src into state buffer b */
static int c32_narrow_state(gp_handle *h, const double *src, int b) {
  const int64_t np = (int64_t)h->N * h->fs, nm = (int64_t)h->NC * h->fs;
  hipLaunchKernelGGL(k_df_narrow, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, h->stream, src,
                     (float *)h->phi[b], c32_plo(h, b), np);
  hipLaunchKernelGGL(k_c32_narrow, dim3((unsigned)((nm + 255) / 256)), dim3(256), 0, h->stream, src + np,
                     (float *)h->mu[b], c32_mlo(h, b), nm);
  HIPCHK(h, hipGetLastError());
  return GP_OK;
}

/* One step (or a delta-phi / smoothing request) of the compensated mode. The steps take the cell-classified
 * kernels with the extensions; the smoothing steps and the delta-phi requests run the FP64 two-pass kernels on
 * the decoded state (they are rare: the pre-loop smoothing, downloads of delta-phi). */
template <int N, int NC, bool P2>
static int launch_step_c(gp_handle *h, int mode) {
  const int64_t mx = h->p.mesh_x, my = h->p.mesh_y, fs = h->fs;
  const Model<double> *md = (const Model<double> *)h->model64;
  double *dphi = (double *)h->dphi;
  dim3 blk(kBX, kBY);
  dim3 g1((unsigned)((my + 2 + kBX - 1) / kBX), (unsigned)((mx + 2 + kBY - 1) / kBY));
  dim3 g2((unsigned)((my + kBX - 1) / kBX), (unsigned)((mx + kBY - 1) / kBY));
  int rc;
  if (mode == MODE_DPHI_PREV) {
    const int b = 1 - h->cur;
    if ((rc = c32_wide_buffer(h, 0)) || (rc = c32_widen_state(h, b, h->wide[0]))) return rc;
    hipLaunchKernelGGL((k_phi<N, NC, double, false, P2>), g1, blk, 0, h->stream, md, (const double *)h->wide[0],
                       (const double *)(h->wide[0] + N * fs), dphi, fs, h->ps, mx, my, (const RowThermo<double> *)nullptr);
    HIPCHK(h, hipGetLastError());
    return GP_OK;
  }
  if (mode == MODE_SMOOTH) {
    if ((rc = c32_wide_buffer(h, 0)) || (rc = c32_wide_buffer(h, 1))) return rc;
    if ((rc = c32_widen_state(h, h->cur, h->wide[0]))) return rc;
    HIPCHK(h, hipMemcpyAsync(h->wide[1], h->wide[0], state_elems(h) * sizeof(double), hipMemcpyDeviceToDevice,
                             h->stream));
    const double *wp = h->wide[0], *wm = h->wide[0] + N * fs;
    {
      ProfScope ps(h, GP_KERNEL_PHI);
      hipLaunchKernelGGL((k_phi<N, NC, double, true, P2>), g1, blk, 0, h->stream, md, wp, wm, dphi, fs, h->ps, mx, my, (const RowThermo<double> *)nullptr);
    }
    {
      ProfScope ps(h, GP_KERNEL_MU);
      hipLaunchKernelGGL((k_mu<N, NC, double, true, P2>), g2, blk, 0, h->stream, md, wp, wm, (const double *)dphi,
                         h->wide[1], h->wide[1] + N * fs, h->maps, h->partials, fs, h->ps, mx, my, (const RowThermo<double> *)nullptr);
    }
    HIPCHK(h, hipGetLastError());
    if ((rc = c32_narrow_state(h, h->wide[1], 1 - h->cur))) return rc;
    h->cur = 1 - h->cur;
    h->codes_valid = false;
    h->flags_valid = false;
    return GP_OK;
  }
  if (mode != MODE_MARCH && mode != MODE_MARCH_STATS)
    return gp_fail(h, GP_E_ARG, "the compensated FP32 mode runs the cell-classified step only");
  const float *phi = (const float *)h->phi[h->cur], *mu = (const float *)h->mu[h->cur];
  float *phi_o = (float *)h->phi[1 - h->cur], *mu_o = (float *)h->mu[1 - h->cur];
  const float *plo = c32_plo(h, h->cur);
  const int8_t *mlo = c32_mlo(h, h->cur);
  float *plo_o = c32_plo(h, 1 - h->cur);
  int8_t *mlo_o = c32_mlo(h, 1 - h->cur);
  const Model<float> *m = (const Model<float> *)h->model;
  const int8_t *codes = h->codes[h->ccur];
  int8_t *codes_o = h->codes[1 - h->ccur];
  const bool stats = mode == MODE_MARCH_STATS;
  if (!h->codes_valid)
    hipLaunchKernelGGL((k_c32_codes<N>), dim3((unsigned)((h->ry + 63) / 64), (unsigned)((h->rx + 3) / 4)), dim3(256),
                       0, h->stream, phi, plo, h->codes[h->ccur], fs, h->ps, h->rx, h->ry);
  const int ct = (int)h->cell_nwaves;
  const unsigned cblk = (unsigned)((ct + 3) / 4);
  if (!h->lgrid_d) {
    h->lgrid_d = list_grid(h, k_dlist_c<N, NC, P2>);
    h->lgrid_i = list_grid(h, k_ilist_c<N, NC, P2, true>);
  }
  const int64_t cstride = (int64_t)ct + 4 * h->lgrid_i;
  int *lc = h->lcount + 2 * h->lsel, *lnext = h->lcount + 2 * (1 - h->lsel);
  h->lsel = 1 - h->lsel;
  const int64_t nring = 2 * (my + 2) + 2 * mx;
  const unsigned rblk = (unsigned)((nring + 255) / 256);
  {
    ProfScope ps(h, GP_KERNEL_CBULK);
    if (stats)
      hipLaunchKernelGGL((k_cbulk<N, NC, float, P2, true, true>), dim3(cblk + rblk), dim3(256), 0, h->stream, m, mu,
                         codes, phi_o, mu_o, codes_o, h->partials, cstride, fs, h->ps, mx, my, h->cell_seg,
                         h->march_nstrips, ct, (int)h->prev_valid, h->ilist, h->clist, lc, lnext, (int)cblk, mlo, plo_o,
                         mlo_o);
    else
      hipLaunchKernelGGL((k_cbulk<N, NC, float, P2, false, true>), dim3(cblk + rblk), dim3(256), 0, h->stream, m, mu,
                         codes, phi_o, mu_o, codes_o, h->partials, cstride, fs, h->ps, mx, my, h->cell_seg,
                         h->march_nstrips, ct, (int)h->prev_valid, h->ilist, h->clist, lc, lnext, (int)cblk, mlo, plo_o,
                         mlo_o);
  }
  {
    ProfScope ps(h, GP_KERNEL_DPHI);
    hipLaunchKernelGGL((k_dlist_c<N, NC, P2>), dim3(h->lgrid_d), dim3(256), 0, h->stream, md, phi, plo, mu, mlo, dphi,
                       fs, h->ps, (const int32_t *)h->clist, (const int *)lc + 1);
  }
  {
    ProfScope ps(h, GP_KERNEL_CIFACE);
    if (stats)
      hipLaunchKernelGGL((k_ilist_c<N, NC, P2, true>), dim3(h->lgrid_i), dim3(256), 0, h->stream, md, phi, plo, mu,
                         mlo, (const double *)dphi, codes, phi_o, plo_o, mu_o, mlo_o, codes_o, h->partials + ct,
                         cstride, fs, h->ps, (const int32_t *)h->ilist, (const int *)lc);
    else
      hipLaunchKernelGGL((k_ilist_c<N, NC, P2, false>), dim3(h->lgrid_i), dim3(256), 0, h->stream, md, phi, plo, mu,
                         mlo, (const double *)dphi, codes, phi_o, plo_o, mu_o, mlo_o, codes_o, h->partials + ct,
                         cstride, fs, h->ps, (const int32_t *)h->ilist, (const int *)lc);
  }
  const int64_t nh = halo_cells(mx, my);
  if (stats) {
    ProfScope ps(h, GP_KERNEL_STATS);
    hipLaunchKernelGGL((k_stats_halo<N, NC, float, true>),
                       dim3((unsigned)(3 * N + 3 * NC + 1 +
                                       kHaloPlanes<N, NC, true> * ((nh + kStatThreads - 1) / kStatThreads))),
                       dim3(kStatThreads), 0, h->stream, h->partials, cstride, cstride, h->stats_dev, phi_o, mu_o,
                       codes_o, plo_o, mlo_o, h->maps, fs, h->ps, mx, my);
  } else {
    ProfScope ps(h, GP_KERNEL_HALO);
    hipLaunchKernelGGL((k_halo<N, NC, float, true>), dim3((unsigned)((nh + 255) / 256), kHaloPlanes<N, NC, true>),
                       dim3(256), 0, h->stream,
                       phi_o, mu_o, codes_o, plo_o, mlo_o, h->maps, fs, h->ps, mx, my);
  }
  HIPCHK(h, hipGetLastError());
  h->cur = 1 - h->cur;
  h->ccur = 1 - h->ccur;
  h->flags_valid = false;
  h->codes_valid = true;
  return GP_OK;
}

template <int N, int NC>
static int launch_step_c_any(gp_handle *h, int mode) {
  return h->pow2 ? launch_step_c<N, NC, true>(h, mode) : launch_step_c<N, NC, false>(h, mode);
}

/* solute and energy of the current (decoded) state: FP64 kernels on its FP64 copy */
template <int N, int NC>
static int launch_invariants_c(gp_handle *h) {
  int rc;
  if ((rc = c32_wide_buffer(h, 0)) || (rc = c32_widen_state(h, h->cur, h->wide[0]))) return rc;
  return launch_invariants_model<N, NC, double>(h, h->model64, h->wide[0], h->wide[0] + (size_t)N * h->fs, h->fs,
                                                h->ps, kYOff, h->partials, h->inv_dev, h->stream);
}
/* the same of a caller's FP64 state (gp_invariants_noalloc) */
template <int N, int NC>
static int launch_invariants_on_c(gp_handle *h, const void *phi, const void *mu, int64_t fs, int64_t ps, int64_t yoff,
                                  double *partials, double *out, hipStream_t stream) {
  return launch_invariants_model<N, NC, double>(h, h->model64, phi, mu, fs, ps, yoff, partials, out, stream);
}

/* the adjoint's FP64 evaluation point of a compensated state: the decoded state (its delta-phi then comes from
 * adj_dphi_t) */
template <int N, int NC, bool P2>
static int adj_widen_c(gp_handle *h) {
  return c32_widen_state(h, h->cur, h->adj->st64);
}
template <int N, int NC>
static int adj_widen_c_any(gp_handle *h) {
  return h->pow2 ? adj_widen_c<N, NC, true>(h) : adj_widen_c<N, NC, false>(h);
}

/* the kernel set of (N, NC) in the given precision */
template <int N, int NC>
const GpOps *make_ops(int precision) {
  static const GpOps f64 = {launch_step<N, NC, double>, launch_invariants<N, NC, double>, adj_step_any<N, NC>,
                            adj_widen_any<N, NC>, launch_invariants_on<N, NC, double>, adj_dphi_any<N, NC>};
  static const GpOps f32 = {launch_step<N, NC, float>, launch_invariants<N, NC, float>, adj_step_any<N, NC>,
                            adj_widen_any<N, NC>, launch_invariants_on<N, NC, float>, adj_dphi_any<N, NC>};
  static const GpOps c32 = {launch_step_c_any<N, NC>, launch_invariants_c<N, NC>, adj_step_any<N, NC>,
                            adj_widen_c_any<N, NC>, launch_invariants_on_c<N, NC>, adj_dphi_any<N, NC>};
  return precision == GP_FP64 ? &f64 : precision == GP_FP32C ? &c32 : &f32;
}

}  // namespace gpk
