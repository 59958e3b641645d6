/*
 * gp_hip.hip -- libgp_hip: the C ABI of include/gp_hip.h over hand-written CDNA4 (gfx950) kernels.
 *
 * Device layout (one handle):
 *   every field component is a plane of rows_x = mesh_x+6 rows; a row holds the halo-inclusive y range
 *   0..mesh_y+5 starting at element kYOff, so the first interior cell y = 3 sits on a 64-element boundary;
 *   the row pitch ps is a multiple of 64 elements. phi (N planes) and mu (NC planes) are ping-ponged,
 *   dphi (N planes) holds delta-phi when it is asked for.
 * A time step (default, gp_cells.h) classifies cells by the phase codes of the current state: k_cbulk updates the
 * bulk cells (streaming mu; phi stays e_b) and lists the interface cells and the delta-phi candidates (its blocks
 * past the tiles add the candidates of the boundary ring); k_dlist evaluates delta-phi of the listed candidates and k_ilist the
 * general update of the listed interface cells, both with full wavefronts; k_stats folds the statistics when they
 * are asked for. Every kernel writes the halo images of the boundary cells it updates.
 * GP_PATH=march selects the tiled march (gp_march.h: k_dphi, then k_march near an interface and k_bulk on deep
 * tiles); GP_PATH=twopass the two-pass form (k_phi: delta-phi on x,y in [2, mesh+3]; k_mu: currents and update
 * on the interior), which also runs the smoothing steps and recomputes delta-phi for gp_download.
 * Reference routines replaced: see include/gp_hip.h.
 */
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "gp_common.h"
#include "gp_hip.h"
#include "gp_cells.h"
#include "gp_march.h"
#include "gp_model.h"
#include "gp_adjoint.h"

using namespace gpk;

namespace {

constexpr int kBX = 64;          /* threads along y (one wavefront) */
constexpr int kBY = 4;           /* rows per block */
constexpr int kNStatMax = 3 * GP_MAXP + 3 * GP_MAXNC + 1;
#ifndef GP_LIST_BLOCKS
#define GP_LIST_BLOCKS 2048
#endif
constexpr int kListBlocks = GP_LIST_BLOCKS; /* grid of the list-driven kernels (grid-stride loops), 4 wavefronts each */

char g_create_error[512];

}  // namespace

struct gp_handle {
  gp_params p;
  int precision;
  int N, NC;
  int64_t rx, ry, ps, fs;   /* rows, row pitch (elements), plane size (elements) */
  hipStream_t stream;
  int device;
  void *phi[2], *mu[2], *dphi;
  void *model;              /* Model<R> in device memory */
  double *partials;         /* [statistic][nparts_max], statistic-major */
  double *stats_dev;        /* cumulative stats (kNStatMax) */
  double *inv_dev;          /* solute[NC], energy of the current state */
  int cur;
  int64_t nblocks_mu;
  int64_t steps;
  int64_t nparts_max;       /* rows of the partials buffer */
  int march_seg, march_nstrips, march_nsegs; /* tile grid of the single-pass step (see gp_march.h) */
  int64_t march_nwaves;     /* tiles = wavefronts of k_march and of k_bulk */
  int cell_seg, cell_nsegs;  /* tile grid of k_cbulk (same strips as the march, shorter segments) */
  int64_t cell_nwaves;
  int32_t *clist, *ilist;    /* delta-phi candidates and interface cells of the current step (cell offsets) */
  int *lcount;               /* their lengths: two pairs, used by alternate steps (lcount + 2 * lsel) */
  int lsel;
  int *flags[2];            /* per-tile purity flags of the last two steps */
  int8_t *codes[2];         /* phase codes (pure phase b, or -1) of every cell: codes[ccur] of the current state */
  int ccur;
  bool codes_valid;         /* codes match the current state */
  int fcur;                 /* flags[fcur] describes the current state */
  bool flags_valid;         /* flags[fcur] is valid for the current state (else no tile is taken as deep) */
  bool use_march;           /* single-pass steps (default) or the two-pass form */
  bool use_cells;           /* single-pass update by cell class (k_cbulk + k_ilist, default) or by march tiles */
  bool prev_valid;          /* the other buffer holds the state before the last step */
  Maps maps;
  bool pow2;
  bool profiling;
  struct Adjoint *adj;      /* checkpoints and buffers of the adjoint (gp_adjoint_forward / _backward) */
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> prof;
  char err[512];
};

/* bracket one launch with events when profiling (kind = GP_KERNEL_*) */
struct ProfScope {
  gp_handle *h;
  hipEvent_t e0, e1;
  int kind;
  ProfScope(gp_handle *hh, int k) : h(hh), e0(nullptr), e1(nullptr), kind(k) {
    if (h->profiling && hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess &&
        hipEventRecord(e0, h->stream) != hipSuccess)
      release();
  }
  ~ProfScope() {
    if (e1) {
      if (hipEventRecord(e1, h->stream) == hipSuccess) h->prof.push_back({kind, {e0, e1}});
      else release();
    }
  }
  /* a launch whose events could not be recorded is simply not timed */
  void release() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    e0 = e1 = nullptr;
  }
};

static int fail(gp_handle *h, int code, const char *fmt, ...) {
  char *dst = h ? h->err : g_create_error;
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(dst, 512, fmt, ap);
  va_end(ap);
  return code;
}

#define HIPCHK(h, call)                                                                         \
  do {                                                                                          \
    hipError_t e_ = (call);                                                                     \
    if (e_ != hipSuccess) return fail((h), GP_E_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

/* ------------------------------------------------------------------------------------------------ kernels */

template <int N, int NC, typename R, bool SMOOTH, bool P2>
__global__ __launch_bounds__(kBX *kBY) void k_phi(const Model<R> *__restrict__ mp, const R *__restrict__ phi,
                                                  const R *__restrict__ mu, R *__restrict__ dphi, int64_t fs,
                                                  int64_t ps, int64_t mx, int64_t my) {
  const int64_t y = 2 + (int64_t)blockIdx.x * kBX + threadIdx.x;
  const int64_t x = 2 + (int64_t)blockIdx.y * kBY + threadIdx.y;
  if (x > mx + 3 || y > my + 3) return;
  const Model<R> &m = *mp;
  const int64_t c = x * ps + y + kYOff;
  View<R, P2> v{phi, mu, nullptr, fs, ps, c};
  R d[N];
  phi_delta<N, NC>(m, v, SMOOTH, d);
  for (int a = 0; a < N; a++) dphi[a * fs + c] = d[a];
}

template <int N, int NC, typename R, bool SMOOTH, bool P2>
__global__ __launch_bounds__(kBX *kBY) void k_mu(const Model<R> *__restrict__ mp, const R *__restrict__ phi,
                                                 const R *__restrict__ mu, const R *__restrict__ dphi,
                                                 R *__restrict__ phi_o, R *__restrict__ mu_o, Maps maps,
                                                 double *__restrict__ partials, int64_t fs, int64_t ps, int64_t mx,
                                                 int64_t my) {
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
    View<R, P2> v{phi, mu, dphi, fs, ps, c};
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
__global__ __launch_bounds__(kStatThreads) void k_stats(const double *__restrict__ partials, int64_t stride,
                                                        int64_t nparts, double *__restrict__ cum) {
  constexpr int NS = 3 * N + 3 * NC + 1;
  __shared__ double red[kStatThreads];
  const int i = blockIdx.x;
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

/* Global solute and grand-potential energy of the current state (no reference counterpart; the FP64 oracle's
 * oracle_invariants defines them): per interior cell
 *   c_k    = sum_a h_a(phi) c_a^k(mu) dx dy,
 *   energy = [ sum_a h_a psi_a(mu) + eps sum_{a<b} Gamma_ab ac(q_ab)^2 |q_ab|^2 + w(phi) / eps ] dx dy,
 * with q_ab from the centred gradients. Cells are evaluated in the state's precision and accumulated in FP64:
 * a block sum per quantity into partials[i * nblocks + block], folded by k_fold_sum. */
template <int N, int NC, typename R>
__global__ __launch_bounds__(kBX *kBY) void k_invariants(const Model<R> *__restrict__ mp, const R *__restrict__ phi,
                                                         const R *__restrict__ mu, double *__restrict__ partials,
                                                         int64_t fs, int64_t ps, int64_t mx, int64_t my) {
  const int64_t y = 3 + (int64_t)blockIdx.x * kBX + threadIdx.x;
  const int64_t x = 3 + (int64_t)blockIdx.y * kBY + threadIdx.y;
  double v[NC + 1];
  for (int i = 0; i <= NC; i++) v[i] = 0.0;
  if (x <= mx + 2 && y <= my + 2) {
    const Model<R> &m = *mp;
    const int64_t c = x * ps + y + kYOff;
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
__global__ __launch_bounds__(kStatThreads) void k_fold_sum(const double *__restrict__ partials, int64_t count,
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
template <int N, int NC, typename R>
static int launch_invariants(gp_handle *h) {
  const int64_t mx = h->p.mesh_x, my = h->p.mesh_y;
  dim3 blk(kBX, kBY);
  dim3 g2((unsigned)((my + kBX - 1) / kBX), (unsigned)((mx + kBY - 1) / kBY));
  hipLaunchKernelGGL((k_invariants<N, NC, R>), g2, blk, 0, h->stream, (const Model<R> *)h->model,
                     (const R *)h->phi[h->cur], (const R *)h->mu[h->cur], h->partials, h->fs, h->ps, mx, my);
  hipLaunchKernelGGL(k_fold_sum, dim3(NC + 1), dim3(kStatThreads), 0, h->stream, h->partials, h->nblocks_mu,
                     h->inv_dev);
  HIPCHK(h, hipGetLastError());
  return GP_OK;
}

/* ------------------------------------------------------------------------------------------------ dispatch */

enum { MODE_SMOOTH = 0, MODE_TWOPASS = 1, MODE_MARCH = 2, MODE_DPHI_PREV = 3, MODE_MARCH_STATS = 4, MODE_DPHI_CUR = 5 };

/* Tile grid of the single-pass step: 60-cell y-strips by x-segments of 32 rows (16 when that gives fewer than
 * 4096 tiles, to fill the device on small grids); GP_SEG_LEN overrides the segment length. Each segment
 * recomputes two warm-up rows of delta-phi. */
static void march_plan(gp_handle *h) {
  const int64_t mx = h->p.mesh_x, my = h->p.mesh_y;
  const int64_t nstrips = (my + kMarchOut - 1) / kMarchOut;
  int64_t seg = 32;
  if (nstrips * ((mx + seg - 1) / seg) < 4096) seg = 16;
  const char *e = getenv("GP_SEG_LEN");
  if (e && atoi(e) > 0) seg = atoi(e);
  if (seg > mx) seg = mx;
  h->march_seg = (int)seg;
  h->march_nstrips = (int)nstrips;
  h->march_nsegs = (int)((mx + seg - 1) / seg);
  h->march_nwaves = nstrips * h->march_nsegs;
  /* the cell-classified kernels have almost no per-tile warm-up: 16-row segments (more, smaller tiles) even out
   * the interface work between wavefronts; GP_CELL_SEG overrides (at most kCellSegMax) */
  int64_t cseg = 16;
  const char *ce = getenv("GP_CELL_SEG");
  if (ce && atoi(ce) > 0) cseg = atoi(ce);
  if (cseg > kCellSegMax) cseg = kCellSegMax;
  if (cseg > mx) cseg = mx;
  h->cell_seg = (int)cseg;
  h->cell_nsegs = (int)((mx + cseg - 1) / cseg);
  h->cell_nwaves = nstrips * h->cell_nsegs;
}

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
  if (mode == MODE_DPHI_CUR) {
    /* delta-phi of the current state (the adjoint's recompute) */
    hipLaunchKernelGGL((k_phi<N, NC, R, false, P2>), g1, blk, 0, h->stream, m, phi, mu, dphi, h->fs, h->ps, mx, my);
    HIPCHK(h, hipGetLastError());
    return GP_OK;
  }
  if (mode == MODE_DPHI_PREV) {
    /* delta-phi of the last step, recomputed from the previous state (still in the other buffer) */
    const R *pp = (const R *)h->phi[1 - h->cur], *mp = (const R *)h->mu[1 - h->cur];
    hipLaunchKernelGGL((k_phi<N, NC, R, false, P2>), g1, blk, 0, h->stream, m, pp, mp, dphi, h->fs, h->ps, mx, my);
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
      const int64_t cstride = (int64_t)ct + 4 * kListBlocks;
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
                             phi_o, mu_o, codes_o, h->maps, h->partials, cstride, h->fs, h->ps, mx, my, h->cell_seg,
                             h->march_nstrips, ct, (int)h->prev_valid, h->ilist, h->clist, lc, lnext, (int)cblk);
        else
          hipLaunchKernelGGL((k_cbulk<N, NC, R, P2, false>), dim3(cblk + rblk), dim3(256), 0, h->stream, m, mu, codes,
                             phi_o, mu_o, codes_o, h->maps, h->partials, cstride, h->fs, h->ps, mx, my, h->cell_seg,
                             h->march_nstrips, ct, (int)h->prev_valid, h->ilist, h->clist, lc, lnext, (int)cblk);
      }
      {
        /* delta-phi of the listed candidates, full wavefronts */
        ProfScope ps(h, GP_KERNEL_DPHI);
        hipLaunchKernelGGL((k_dlist<N, NC, R, P2>), dim3(kListBlocks), dim3(256), 0, h->stream, m, phi, mu, dphi,
                           h->fs, h->ps, (const int32_t *)h->clist, (const int *)lc + 1);
      }
      {
        ProfScope ps(h, GP_KERNEL_CIFACE);
        if (stats)
          hipLaunchKernelGGL((k_ilist<N, NC, R, P2, true>), dim3(kListBlocks), dim3(256), 0, h->stream, m, phi, mu,
                             dphi, codes, phi_o, mu_o, codes_o, h->maps, h->partials + ct, cstride, h->fs, h->ps, mx,
                             my, (const int32_t *)h->ilist, (const int *)lc);
        else
          hipLaunchKernelGGL((k_ilist<N, NC, R, P2, false>), dim3(kListBlocks), dim3(256), 0, h->stream, m, phi, mu,
                             dphi, codes, phi_o, mu_o, codes_o, h->maps, h->partials + ct, cstride, h->fs, h->ps, mx,
                             my, (const int32_t *)h->ilist, (const int *)lc);
      }
      if (stats) {
        ProfScope ps(h, GP_KERNEL_STATS);
        hipLaunchKernelGGL((k_stats<N, NC>), dim3(3 * N + 3 * NC + 1), dim3(kStatThreads), 0, h->stream,
                           h->partials, cstride, cstride, h->stats_dev);
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
  if (smooth) {
    {
      ProfScope ps(h, GP_KERNEL_PHI);
      hipLaunchKernelGGL((k_phi<N, NC, R, true, P2>), g1, blk, 0, h->stream, m, phi, mu, dphi, h->fs, h->ps, mx, my);
    }
    {
      ProfScope ps(h, GP_KERNEL_MU);
      hipLaunchKernelGGL((k_mu<N, NC, R, true, P2>), g2, blk, 0, h->stream, m, phi, mu, dphi, phi_o, mu_o, h->maps,
                         h->partials, h->fs, h->ps, mx, my);
    }
  } else {
    {
      ProfScope ps(h, GP_KERNEL_PHI);
      hipLaunchKernelGGL((k_phi<N, NC, R, false, P2>), g1, blk, 0, h->stream, m, phi, mu, dphi, h->fs, h->ps, mx, my);
    }
    {
      ProfScope ps(h, GP_KERNEL_MU);
      hipLaunchKernelGGL((k_mu<N, NC, R, false, P2>), g2, blk, 0, h->stream, m, phi, mu, dphi, phi_o, mu_o, h->maps,
                         h->partials, h->fs, h->ps, mx, my);
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

typedef int (*step_fn)(gp_handle *, int);
typedef int (*inv_fn)(gp_handle *);

template <typename R>
static inv_fn pick_inv(int N, int NC) {
#define GP_CASE(a, b) \
  if (N == a && NC == b) return launch_invariants<a, b, R>;
  GP_CASE(2, 1) GP_CASE(3, 2) GP_CASE(3, 1) GP_CASE(2, 2) GP_CASE(4, 2) GP_CASE(4, 1) GP_CASE(3, 3)
#undef GP_CASE
  return nullptr;
}

template <typename R>
static step_fn pick(int N, int NC) {
#define GP_CASE(a, b) \
  if (N == a && NC == b) return launch_step<a, b, R>;
  GP_CASE(2, 1) GP_CASE(3, 2) GP_CASE(3, 1) GP_CASE(2, 2) GP_CASE(4, 2) GP_CASE(4, 1) GP_CASE(3, 3)
#undef GP_CASE
  return nullptr;
}

static step_fn pick_any(const gp_handle *h) {
  return h->precision == GP_FP64 ? pick<double>(h->N, h->NC) : pick<float>(h->N, h->NC);
}

/* ------------------------------------------------------------------------------------------------ helpers */

static size_t esize(const gp_handle *h) { return h->precision == GP_FP64 ? sizeof(double) : sizeof(float); }

static void stats_reset_host(gp_handle *h, double *s) {
  const int N = h->N, NC = h->NC;
  for (int a = 0; a < N; a++) {
    s[a] = 1.0;      /* phi_max */
    s[N + a] = 0.0;  /* phi_min */
    s[2 * N + 2 * NC + a] = 0.0;
  }
  for (int k = 0; k < NC; k++) {
    s[2 * N + k] = 1.0;       /* mu_max */
    s[2 * N + NC + k] = 1.0;  /* mu_min */
    s[3 * N + 2 * NC + k] = 0.0;
  }
  s[3 * N + 3 * NC] = 0.0;
}

/* halo-inclusive host planes (rows_x*rows_y, y fastest) -> device planes, with the BC maps applied */
static int upload_planes(gp_handle *h, void *dst, const double *src, int ncomp, int field) {
  const int64_t rx = h->rx, ry = h->ry, n = rx * ry;
  std::vector<double> tmp(src, src + (size_t)ncomp * n);
  const int64_t hx[6] = {0, 1, 2, rx - 3, rx - 2, rx - 1}, hy[6] = {0, 1, 2, ry - 3, ry - 2, ry - 1};
  for (int c = 0; c < ncomp; c++) {
    double *v = tmp.data() + (size_t)c * n;
    for (int s = 0; s < 6; s++)
      for (int64_t y = 0; y < ry; y++) v[hx[s] * ry + y] = v[(int64_t)h->p.halo_src_x[field][s] * ry + y];
    for (int s = 0; s < 6; s++)
      for (int64_t x = 0; x < rx; x++) v[x * ry + hy[s]] = v[x * ry + h->p.halo_src_y[field][s]];
  }
  const size_t es = esize(h);
  std::vector<char> stage((size_t)ncomp * h->fs * es, 0);
  for (int c = 0; c < ncomp; c++)
    for (int64_t x = 0; x < rx; x++)
      for (int64_t y = 0; y < ry; y++) {
        const double v = tmp[(size_t)c * n + x * ry + y];
        const size_t o = ((size_t)c * h->fs + x * h->ps + y + kYOff);
        if (h->precision == GP_FP64) ((double *)stage.data())[o] = v;
        else ((float *)stage.data())[o] = (float)v;
      }
  HIPCHK(h, hipMemcpyAsync(dst, stage.data(), stage.size(), hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return GP_OK;
}

static int download_planes(gp_handle *h, double *dst, const void *src, int ncomp) {
  const int64_t rx = h->rx, ry = h->ry, n = rx * ry;
  const size_t es = esize(h);
  std::vector<char> stage((size_t)ncomp * h->fs * es);
  HIPCHK(h, hipMemcpyAsync(stage.data(), src, stage.size(), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  for (int c = 0; c < ncomp; c++)
    for (int64_t x = 0; x < rx; x++)
      for (int64_t y = 0; y < ry; y++) {
        const size_t o = ((size_t)c * h->fs + x * h->ps + y + kYOff);
        dst[(size_t)c * n + x * ry + y] =
            h->precision == GP_FP64 ? ((double *)stage.data())[o] : (double)((float *)stage.data())[o];
      }
  return GP_OK;
}

static int validate(const gp_params *p, char *err) {
  if (p->nphases < 2 || p->nphases > GP_MAXP) return snprintf(err, 512, "nphases %d outside [2,%d]", p->nphases, GP_MAXP), GP_E_ARG;
  if (p->ncomp < 2 || p->ncomp - 1 > GP_MAXNC) return snprintf(err, 512, "ncomp %d outside [2,%d]", p->ncomp, GP_MAXNC + 1), GP_E_ARG;
  if (p->mesh_x < 4 || p->mesh_y < 4) return snprintf(err, 512, "mesh must be at least 4x4"), GP_E_ARG;
  if (p->mesh_x > (1 << 20) || p->mesh_y > (1 << 20)) return snprintf(err, 512, "mesh too large"), GP_E_ARG;
  if (p->fn_w != 1 && p->fn_w != 2) return snprintf(err, 512, "fn_w must be 1 (obstacle) or 2 (well)"), GP_E_ARG;
  if (p->fn_aniso != 0 && p->fn_aniso != 1) return snprintf(err, 512, "fn_aniso must be 0 or 1"), GP_E_ARG;
  if (p->fn_aniso == 1 && p->fold != 4) return snprintf(err, 512, "anisotropy type %d not supported (4 only)", p->fold), GP_E_ARG;
  if (!(p->dx > 0 && p->dy > 0 && p->dt > 0 && p->epsilon > 0)) return snprintf(err, 512, "dx, dy, dt, epsilon must be > 0"), GP_E_ARG;
  if (p->solve < 0 || p->solve > 3) return snprintf(err, 512, "unknown solve mode %d", p->solve), GP_E_ARG;
  if (p->solve == GP_SOLVE_BINARY && p->ncomp != 2) return snprintf(err, 512, "BINARY solve needs ncomp = 2"), GP_E_ARG;
  if (p->solve == GP_SOLVE_TERNARY && p->ncomp != 3) return snprintf(err, 512, "TERNARY solve needs ncomp = 3"), GP_E_ARG;
  const int64_t ex[2] = {p->mesh_x + 2, p->mesh_y + 2};
  for (int f = 0; f < 2; f++)
    for (int k = 0; k < 6; k++) {
      if (p->halo_src_x[f][k] < 3 || p->halo_src_x[f][k] > ex[0] || p->halo_src_y[f][k] < 3 || p->halo_src_y[f][k] > ex[1])
        return snprintf(err, 512, "halo map of field %d entry %d points outside the interior", f, k), GP_E_ARG;
    }
  return GP_OK;
}

/* ------------------------------------------------------------------------------------------------ adjoint */

/* Device buffers of the adjoint: checkpoints every `every` steps, the recomputed states of one segment, the
 * state adjoints of two consecutive steps, the delta-phi adjoint, the parameter gradient and the dual models of
 * every direction chunk of stages Q and P. */
struct Adjoint {
  int64_t nsteps = 0, every = 0;
  std::vector<double *> ckpt;
  std::vector<double *> seg;
  double *g[2] = {nullptr, nullptr};
  double *gdphi = nullptr, *gth = nullptr, *cbar = nullptr;
  int32_t *qlist = nullptr, *plist = nullptr; /* cells of the dual-number stages Q and P */
  /* FP32 handles: the current state widened to FP64, its delta-phi in FP64 and the FP64 model, at which every
   * transposed step is evaluated (the FP64 model's derivative at the FP32 trajectory) */
  double *st64 = nullptr, *dphi64 = nullptr;
  void *model64 = nullptr;
  int *counts = nullptr;                      /* their lengths */
  void *qmodels = nullptr, *pmodels = nullptr;
  int nq = 0, np = 0;
  ~Adjoint() {
    for (double *b : ckpt) (void)hipFree(b);
    for (double *b : seg) (void)hipFree(b);
    void *bufs[] = {g[0], g[1], gdphi, gth, cbar, qlist, plist, counts, qmodels, pmodels, st64, dphi64, model64};
    for (void *b : bufs)
      if (b) (void)hipFree(b);
  }
};

static size_t state_elems(const gp_handle *h) { return (size_t)(h->N + h->NC) * h->fs; }

/* the current state into / from a state buffer laid out phi planes then mu planes */
static int state_save(gp_handle *h, double *dst) {
  const size_t es = esize(h), pb = (size_t)h->N * h->fs * es, mb = (size_t)h->NC * h->fs * es;
  HIPCHK(h, hipMemcpyAsync(dst, h->phi[h->cur], pb, hipMemcpyDeviceToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync((char *)dst + pb, h->mu[h->cur], mb, hipMemcpyDeviceToDevice, h->stream));
  return GP_OK;
}
static int state_load(gp_handle *h, const double *src) {
  const size_t es = esize(h), pb = (size_t)h->N * h->fs * es, mb = (size_t)h->NC * h->fs * es;
  HIPCHK(h, hipMemcpyAsync(h->phi[h->cur], src, pb, hipMemcpyDeviceToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(h->mu[h->cur], (const char *)src + pb, mb, hipMemcpyDeviceToDevice, h->stream));
  h->flags_valid = false;
  h->codes_valid = false;
  h->prev_valid = false;
  return GP_OK;
}

/* halo-inclusive host planes (rows_x*rows_y, y fastest) <-> device planes, without boundary copies: one strided
 * copy (the planes are rows_x rows apart on both sides) */
static int planes_to_device(gp_handle *h, double *dst, const double *src, int ncomp) {
  HIPCHK(h, hipMemcpy2DAsync(dst + kYOff, h->ps * sizeof(double), src, h->ry * sizeof(double), h->ry * sizeof(double),
                             (size_t)ncomp * h->rx, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return GP_OK;
}
static int planes_to_host(gp_handle *h, double *dst, const double *src, int ncomp) {
  HIPCHK(h, hipMemcpy2DAsync(dst, h->ry * sizeof(double), src + kYOff, h->ps * sizeof(double), h->ry * sizeof(double),
                             (size_t)ncomp * h->rx, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return GP_OK;
}

/* the dual models of every chunk of stage Q (directions 9N + 5NC + 5N + theta) and stage P (9N + NC + theta) */
static int adj_models(gp_handle *h) {
  Adjoint *A = h->adj;
  if (A->qmodels) return GP_OK;
  const int N = h->N, NC = h->NC;
  const int dq = 9 * N + 5 * NC + 5 * N, dp = 9 * N + NC;
  const int tq = stage_theta_count(kStageQ, N, NC), tp = stage_theta_count(kStageP, N, NC);
  A->nq = (dq + tq + kAdjK - 1) / kAdjK;
  A->np = (dp + tp + kAdjK - 1) / kAdjK;
  using MD = Model<Dual<kAdjK>>;
  std::vector<MD> q(A->nq), pm(A->np);
  /* theta entry t carries direction base + j of the chunk when stage_theta(stage, j) == t */
  auto seeds = [&](int stage, int base, int nt, int chunk0) {
    return [=](int t, int k) {
      for (int j = 0; j < nt; j++)
        if (stage_theta(stage, j, N, NC) == t) return base + j - chunk0 == k ? 1.0 : 0.0;
      return 0.0;
    };
  };
  for (int i = 0; i < A->nq; i++) make_dual_model_w<kAdjK>(&h->p, seeds(kStageQ, dq, tq, i * kAdjK), &q[i]);
  for (int i = 0; i < A->np; i++) make_dual_model_w<kAdjK>(&h->p, seeds(kStageP, dp, tp, i * kAdjK), &pm[i]);
  if (hipMalloc(&A->qmodels, sizeof(MD) * A->nq) != hipSuccess || hipMalloc(&A->pmodels, sizeof(MD) * A->np) != hipSuccess)
    return fail(h, GP_E_OOM, "adjoint: dual models do not fit in device memory");
  HIPCHK(h, hipMemcpyAsync(A->qmodels, q.data(), sizeof(MD) * A->nq, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(A->pmodels, pm.data(), sizeof(MD) * A->np, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return GP_OK;
}

/* one adjoint step: the current state is s_t (delta-phi of s_t in h->dphi); A->g[1] holds the adjoint of
 * s_{t+1} and becomes, in A->g[0] (then swapped into g[1]), the adjoint of s_t */
template <int N, int NC, bool P2>
static int adj_step_t(gp_handle *h, const double *phi, const double *mu, const double *dphi, const void *model) {
  Adjoint *A = h->adj;
  const int64_t mx = h->p.mesh_x, my = h->p.mesh_y, fs = h->fs;
  double *gn = A->g[1], *gc = A->g[0];
  const dim3 blk(256);
  const dim3 gall((unsigned)((h->ry + 63) / 64), (unsigned)((h->rx + 3) / 4));
  hipLaunchKernelGGL(k_adj_bc, gall, blk, 0, h->stream, gn, N, h->maps, 0, fs, h->ps, h->rx, h->ry);
  hipLaunchKernelGGL(k_adj_bc, gall, blk, 0, h->stream, gn + (size_t)N * fs, NC, h->maps, 1, fs, h->ps, h->rx, h->ry);
  HIPCHK(h, hipMemsetAsync(gc, 0, state_elems(h) * sizeof(double), h->stream));
  HIPCHK(h, hipMemsetAsync(A->gdphi, 0, (size_t)N * fs * sizeof(double), h->stream));
  /* bulk cells of stage Q in closed form (plain stores), then the dual kernels add the rest */
  HIPCHK(h, hipMemsetAsync(A->counts, 0, 2 * sizeof(int), h->stream));
  const Model<double> *md = (const Model<double> *)model;
  hipLaunchKernelGGL((k_adj_cbar<N, NC>), gall, blk, 0, h->stream, md, phi, dphi, gn, gn + (size_t)N * fs, A->cbar,
                     A->qlist, A->counts, fs, h->ps, mx, my);
  const dim3 gb((unsigned)((my + 2 + 63) / 64), (unsigned)((mx + 2 + 3) / 4));
  hipLaunchKernelGGL((k_adj_bulk<N, NC>), gb, blk, 0, h->stream, md, phi, mu, dphi, (const double *)A->cbar, gn, gc,
                     gc + (size_t)N * fs, A->gdphi, A->gth, 3 * (N * (N - 1) / 2), fs, h->ps, mx, my);
  /* the other cells by dual numbers, over compacted lists (a grid-stride loop of one resident round) */
  using MD = Model<Dual<kAdjK>>;
  const dim3 gl(512);
  for (int i = 0; i < A->nq; i++)
    hipLaunchKernelGGL((k_adj_q<N, NC, P2>), gl, blk, 0, h->stream, (const MD *)A->qmodels + i, phi, mu, dphi, gn,
                       gn + (size_t)N * fs, gc, gc + (size_t)N * fs, A->gdphi, A->gth, (const int32_t *)A->qlist,
                       (const int *)A->counts, fs, h->ps, i * kAdjK);
  hipLaunchKernelGGL((k_adj_plist<N>), gb, blk, 0, h->stream, phi, (const double *)A->gdphi, A->plist, A->counts + 1, fs,
                     h->ps, mx, my);
  for (int i = 0; i < A->np; i++)
    hipLaunchKernelGGL((k_adj_p<N, NC, P2>), gl, blk, 0, h->stream, (const MD *)A->pmodels + i, phi, mu, A->gdphi, gc,
                       gc + (size_t)N * fs, A->gth, (const int32_t *)A->plist, (const int *)A->counts + 1, fs, h->ps,
                       i * kAdjK);
  HIPCHK(h, hipGetLastError());
  std::swap(A->g[0], A->g[1]);
  return GP_OK;
}

/* fold the halo adjoints of A->g[1] into their source cells */
static int adj_fold(gp_handle *h) {
  Adjoint *A = h->adj;
  const dim3 gall((unsigned)((h->ry + 63) / 64), (unsigned)((h->rx + 3) / 4));
  hipLaunchKernelGGL(k_adj_bc, gall, dim3(256), 0, h->stream, A->g[1], h->N, h->maps, 0, h->fs, h->ps, h->rx, h->ry);
  hipLaunchKernelGGL(k_adj_bc, gall, dim3(256), 0, h->stream, A->g[1] + (size_t)h->N * h->fs, h->NC, h->maps, 1, h->fs,
                     h->ps, h->rx, h->ry);
  HIPCHK(h, hipGetLastError());
  return GP_OK;
}

typedef int (*adj_fn)(gp_handle *, const double *, const double *, const double *, const void *);

/* FP32 handles: the current state widened into A->st64 and its delta-phi (FP64 model) into A->dphi64 */
__global__ __launch_bounds__(256) void k_widen(const float *__restrict__ src, double *__restrict__ dst, int64_t n) {
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
  const int64_t mx = h->p.mesh_x, my = h->p.mesh_y;
  dim3 g1((unsigned)((my + 2 + kBX - 1) / kBX), (unsigned)((mx + 2 + kBY - 1) / kBY));
  hipLaunchKernelGGL((k_phi<N, NC, double, false, P2>), g1, dim3(kBX, kBY), 0, h->stream,
                     (const Model<double> *)A->model64, (const double *)A->st64, (const double *)(A->st64 + np),
                     A->dphi64, h->fs, h->ps, mx, my);
  HIPCHK(h, hipGetLastError());
  return GP_OK;
}
typedef int (*widen_fn)(gp_handle *);
static widen_fn pick_widen(const gp_handle *h) {
#define GP_CASE(a, b) \
  if (h->N == a && h->NC == b) return h->pow2 ? adj_widen_t<a, b, true> : adj_widen_t<a, b, false>;
  GP_CASE(2, 1) GP_CASE(3, 2) GP_CASE(3, 1) GP_CASE(2, 2) GP_CASE(4, 2) GP_CASE(4, 1) GP_CASE(3, 3)
#undef GP_CASE
  return nullptr;
}

static adj_fn pick_adj(const gp_handle *h) {
#define GP_CASE(a, b) \
  if (h->N == a && h->NC == b) return h->pow2 ? adj_step_t<a, b, true> : adj_step_t<a, b, false>;
  GP_CASE(2, 1) GP_CASE(3, 2) GP_CASE(3, 1) GP_CASE(2, 2) GP_CASE(4, 2) GP_CASE(4, 1) GP_CASE(3, 3)
#undef GP_CASE
  return nullptr;
}

/* ------------------------------------------------------------------------------------------------ C ABI */

extern "C" {

const char *gp_version(void) { return "gp_hip 0.1 gfx950"; }

const char *gp_last_error(const gp_handle *h) { return h ? h->err : g_create_error; }

int gp_create(const gp_params *params, int precision, gp_handle **out) {
  if (!params || !out) return fail(nullptr, GP_E_ARG, "null argument");
  *out = nullptr;
  int rc = validate(params, g_create_error);
  if (rc) return rc;
  if (precision != GP_FP64 && precision != GP_FP32) return fail(nullptr, GP_E_ARG, "unknown precision %d", precision);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(nullptr, GP_E_NODEV, "no HIP device");
  gp_handle *h = new gp_handle();
  memset(h->err, 0, sizeof h->err);
  h->p = *params;
  h->precision = precision;
  h->N = params->nphases;
  h->NC = params->ncomp - 1;
  h->rx = params->mesh_x + 6;
  h->ry = params->mesh_y + 6;
  h->ps = ((h->ry + kYOff + 63) / 64) * 64;
  h->fs = h->rx * h->ps;
  h->cur = 0;
  h->steps = 0;
  if (!pick_any(h)) {
    delete h;
    return fail(nullptr, GP_E_ARG, "no kernel instantiated for %d phases / %d components", params->nphases, params->ncomp);
  }
  const int64_t hx[6] = {0, 1, 2, h->rx - 3, h->rx - 2, h->rx - 1}, hy[6] = {0, 1, 2, h->ry - 3, h->ry - 2, h->ry - 1};
  for (int f = 0; f < 2; f++)
    for (int k = 0; k < 6; k++) {
      h->maps.sx[f][k] = params->halo_src_x[f][k];
      h->maps.sy[f][k] = params->halo_src_y[f][k];
    }
  for (int k = 0; k < 6; k++) {
    h->maps.hx[k] = (int32_t)hx[k];
    h->maps.hy[k] = (int32_t)hy[k];
  }
  if (hipGetDevice(&h->device) != hipSuccess) {
    delete h;
    return fail(nullptr, GP_E_NODEV, "hipGetDevice failed");
  }
  h->nblocks_mu = ((params->mesh_y + kBX - 1) / kBX) * ((params->mesh_x + kBY - 1) / kBY);
  march_plan(h);
  h->nparts_max = std::max(h->nblocks_mu, std::max(2 * h->march_nwaves, h->cell_nwaves + 4 * (int64_t)kListBlocks));
  const char *path = getenv("GP_PATH");
  h->use_march = !(path && strcmp(path, "twopass") == 0);
  /* the cell lists hold 32-bit cell indices: grids of 2^31 cells or more (about 46000^2) take the tiled march,
   * which gives the same results */
  h->use_cells = h->use_march && !(path && strcmp(path, "march") == 0) && h->fs <= 0x7fffffff;
  h->prev_valid = false;
  h->fcur = 0;
  h->flags_valid = false;
  h->codes_valid = false;
  const size_t es = precision == GP_FP64 ? 8 : 4;
  const size_t pn = (size_t)h->N * h->fs * es, mn = (size_t)h->NC * h->fs * es;
  hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
  bool ok = e == hipSuccess;
  ok = ok && hipMalloc(&h->phi[0], pn) == hipSuccess && hipMalloc(&h->phi[1], pn) == hipSuccess;
  ok = ok && hipMalloc(&h->mu[0], mn) == hipSuccess && hipMalloc(&h->mu[1], mn) == hipSuccess;
  ok = ok && hipMalloc(&h->dphi, pn) == hipSuccess;
  ok = ok && hipMalloc((void **)&h->partials, sizeof(double) * kNStatMax * h->nparts_max) == hipSuccess;
  ok = ok && hipMalloc((void **)&h->stats_dev, sizeof(double) * kNStatMax) == hipSuccess;
  ok = ok && hipMalloc((void **)&h->inv_dev, sizeof(double) * (GP_MAXNC + 1)) == hipSuccess;
  ok = ok && hipMalloc((void **)&h->flags[0], sizeof(int) * h->march_nwaves) == hipSuccess;
  ok = ok && hipMalloc((void **)&h->flags[1], sizeof(int) * h->march_nwaves) == hipSuccess;
  ok = ok && hipMalloc((void **)&h->clist, (size_t)h->fs * sizeof(int32_t)) == hipSuccess &&
       hipMalloc((void **)&h->ilist, (size_t)h->fs * sizeof(int32_t)) == hipSuccess &&
       hipMalloc((void **)&h->lcount, 4 * sizeof(int)) == hipSuccess;
  ok = ok && hipMalloc((void **)&h->codes[0], (size_t)h->rx * h->ps) == hipSuccess &&
       hipMalloc((void **)&h->codes[1], (size_t)h->rx * h->ps) == hipSuccess;
  h->ccur = 0;
  const size_t msize = precision == GP_FP64 ? sizeof(Model<double>) : sizeof(Model<float>);
  ok = ok && hipMalloc(&h->model, msize) == hipSuccess;
  if (!ok) {
    gp_destroy(h);
    return fail(nullptr, GP_E_OOM, "device allocation failed (%zu bytes per phi copy)", pn);
  }
  ok = hipMemsetAsync(h->phi[0], 0, pn, h->stream) == hipSuccess && hipMemsetAsync(h->phi[1], 0, pn, h->stream) == hipSuccess &&
       hipMemsetAsync(h->mu[0], 0, mn, h->stream) == hipSuccess && hipMemsetAsync(h->mu[1], 0, mn, h->stream) == hipSuccess &&
       hipMemsetAsync(h->dphi, 0, pn, h->stream) == hipSuccess &&
       hipMemsetAsync(h->lcount, 0, 4 * sizeof(int), h->stream) == hipSuccess;
  h->lsel = 0;
  if (precision == GP_FP64) {
    Model<double> m;
    make_model(params, &m);
    h->pow2 = m.pow2 != 0;
    ok = ok && hipMemcpyAsync(h->model, &m, sizeof m, hipMemcpyHostToDevice, h->stream) == hipSuccess;
    ok = ok && hipStreamSynchronize(h->stream) == hipSuccess;
  } else {
    Model<float> m;
    make_model(params, &m);
    h->pow2 = m.pow2 != 0;
    ok = ok && hipMemcpyAsync(h->model, &m, sizeof m, hipMemcpyHostToDevice, h->stream) == hipSuccess;
    ok = ok && hipStreamSynchronize(h->stream) == hipSuccess;
  }
  if (!ok) {
    gp_destroy(h);
    return fail(nullptr, GP_E_HIP, "device initialisation failed");
  }
  double s[kNStatMax] = {0};
  stats_reset_host(h, s);
  if (hipMemcpy(h->stats_dev, s, sizeof s, hipMemcpyHostToDevice) != hipSuccess) {
    gp_destroy(h);
    return fail(nullptr, GP_E_HIP, "device initialisation failed");
  }
  *out = h;
  return GP_OK;
}

void gp_destroy(gp_handle *h) {
  if (!h) return;
  /* teardown: errors here have no caller to report to, so they are ignored explicitly */
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  for (auto &r : h->prof) {
    (void)hipEventDestroy(r.second.first);
    (void)hipEventDestroy(r.second.second);
  }
  h->prof.clear();
  delete h->adj;
  void *bufs[] = {h->phi[0], h->phi[1], h->mu[0], h->mu[1], h->dphi, h->partials, h->stats_dev, h->model,
                  h->flags[0], h->flags[1], h->inv_dev, h->codes[0], h->codes[1], h->clist, h->ilist, h->lcount};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

int gp_upload(gp_handle *h, const double *phi, const double *mu) {
  if (!h || !phi || !mu) return h ? fail(h, GP_E_ARG, "null argument") : GP_E_ARG;
  int rc = upload_planes(h, h->phi[h->cur], phi, h->N, 0);
  if (rc) return rc;
  rc = upload_planes(h, h->mu[h->cur], mu, h->NC, 1);
  if (rc) return rc;
  /* the other copy starts identical so its halo rows are defined */
  HIPCHK(h, hipMemcpyAsync(h->phi[1 - h->cur], h->phi[h->cur], (size_t)h->N * h->fs * esize(h), hipMemcpyDeviceToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(h->mu[1 - h->cur], h->mu[h->cur], (size_t)h->NC * h->fs * esize(h), hipMemcpyDeviceToDevice, h->stream));
  HIPCHK(h, hipMemsetAsync(h->dphi, 0, (size_t)h->N * h->fs * esize(h), h->stream));
  double s[kNStatMax] = {0};
  stats_reset_host(h, s);
  HIPCHK(h, hipMemcpyAsync(h->stats_dev, s, sizeof s, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->steps = 0;
  h->prev_valid = false;
  h->flags_valid = false;
  h->codes_valid = false;
  return GP_OK;
}

int gp_download(gp_handle *h, double *phi, double *mu, double *dphi) {
  if (!h) return GP_E_ARG;
  int rc;
  if (phi && (rc = download_planes(h, phi, h->phi[h->cur], h->N))) return rc;
  if (mu && (rc = download_planes(h, mu, h->mu[h->cur], h->NC))) return rc;
  if (dphi) {
    if (h->prev_valid) {
      /* single-pass steps do not store delta-phi: recompute the last step's from the previous state */
      if ((rc = pick_any(h)(h, MODE_DPHI_PREV))) return rc;
    }
    if ((rc = download_planes(h, dphi, h->dphi, h->N))) return rc;
  }
  return GP_OK;
}

int gp_upload_fields(gp_handle *h, const gp_cell *cells, int64_t ncells) {
  if (!h || !cells) return h ? fail(h, GP_E_ARG, "null argument") : GP_E_ARG;
  const int64_t n = h->rx * h->ry;
  if (ncells != n) return fail(h, GP_E_ARG, "ncells %lld != rows_x*rows_y %lld", (long long)ncells, (long long)n);
  std::vector<double> phi((size_t)h->N * n), mu((size_t)h->NC * n);
  for (int64_t i = 0; i < n; i++) {
    for (int a = 0; a < h->N; a++) phi[(size_t)a * n + i] = cells[i].phia[a];
    for (int k = 0; k < h->NC; k++) mu[(size_t)k * n + i] = cells[i].compi[k];
  }
  return gp_upload(h, phi.data(), mu.data());
}

int gp_download_fields(gp_handle *h, gp_cell *cells, int64_t ncells) {
  if (!h || !cells) return h ? fail(h, GP_E_ARG, "null argument") : GP_E_ARG;
  const int64_t n = h->rx * h->ry;
  if (ncells != n) return fail(h, GP_E_ARG, "ncells %lld != rows_x*rows_y %lld", (long long)ncells, (long long)n);
  std::vector<double> phi((size_t)h->N * n), mu((size_t)h->NC * n), dphi((size_t)h->N * n);
  int rc = gp_download(h, phi.data(), mu.data(), dphi.data());
  if (rc) return rc;
  for (int64_t i = 0; i < n; i++) {
    for (int a = 0; a < h->N; a++) {
      cells[i].phia[a] = phi[(size_t)a * n + i];
      if (cells[i].deltaphi) cells[i].deltaphi[a] = dphi[(size_t)a * n + i];
    }
    for (int k = 0; k < h->NC; k++) cells[i].compi[k] = mu[(size_t)k * n + i];
  }
  return GP_OK;
}

int gp_smooth(gp_handle *h, int64_t nsteps) {
  if (!h || nsteps < 0) return h ? fail(h, GP_E_ARG, "bad nsteps") : GP_E_ARG;
  step_fn f = pick_any(h);
  for (int64_t t = 0; t < nsteps; t++) {
    int rc = f(h, MODE_SMOOTH);
    if (rc) return rc;
  }
  if (nsteps > 0) h->prev_valid = false; /* dphi now holds the smoothing increment */
  return GP_OK;
}

int gp_get_stats(gp_handle *h, gp_stats *st) {
  if (!h || !st) return h ? fail(h, GP_E_ARG, "null argument") : GP_E_ARG;
  double s[kNStatMax], inv[GP_MAXNC + 1];
  inv_fn fi = h->precision == GP_FP64 ? pick_inv<double>(h->N, h->NC) : pick_inv<float>(h->N, h->NC);
  int rc = fi(h);
  if (rc) return rc;
  HIPCHK(h, hipMemcpyAsync(s, h->stats_dev, sizeof s, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipMemcpyAsync(inv, h->inv_dev, sizeof(double) * (h->NC + 1), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  memset(st, 0, sizeof *st);
  const int N = h->N, NC = h->NC;
  for (int a = 0; a < N; a++) {
    st->phi_max[a] = s[a];
    st->phi_min[a] = s[N + a];
    st->rel_change_phi[a] = s[2 * N + 2 * NC + a];
  }
  for (int k = 0; k < NC; k++) {
    st->mu_max[k] = s[2 * N + k];
    st->mu_min[k] = s[2 * N + NC + k];
    st->rel_change_mu[k] = s[3 * N + 2 * NC + k];
  }
  st->nonfinite = (int64_t)s[3 * N + 3 * NC];
  for (int k = 0; k < NC; k++) st->solute[k] = inv[k];
  st->energy = inv[NC];
  st->step = h->steps;
  return GP_OK;
}

int gp_step(gp_handle *h, int64_t nsteps, gp_stats *stats) {
  if (!h || nsteps < 0) return h ? fail(h, GP_E_ARG, "bad nsteps") : GP_E_ARG;
  step_fn f = pick_any(h);
  for (int64_t t = 0; t < nsteps; t++) {
    int rc = f(h, h->use_march ? (stats ? MODE_MARCH_STATS : MODE_MARCH) : MODE_TWOPASS);
    if (rc) return rc;
    h->steps++;
    h->prev_valid = h->use_march;
  }
  if (stats) {
    int rc = gp_get_stats(h, stats);
    if (rc) return rc;
    if (stats->nonfinite) return fail(h, GP_E_NAN, "%lld non-finite values after step %lld", (long long)stats->nonfinite, (long long)h->steps);
  }
  return GP_OK;
}

int gp_theta_count(const gp_handle *h) { return h ? theta_count(h->N, h->NC) : GP_E_ARG; }

int gp_theta_values(const gp_handle *h, double *theta) {
  if (!h || !theta) return GP_E_ARG;
  const gp_params *p = &h->p;
  const int N = h->N, NC = h->NC, L = N - 1;
  int t = 0;
  for (int a = 0; a < N; a++)
    for (int b = a + 1; b < N; b++) {
      theta[t++] = p->dab[a * GP_MAXP + b];
      theta[t++] = p->tau_ab[a * GP_MAXP + b];
      theta[t++] = p->gamma[a * GP_MAXP + b];
    }
  for (int k = 0; k < NC; k++) theta[t++] = p->diffusivity[(L * GP_MAXNC + k) * GP_MAXNC + k];
  theta[t++] = p->tau;
  return GP_OK;
}

int gp_adjoint_forward(gp_handle *h, int64_t nsteps, int64_t every) {
  if (!h || nsteps < 0) return h ? fail(h, GP_E_ARG, "bad nsteps") : GP_E_ARG;
  if (!pick_adj(h) || !pick_widen(h)) return fail(h, GP_E_ARG, "no adjoint instantiated for %d phases / %d components", h->N, h->NC);
  if (!h->use_march) return fail(h, GP_E_ARG, "the adjoint records single-pass steps (GP_PATH=twopass is set)");
  if (h->fs > 0x7fffffff) return fail(h, GP_E_ARG, "the adjoint indexes cells with 32 bits: grid too large");
  delete h->adj;
  h->adj = new Adjoint();
  Adjoint *A = h->adj;
  if (every <= 0) every = (int64_t)ceil(sqrt((double)(nsteps > 0 ? nsteps : 1)));
  A->nsteps = nsteps;
  A->every = every;
  const int64_t nck = (nsteps + every - 1) / every;
  for (int64_t k = 0; k < nck; k++) {
    double *b = nullptr;
    if (hipMalloc((void **)&b, state_elems(h) * esize(h)) != hipSuccess)
      return fail(h, GP_E_OOM, "adjoint: checkpoint %lld does not fit in device memory", (long long)k);
    A->ckpt.push_back(b);
    int rc = state_save(h, b);
    if (!rc) rc = gp_step(h, std::min(every, nsteps - k * every), nullptr);
    if (rc) return rc;
  }
  return GP_OK;
}

int gp_adjoint_backward(gp_handle *h, const double *gphi_T, const double *gmu_T, double *gphi_0, double *gmu_0,
                        double *gtheta) {
  if (!h || !gphi_T || !gmu_T || !gphi_0 || !gmu_0 || !gtheta) return h ? fail(h, GP_E_ARG, "null argument") : GP_E_ARG;
  Adjoint *A = h->adj;
  if (!A) return fail(h, GP_E_ARG, "gp_adjoint_forward must run first");
  const int nth = theta_count(h->N, h->NC);
  const size_t se = state_elems(h);
  if (!A->g[0]) {
    bool ok = hipMalloc((void **)&A->g[0], se * sizeof(double)) == hipSuccess &&
              hipMalloc((void **)&A->g[1], se * sizeof(double)) == hipSuccess &&
              hipMalloc((void **)&A->gdphi, (size_t)h->N * h->fs * sizeof(double)) == hipSuccess &&
              hipMalloc((void **)&A->gth, sizeof(double) * nth) == hipSuccess &&
              hipMalloc((void **)&A->cbar, (size_t)h->NC * h->fs * sizeof(double)) == hipSuccess &&
              hipMalloc((void **)&A->qlist, (size_t)h->fs * sizeof(int32_t)) == hipSuccess &&
              hipMalloc((void **)&A->plist, (size_t)h->fs * sizeof(int32_t)) == hipSuccess &&
              hipMalloc((void **)&A->counts, 2 * sizeof(int)) == hipSuccess;
    for (int64_t j = 0; ok && j < std::min(A->every, A->nsteps); j++) {
      double *b = nullptr;
      ok = hipMalloc((void **)&b, se * esize(h)) == hipSuccess;
      if (ok) A->seg.push_back(b);
    }
    if (ok && h->precision == GP_FP32) {
      Model<double> m;
      make_model(&h->p, &m);
      ok = hipMalloc((void **)&A->st64, se * sizeof(double)) == hipSuccess &&
           hipMalloc((void **)&A->dphi64, (size_t)h->N * h->fs * sizeof(double)) == hipSuccess &&
           hipMalloc(&A->model64, sizeof m) == hipSuccess &&
           hipMemcpyAsync(A->model64, &m, sizeof m, hipMemcpyHostToDevice, h->stream) == hipSuccess &&
           hipStreamSynchronize(h->stream) == hipSuccess;
    }
    if (!ok) return fail(h, GP_E_OOM, "adjoint: buffers do not fit in device memory");
  }
  int rc = adj_models(h);
  if (rc) return rc;
  HIPCHK(h, hipMemsetAsync(A->g[1], 0, se * sizeof(double), h->stream));
  if ((rc = planes_to_device(h, A->g[1], gphi_T, h->N))) return rc;
  if ((rc = planes_to_device(h, A->g[1] + (size_t)h->N * h->fs, gmu_T, h->NC))) return rc;
  HIPCHK(h, hipMemsetAsync(A->gth, 0, sizeof(double) * nth, h->stream));
  step_fn f = pick_any(h);
  adj_fn fa = pick_adj(h);
  widen_fn fw = pick_widen(h);
  const bool wide = h->precision == GP_FP64;
  for (int64_t k = (int64_t)A->ckpt.size() - 1; k >= 0; k--) {
    const int64_t n = std::min(A->every, A->nsteps - k * A->every);
    if ((rc = state_load(h, A->ckpt[k]))) return rc;
    for (int64_t j = 0; j < n; j++) {
      if ((rc = state_save(h, A->seg[j]))) return rc;
      if (j + 1 < n && (rc = gp_step(h, 1, nullptr))) return rc;
    }
    for (int64_t j = n - 1; j >= 0; j--) {
      ProfScope ps(h, GP_KERNEL_ADJ);
      if ((rc = state_load(h, A->seg[j]))) return rc;
      if (wide) {
        if ((rc = f(h, MODE_DPHI_CUR))) return rc;
        rc = fa(h, (const double *)h->phi[h->cur], (const double *)h->mu[h->cur], (const double *)h->dphi, h->model);
      } else {
        if ((rc = fw(h))) return rc;
        rc = fa(h, A->st64, A->st64 + (size_t)h->N * h->fs, A->dphi64, A->model64);
      }
      if (rc) return rc;
    }
  }
  if ((rc = adj_fold(h))) return rc;
  if ((rc = planes_to_host(h, gphi_0, A->g[1], h->N))) return rc;
  if ((rc = planes_to_host(h, gmu_0, A->g[1] + (size_t)h->N * h->fs, h->NC))) return rc;
  HIPCHK(h, hipMemcpyAsync(gtheta, A->gth, sizeof(double) * nth, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->steps = 0; /* the handle holds the initial state again */
  return GP_OK;
}

int gp_profile_enable(gp_handle *h, int enable) {
  if (!h) return GP_E_ARG;
  h->profiling = enable != 0;
  return GP_OK;
}

const char *gp_kernel_name(int kind) {
  static const char *names[GP_NKERNELS] = {"k_phi", "k_mu", "k_stats", "k_march", "k_march_stats", "k_bulk",
                                           "k_dphi", "adj_step", "k_cbulk", "k_ilist"};
  return (kind >= 0 && kind < GP_NKERNELS) ? names[kind] : "?";
}

int gp_profile_read(gp_handle *h, double *avg_ms, int64_t *count) {
  if (!h || !avg_ms || !count) return h ? fail(h, GP_E_ARG, "null argument") : GP_E_ARG;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  double tot[GP_NKERNELS] = {0};
  int64_t cnt[GP_NKERNELS] = {0};
  for (auto &r : h->prof) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, r.second.first, r.second.second) == hipSuccess && r.first >= 0 && r.first < GP_NKERNELS) {
      tot[r.first] += ms;
      cnt[r.first]++;
    }
    (void)hipEventDestroy(r.second.first);
    (void)hipEventDestroy(r.second.second);
  }
  h->prof.clear();
  for (int k = 0; k < GP_NKERNELS; k++) {
    avg_ms[k] = cnt[k] ? tot[k] / (double)cnt[k] : 0.0;
    count[k] = cnt[k];
  }
  return GP_OK;
}

int gp_synchronize(gp_handle *h) {
  if (!h) return GP_E_ARG;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return GP_OK;
}

}  // extern "C"
