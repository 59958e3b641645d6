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
 * pure-phase step and k_bulk, so the results are bit-identical. A tile whose staged codes are all one phase takes
 * that phase as a scalar and skips the per-cell classification. phi of a bulk cell is e_b; when the output buffer
 * already holds e_b there (the cell was purely b two states back) it is not rewritten.
 * k_cbulk also lists the interface cells of its tile (LDS, then one atomic per wavefront into a global list);
 * k_ilist runs the general per-cell update (gp_model.h: jat_x, jat_y, mu_update -- bit-identical to the
 * reference by construction) on the listed cells in a grid-stride loop, so that every lane of every wavefront
 * works on an interface cell (a march row has about one in five) and no wavefront holds more than its share.
 * Every interior cell is written by exactly one of the two. They read the phase codes of the current state and
 * write those of the new one (a second buffer). k_halo then fills the new state's halo from its interior.
 */
#pragma once

#include "gp_march.h"

namespace gpk {

#ifndef GP_CELL_SEG_MAX
#define GP_CELL_SEG_MAX 16
#endif
constexpr int kCellSegMax = GP_CELL_SEG_MAX;   /* longest x-segment the kernels stage (LDS sized for it) */
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
  bool out_lane;
  __device__ CellTile(int w, int lane, int64_t mx, int64_t my, int seg_len, int nstrips) {
    const int strip = w % nstrips, seg = w / nstrips;
    const int64_t ry = my + 6;
    y = 3 + (int64_t)strip * kMarchOut - 2 + lane;
    yl = y < ry - 1 ? y : ry - 1;
    xa = 3 + (int64_t)seg * seg_len;
    xb = (xa + seg_len - 1 < mx + 2) ? xa + seg_len - 1 : mx + 2;
    out_lane = lane >= 2 && lane <= 61 && y <= my + 2;
  }
};

/* a wave-uniform value read as a scalar: a per-lane choice among such values then stays a register select */
__device__ __forceinline__ float scalar_val(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ __forceinline__ double scalar_val(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffll));
  const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

/* FP32 bulk cells take the closed form through Model::Kbulk; FP64 keeps the reference's operations (bit-exact) */
template <typename R>
constexpr bool kLeanBulk = sizeof(R) < sizeof(double);

/* the bulk update of one cell in phase b. UNIFORM: b is the same on every lane (a scalar); every lane computes,
 * only lanes with `store` write, the mu statistics take every lane (the caller resets the lanes without a cell
 * once per tile) and the phi extrema (e_b on every cell) are left to the caller, once per tile. Otherwise the
 * caller calls it on the bulk lanes only, and it folds the full statistics. */
template <int N, int NC, typename R, bool P2, bool STATS, bool UNIFORM, int NSA, bool C = false>
__device__ __forceinline__ void cbulk_cell(const Model<R> &m, int b, int64_t c, const R *m0,
                                           const R *m1, const R *m2, const R *mud, const R *muu, R divjat0,
                                           R *__restrict__ phi_o, R *__restrict__ mu_o,
                                           int8_t *__restrict__ codes_out, int64_t fs, bool kept, bool store,
                                           bool live, R (&st)[NSA], unsigned &seen,
                                           const int8_t *q1 = nullptr, float *__restrict__ plo_o = nullptr,
                                           int8_t *__restrict__ mlo_o = nullptr) {
  using Vw = View<R, P2, false, kFastOf<R>>;
  R pe[N];
#pragma unroll
  for (int a = 0; a < N; a++) pe[a] = a == b ? R(1.0) : R(0.0);
  R deltamu[NC], munew[NC];
  if constexpr (kLeanBulk<R>) {
    /* FP32: the pure-phase update in closed form, dmu = Kbulk[b] . lap(mu) (equal in exact arithmetic to the
     * flux divergence through Mpure[b] followed by Mu_lin; divjat0 is a signed zero). The second differences
     * are formed from first differences, so their rounding is relative to the increment, as in the flux form. */
    R lap[NC];
#pragma unroll
    for (int l = 0; l < NC; l++) {
      const R gx = (m2[l] - m1[l]) - (m1[l] - m0[l]);
      const R gy = (mud[l] - m1[l]) - (m1[l] - muu[l]);
      lap[l] = __builtin_fmaf(gx, m.lap_x, gy * m.lap_y);
    }
#pragma unroll
    for (int k = 0; k < NC; k++) {
      R s = R(0.0);
#pragma unroll
      for (int l = 0; l < NC; l++) {
        R kk;
        if (UNIFORM) {
          kk = m.Kbulk[b][k][l];
        } else {
          /* per-lane phase: pick among the phases' scalar values (kept scalar, so that the choice is a register
           * select and not a per-lane load issued just before its use) */
          kk = scalar_val(m.Kbulk[0][k][l]);
#pragma unroll
          for (int a = 1; a < N; a++)
            if (b == a) kk = scalar_val(m.Kbulk[a][k][l]);
        }
        s = l == 0 ? kk * lap[0] : __builtin_fmaf(kk, lap[l], s);
      }
      deltamu[k] = s;
      munew[k] = m1[k] + s;
    }
  } else {
  R Mb[NC][NC];
#pragma unroll
  for (int k = 0; k < NC; k++)
#pragma unroll
    for (int l = 0; l < NC; l++) {
      if (UNIFORM) {
        Mb[k][l] = m.Mpure[b][k][l];
      } else {
        Mb[k][l] = scalar_val(m.Mpure[0][k][l]);
#pragma unroll
        for (int a = 1; a < N; a++)
          if (b == a) Mb[k][l] = scalar_val(m.Mpure[a][k][l]);
      }
    }
  R divflux[NC];
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
  }
  int8_t qn[NC];
  if constexpr (C) {
    /* compensated mode: the increment is added to the decoded value (hi plus its extension) and the sum is
     * re-encoded, so the bits of the increment below the FP32 part are kept */
#pragma unroll
    for (int k = 0; k < NC; k++) {
      float hn;
      c32_encode(c32_decode((float)m1[k], q1[k]) + (double)deltamu[k], hn, qn[k]);
      munew[k] = (R)hn;
    }
  }
  if (store) {
    const int8_t code = (int8_t)b;
    /* kept: the output buffers hold the state before the current one and its codes; a cell that was purely
     * phase b there already holds phi = e_b, so only mu is written */
    if (!kept) {
#pragma unroll
      for (int a = 0; a < N; a++) phi_o[a * fs + c] = pe[a];
      if constexpr (C) {
#pragma unroll
        for (int a = 0; a < N; a++) plo_o[a * fs + c] = 0.0f;
      }
      codes_out[c] = code;
    }
#pragma unroll
    for (int k = 0; k < NC; k++) mu_o[k * fs + c] = munew[k];
    if constexpr (C) {
#pragma unroll
      for (int k = 0; k < NC; k++) mlo_o[k * fs + c] = qn[k];
    }
  }
  if (STATS && UNIFORM && live) {
#pragma unroll
    for (int k = 0; k < NC; k++) {
      st[2 * N + k] = fmax(st[2 * N + k], munew[k]);
      st[2 * N + NC + k] = fmin(st[2 * N + NC + k], munew[k]);
      st[3 * N + 2 * NC + k] += deltamu[k] * deltamu[k];
      if (!isfinite(munew[k])) st[NSA - 1] += R(1.0);
    }
  } else if (STATS) {
    /* stats_add with delta-phi = 0 and phi = e_b: the squared increments of phi add +0 to sums that start at +0
     * and e_b is finite, so only the extrema of phi and the mu terms remain; the extrema of phi over cells that are
     * each some e_b follow from the set of phases seen (folded once, bulk_seen_fold) */
    seen |= 1u << b;
#pragma unroll
    for (int k = 0; k < NC; k++) {
      st[2 * N + k] = fmax(st[2 * N + k], munew[k]);
      st[2 * N + NC + k] = fmin(st[2 * N + NC + k], munew[k]);
      st[3 * N + 2 * NC + k] += deltamu[k] * deltamu[k];
      if (!isfinite(munew[k])) st[NSA - 1] += R(1.0);
    }
  }
}

/* the phi extrema of a lane's bulk cells of the phases in `seen` (bit b: a cell equal to e_b) into st: over such
 * cells phi_a is 1 where b = a and 0 elsewhere */
template <int N, typename R, int NSA>
__device__ __forceinline__ void bulk_seen_fold(R (&st)[NSA], unsigned seen) {
#pragma unroll
  for (int a = 0; a < N; a++) {
    const bool own = (seen >> a) & 1u, other = (seen & ~(1u << a)) != 0u;
    if (own) st[a] = fmax(st[a], R(1.0));
    else if (other) st[a] = fmax(st[a], R(0.0));
    if (other) st[N + a] = fmin(st[N + a], R(0.0));
    else if (own) st[N + a] = fmin(st[N + a], R(1.0));
  }
}

/* mu rows in flight ahead of the row after the current one, per precision (register budget: FP64 rows take twice
 * the registers); kCellRing = depth + 2 register slots hold rows x-1, x, x+1 and those in flight, and divides the
 * tile's 16 rows */
#ifndef GP_CELL_DEPTH_F32
#define GP_CELL_DEPTH_F32 2
#endif
#ifndef GP_CELL_DEPTH_F64
#define GP_CELL_DEPTH_F64 2
#endif
template <typename R>
constexpr int kCellBulkDepth = sizeof(R) == 8 ? GP_CELL_DEPTH_F64 : GP_CELL_DEPTH_F32;
template <typename R>
constexpr int kCellRing = kCellBulkDepth<R> + 2;

/* k_cbulk keeps its row offsets in 32 bits. The largest it forms is the refill offset of the last tile: row
 * x + D + 1 for a row x <= xb + RG - 1 of the unrolled ring (rows past xb run on clamped rows), i.e. below
 * (xb + RG + D + 1) ps with xb <= mesh_x + 2, before it is clamped to row xb + 1; the in-row offset adds less than
 * ps. A grid takes the cells path only when all of these fit: (rows_x + kCellSlack) * ps <= INT32_MAX. */
constexpr int kCellSlack = 32;
static_assert(kCellRing<float> + kCellBulkDepth<float> + 1 <= kCellSlack - 4 &&
                  kCellRing<double> + kCellBulkDepth<double> + 1 <= kCellSlack - 4,
              "the cells path's 32-bit offset bound needs a larger slack");
__host__ __device__ constexpr bool cells_fit(int64_t rows_x, int64_t ps) {
  return (rows_x + kCellSlack) * ps <= (int64_t)0x7fffffff;
}

/* Delta-phi candidates on the ring x or y in {2, mesh+3} around the interior (k_cbulk's tiles list the interior
 * ones): ring cell i (one per thread of k_cbulk's extra blocks), appended to the same list. */
__device__ __forceinline__ void dring_cell(int64_t i, const int8_t *__restrict__ codes, int32_t *__restrict__ clist,
                                           int *ccount, int64_t ps, int64_t mx, int64_t my) {
  const int lane = threadIdx.x & 63;
  const int64_t nxr = my + 2, nyr = mx; /* rows x = 2, mx+3 over y in [2, my+3]; columns y = 2, my+3 over x in [3, mx+2] */
  int64_t x = -1, y = -1;
  if (i < 2 * nxr) {
    x = (i < nxr) ? 2 : mx + 3;
    y = 2 + (i % nxr);
  } else if (i < 2 * nxr + 2 * nyr) {
    const int64_t k = i - 2 * nxr;
    y = (k < nyr) ? 2 : my + 3;
    x = 3 + (k % nyr);
  }
  bool cand = false;
  const int64_t c = x * ps + y + kYOff;
  if (x >= 0) {
    const int8_t k0 = codes[c];
    cand = k0 < 0 || codes[c - ps] != k0 || codes[c + ps] != k0 || codes[c - 1] != k0 || codes[c + 1] != k0;
  }
  const uint64_t mask = __ballot(cand);
  if (!mask) return;
  const int leader = __ffsll((long long)mask) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(ccount, __popcll(mask));
  base = __shfl(base, leader);
  if (cand) clist[base + __popcll(mask & ((1ull << lane) - 1ull))] = (int32_t)c;
}

/* a floor of waves per SIMD for k_cbulk's register allocation: the FP32 kernels of up to two mu fields fit seven
 * (the LDS bound) with at most one spilled register; the others keep their natural allocation */
template <typename R, int NC>
constexpr int kCbulkWaves = (sizeof(R) == 4 && NC <= 2) ? 7 : 1;

template <int N, int NC, typename R, bool P2, bool STATS, bool C = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kCbulkWaves<R, NC>)))
void k_cbulk(const Model<R> *__restrict__ mp, const R *__restrict__ mu,
                                               const int8_t *__restrict__ codes_in, R *__restrict__ phi_o,
                                               R *__restrict__ mu_o, int8_t *__restrict__ codes_out,
                                               double *__restrict__ partials, int64_t pstride, int64_t fs, int64_t ps,
                                               int64_t mx, int64_t my, int seg_len, int nstrips, int nwaves,
                                               int keep, int32_t *__restrict__ ilist, int32_t *__restrict__ clist,
                                               int *counts, int *counts_next, int tile_blocks,
                                               const int8_t *__restrict__ mlo, float *__restrict__ plo_o,
                                               int8_t *__restrict__ mlo_o) {
  using Vw = View<R, P2, false, kFastOf<R>>;
  constexpr int NS = 3 * N + 3 * NC + 1;
  constexpr int NSA = STATS ? NS : 1;
  if ((int)blockIdx.x >= tile_blocks) {
    /* the blocks past the tiles list the delta-phi candidates of the ring x or y in {2, mesh+3} around the
     * interior (one thread per ring cell; the tiles list the interior ones) */
    dring_cell((int64_t)(blockIdx.x - tile_blocks) * blockDim.x + threadIdx.x, codes_in, clist, counts + 1, ps, mx,
               my);
    return;
  }
  __shared__ int8_t cs_all[4][kCellRows * 64];
  __shared__ uint16_t il_all[4][kCellSegMax * 64], cl_all[4][kCellSegMax * 64];
  __shared__ int wcnt[2][4], wbase[2][4];
  uint16_t *il = il_all[threadIdx.x >> 6], *cl = cl_all[threadIdx.x >> 6];
  int icnt = 0, ccnt = 0;
  const Model<R> &m = *mp;
  const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int w = blockIdx.x * 4 + wv; /* wave-uniform: the tile's rows and the row loop are scalar */
  int8_t *cs = cs_all[wv];
  R st[NSA];
  if (STATS) stats_init<N, NC>(st);
  unsigned seen = 0u; /* phases of the lane's bulk cells in mixed tiles (their phi extrema, folded at the end) */
  const CellTile t(w, lane, mx, my, seg_len, nstrips);
  const int64_t c0 = t.yl + kYOff;
  /* the next step's list lengths start at zero (the kernels that read them last ran before this one) */
  if (blockIdx.x == 0 && threadIdx.x < 2) counts_next[threadIdx.x] = 0;
  /* mu rows in a ring of kCellRing register slots, row r in slot (r - xa) mod kCellRing, so that no slot is
   * ever copied: the row being updated reads rows x-1, x, x+1 where they were loaded, and the slot of x-1 then
   * takes row x + kCellBulkDepth + 1. The phase codes of the output buffers (for `kept`) ride the same ring,
   * kCellBulkDepth rows ahead. A loaded value is first read kCellBulkDepth rows after its load, and no wait
   * in between covers a newer load (the counter of outstanding memory operations is in order). */
  constexpr int RG = kCellRing<R>, D = kCellBulkDepth<R>;
  R ring[RG][NC];
  int8_t qring[RG][NC]; /* compensated mode: the mu extensions of the ring's rows */
  uint8_t kc[RG]; /* output codes as bytes (zero-extending loads, no sign extension after the load) */
  const uint8_t *codes_ou = (const uint8_t *)codes_out;
  if (w < nwaves) {
    /* the tile's loads all issue before the first wait: its phase codes (rows xa-2 .. xa+kCellRows-3, clamped
     * to xb+2), the first mu rows and output codes of the ring (every load is issued on every path; rows past the
     * tile's end are clamped to its last row + 1, which is inside the halo) */
    int8_t cv[kCellRows];
#pragma unroll
    for (int r = 0; r < kCellRows; r++) {
      const int64_t xr = t.xa - 2 + r <= t.xb + 2 ? t.xa - 2 + r : t.xb + 2;
      cv[r] = codes_in[xr * ps + c0];
    }
#pragma unroll
    for (int d = -1; d <= D; d++) {
      const int64_t xr = t.xa + d <= t.xb + 1 ? t.xa + d : t.xb + 1;
#pragma unroll
      for (int k = 0; k < NC; k++) ring[(d + RG) % RG][k] = mu[k * fs + xr * ps + c0];
      if constexpr (C) {
#pragma unroll
        for (int k = 0; k < NC; k++) qring[(d + RG) % RG][k] = mlo[k * fs + xr * ps + c0];
      }
    }
#pragma unroll
    for (int d = 0; d < D; d++) kc[d] = codes_ou[(t.xa + d <= t.xb ? t.xa + d : t.xb) * ps + c0];
#pragma unroll
    for (int r = 0; r < kCellRows; r++) cs[r * 64 + lane] = cv[r];
    /* the prologue's loads complete here, once per tile: the loop's first row then waits on nothing, and the
     * wait counts inside the loop follow from the loop body alone (vmcnt 0, expcnt and lgkmcnt unconstrained) */
    __builtin_amdgcn_s_waitcnt(0x0f70);
  }
  __syncthreads();
  if (w < nwaves) {
    /* a tile whose staged codes are all one phase b has only bulk cells of phase b: no per-cell test */
    int tb = -1;
    {
      const int u = cs[lane];
      bool same = u >= 0;
      const int nr = (int)(t.xb - t.xa) + 5;
      for (int r = 1; r < nr; r++) same = same && cs[r * 64 + lane] == u;
      const int u0 = __builtin_amdgcn_readfirstlane(u);
      if (__all(same && u == u0)) tb = u0;
    }
    const R zj = R(0.0) * m.jat_fac; /* (j - j_prev) * jat_fac with both currents +0 */
    R divjat0 = divx<Vw>(m, zj);
    divjat0 += divy<Vw>(m, zj);
    /* the row loop, instantiated twice so that the tile-uniform form is a loop of its own (no per-row test) */
    /* row offsets in 32 bits (the cells path runs on grids of fewer than 2^31 cells), kept incrementally */
    const int pso = (int)ps, c0i = (int)c0, cyi = (int)(t.y + kYOff);
    const int xa = (int)t.xa, xb = (int)t.xb, mxi = (int)mx;
    const int xlast = (xb + 1) * pso;
    auto rows = [&](auto uniform) {
      constexpr bool UNI = decltype(uniform)::value;
      int xoff = xa * pso;
      for (int x0 = xa; x0 <= xb; x0 += RG) {
#pragma unroll
        for (int j = 0; j < RG; j++) {
          /* rows past the tile's last one (at the grid's end) run on clamped rows and write nothing: the loop
           * has one exit and every path through its body issues every load */
          const int x = x0 + j;
          const bool live = x <= xb;
          const int64_t c = (int64_t)xoff + cyi;
          const R *m0 = ring[(j + RG - 1) % RG], *m1 = ring[j], *m2 = ring[(j + 1) % RG];
          R mud[NC], muu[NC];
#pragma unroll
          for (int k = 0; k < NC; k++) {
            mud[k] = lane_dn(m1[k]);
            muu[k] = lane_up(m1[k]);
          }
          if (UNI) {
            cbulk_cell<N, NC, R, P2, STATS, true, NSA, C>(m, tb, c, m0, m1, m2, mud, muu, divjat0, phi_o, mu_o,
                                                          codes_out, fs, keep && kc[j] == (uint8_t)tb,
                                                          live && t.out_lane, live, st, seen, qring[j], plo_o, mlo_o);
          } else if (live) {
            const int bl = t.out_lane ? bulk_of(cs, x - xa + 2, lane) : -1;
            {
              /* the row's interface cells (for k_ilist) and delta-phi candidates (for k_dlist) */
              const int r = x - xa + 2;
              const int8_t k0 = cs[r * 64 + lane];
              const bool cand = t.out_lane && (k0 < 0 || cs[(r - 1) * 64 + lane] != k0 ||
                                               cs[(r + 1) * 64 + lane] != k0 || cs[r * 64 + lane - 1] != k0 ||
                                               cs[r * 64 + lane + 1] != k0);
              const bool iface = t.out_lane && bl < 0;
              const uint16_t e = (uint16_t)((x - xa) * 64 + lane);
              const uint64_t lt = (1ull << lane) - 1ull;
              const uint64_t mi = __ballot(iface), mc = __ballot(cand);
              if (iface) il[icnt + __popcll(mi & lt)] = e;
              if (cand) cl[ccnt + __popcll(mc & lt)] = e;
              icnt += __popcll(mi);
              ccnt += __popcll(mc);
            }
            if (bl >= 0)
              cbulk_cell<N, NC, R, P2, STATS, false, NSA, C>(m, bl, c, m0, m1, m2, mud, muu, divjat0, phi_o, mu_o,
                                                             codes_out, fs, keep && kc[j] == (uint8_t)bl, true, true,
                                                             st, seen, qring[j], plo_o, mlo_o);
          }
          /* refill: row x + D + 1 into the slot of row x - 1, the output codes of row x + D */
          const int xn = min(xoff + (D + 1) * pso, xlast);
#pragma unroll
          for (int k = 0; k < NC; k++) ring[(j + RG - 1) % RG][k] = mu[k * fs + (int64_t)(xn + c0i)];
          if constexpr (C) {
#pragma unroll
            for (int k = 0; k < NC; k++) qring[(j + RG - 1) % RG][k] = mlo[k * fs + (int64_t)(xn + c0i)];
          }
          kc[(j + D) % RG] = codes_ou[(int64_t)(xn - pso + c0i)];
          xoff += pso;
        }
      }
    };
    if (tb >= 0)
      rows(std::true_type{});
    else
      rows(std::false_type{});
    if (STATS && tb >= 0) {
      /* the tile-uniform rows folded the mu statistics of every lane: keep those of the tile's cells, and add
       * the phi extrema of its cells (e_tb on each) once */
      if (!t.out_lane) {
        stats_init<N, NC>(st);
      } else {
#pragma unroll
        for (int a = 0; a < N; a++) {
          st[a] = fmax(st[a], a == tb ? R(1.0) : R(0.0));
          st[N + a] = fmin(st[N + a], a == tb ? R(1.0) : R(0.0));
        }
      }
    }
  }
  /* one atomic per block and list: reserve the block's interface cells and candidates in the global lists */
  if (lane == 0) {
    wcnt[0][wv] = icnt;
    wcnt[1][wv] = ccnt;
  }
  __syncthreads();
  if (threadIdx.x == 0)
    for (int l = 0; l < 2; l++) {
      int tot = 0;
      for (int v = 0; v < 4; v++) {
        wbase[l][v] = tot;
        tot += wcnt[l][v];
      }
      const int b = tot > 0 ? atomicAdd(&counts[l], tot) : 0;
      for (int v = 0; v < 4; v++) wbase[l][v] += b;
    }
  __syncthreads();
  for (int i = lane; i < icnt; i += 64) {
    const int e = il[i];
    ilist[wbase[0][wv] + i] = (int32_t)((t.xa + (e >> 6)) * ps + (t.y - lane + (e & 63)) + kYOff);
  }
  for (int i = lane; i < ccnt; i += 64) {
    const int e = cl[i];
    clist[wbase[1][wv] + i] = (int32_t)((t.xa + (e >> 6)) * ps + (t.y - lane + (e & 63)) + kYOff);
  }
  if (!STATS) return;
  bulk_seen_fold<N>(st, seen);
  stats_partials<N, NC, true>(st, partials, pstride, w, nwaves, lane);
}


/* The halo of the new state, after k_cbulk and k_ilist wrote every interior cell: each halo cell takes the value
 * of its proxy, as write_images scatters it (apply_boundary_conditions as a gather). Halo row X = hx[k] has the
 * proxy row sx[f][k] when that is an interior row and at most one k' < k shares it (write_images takes the first
 * two images of a row); an interior row is its own (likewise for columns). Field group f = 0 holds phi, its low
 * parts and the phase codes, f = 1 mu and its extensions. A halo cell without a proxy is left as it is. */
__device__ __forceinline__ int64_t halo_source(const int32_t *src, const int32_t *h, int64_t X, int64_t lo,
                                               int64_t hi) {
  if (X >= lo && X <= hi) return X;
  int64_t s = -1;
#pragma unroll
  for (int k = 0; k < 6; k++) {
    if (h[k] != X) continue;
    int before = 0;
    for (int j = 0; j < k; j++) before += src[j] == src[k];
    if (src[k] >= lo && src[k] <= hi && before < 2) s = src[k];
  }
  return s;
}

/* planes of the state the halo pass copies: the phase codes, phi, (C: phi's low parts), mu, (C: mu's extensions) */
template <int N, int NC, bool C>
constexpr int kHaloPlanes = 1 + N + NC + (C ? N + NC : 0);

/* halo cell j (the six halo rows over the full width, then the six halo cells of each interior row) of plane p */
template <int N, int NC, typename R, bool C>
__device__ __forceinline__ void halo_cell(int64_t j, int p, R *__restrict__ phi, R *__restrict__ mu,
                                          int8_t *__restrict__ codes, float *__restrict__ plo,
                                          int8_t *__restrict__ mlo, const Maps &maps, int64_t fs, int64_t ps,
                                          int64_t mx, int64_t my) {
  const int64_t ry = my + 6, nrow = 6 * ry, ncol = 6 * mx;
  if (j >= nrow + ncol) return;
  /* the index arithmetic in 32 bits (the halo has fewer than 2^31 cells; 64-bit division is a long sequence) */
  int64_t X, Y;
  if (j < nrow) {
    const uint32_t iu = (uint32_t)j, ryu = (uint32_t)ry, r = iu / ryu;
    X = r < 3u ? (int64_t)r : mx + r;
    Y = iu - r * ryu;
  } else {
    /* six consecutive threads take the six halo cells of one interior row (two runs of three adjacent elements) */
    const uint32_t k = (uint32_t)(j - nrow), row = k / 6u, q = k - 6u * row;
    Y = q < 3u ? (int64_t)q : my + q;
    X = 3 + (int64_t)row;
  }
  constexpr int P0 = 1 + N + (C ? N : 0); /* planes of field group 0 (codes, phi, plo); the rest are group 1 */
  const int f = p < P0 ? 0 : 1;
  const int64_t sx = halo_source(maps.sx[f], maps.hx, X, 3, mx + 2);
  const int64_t sy = halo_source(maps.sy[f], maps.hy, Y, 3, my + 2);
  if (sx < 0 || sy < 0) return;
  const int64_t c = X * ps + Y + kYOff, s = sx * ps + sy + kYOff;
  if (p == 0) {
    codes[c] = codes[s];
  } else if (p <= N) {
    phi[(p - 1) * fs + c] = phi[(p - 1) * fs + s];
  } else if (C && p < P0) {
    plo[(p - 1 - N) * fs + c] = plo[(p - 1 - N) * fs + s];
  } else if (p < P0 + NC) {
    mu[(p - P0) * fs + c] = mu[(p - P0) * fs + s];
  } else if (C) {
    mlo[(p - P0 - NC) * fs + c] = mlo[(p - P0 - NC) * fs + s];
  }
}

/* number of halo cells halo_cell covers */
__host__ __device__ constexpr int64_t halo_cells(int64_t mx, int64_t my) { return 6 * (my + 6) + 6 * mx; }

/* one thread per halo cell and plane (blockIdx.y is the plane) */
template <int N, int NC, typename R, bool C>
__global__ __launch_bounds__(256) void k_halo(R *__restrict__ phi, R *__restrict__ mu, int8_t *__restrict__ codes,
                                              float *__restrict__ plo, int8_t *__restrict__ mlo, Maps maps,
                                              int64_t fs, int64_t ps, int64_t mx, int64_t my) {
  halo_cell<N, NC, R, C>((int64_t)blockIdx.x * blockDim.x + threadIdx.x, (int)blockIdx.y, phi, mu, codes, plo, mlo,
                         maps, fs, ps, mx, my);
}

/* The general update of the interface cells listed by k_cbulk (any order): a grid-stride loop of full
 * wavefronts. Each cell gathers its stencil from memory in one batch: phi on the 3x3 neighbourhood, mu and
 * delta-phi on the 5-point one, and the phase codes of its radius-2 diamond, which decide which of the five
 * delta-phi values are k_dphi's (candidates) and which are exactly 0. Statistics: one partial per wavefront of the
 * grid (the grid is fixed, the order of the cells within it is not). */
template <int N, int NC, typename R, bool P2, bool STATS>
__global__ __launch_bounds__(256) void k_ilist(const Model<R> *__restrict__ mp, const R *__restrict__ phi,
                                               const R *__restrict__ mu, const R *__restrict__ dphi,
                                               const int8_t *__restrict__ codes_in, R *__restrict__ phi_o,
                                               R *__restrict__ mu_o, int8_t *__restrict__ codes_out,
                                               double *__restrict__ partials, int64_t pstride, int64_t fs,
                                               int64_t ps, const int32_t *__restrict__ ilist,
                                               const int *__restrict__ icount) {
  constexpr int NS = 3 * N + 3 * NC + 1;
  constexpr int NSA = STATS ? NS : 1;
  const Model<R> &m = *mp;
  const int lane = threadIdx.x & 63;
  const int w = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int nw = gridDim.x * 4;
  R st[NSA];
  if (STATS) stats_init<N, NC>(st);
  const int n = *icount;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int64_t c = ilist[i];
    const int64_t y = (c - kYOff) % ps;
    /* codes of the diamond: slot (di, dj) for |di| + |dj| <= 2 */
    int8_t cd[5][5];
#pragma unroll
    for (int di = -2; di <= 2; di++)
#pragma unroll
      for (int dj = -2; dj <= 2; dj++)
        if ((di < 0 ? -di : di) + (dj < 0 ? -dj : dj) <= 2) cd[di + 2][dj + 2] = codes_in[c + di * ps + dj];
    RegStore<N, NC, R> rs;
    const int oi[5] = {0, 1, -1, 0, 0}, oj[5] = {0, 0, 0, 1, -1};
#pragma unroll
    for (int a = 0; a < N; a++)
#pragma unroll
      for (int di = 0; di < 3; di++)
#pragma unroll
        for (int dj = 0; dj < 3; dj++) rs.p[a][di][dj] = phi[a * fs + c + (di - 1) * ps + (dj - 1)];
#pragma unroll
    for (int o = 0; o < 5; o++) {
      const int64_t co = c + oi[o] * ps + oj[o];
#pragma unroll
      for (int k = 0; k < NC; k++) rs.u[k][o] = mu[k * fs + co];
      const int r = oi[o] + 2, q = oj[o] + 2;
      const int8_t k0 = cd[r][q];
      const bool cand = k0 < 0 || cd[r - 1][q] != k0 || cd[r + 1][q] != k0 || cd[r][q - 1] != k0 || cd[r][q + 1] != k0;
#pragma unroll
      for (int a = 0; a < N; a++) {
        const R v = dphi[a * fs + co];
        rs.d[a][o] = cand ? v : R(0.0);
      }
    }
    const RegView<N, NC, R, P2, kFastOf<R>> v{&rs, 0, 0};
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
    codes_out[c] = (int8_t)pure_phase<N>(pn);
    if (STATS) stats_add<N, NC>(st, pn, d, mn, dm);
  }
  if (!STATS) return;
  stats_partials<N, NC>(st, partials, pstride, w, nw, lane);
}

}  // namespace gpk
