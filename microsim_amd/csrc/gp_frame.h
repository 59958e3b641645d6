/*
 * gp_frame.h -- non-isothermal runs (temperature gradient along y) and the moving frame.
 *
 * With ISOTHERMAL = 0 the serial driver keeps a temperature per cell that depends on y only: after every step it
 * sets T(y) = temp_bottom + y GRADIENT on every allocated cell (apply_temperature_gradientY,
 * GPs/functions/Temperature_gradient.h:4-29, GPs/microsim_gp.c:108-113), and the boundary copies of the next step
 * copy T like phi and mu (GPs/solverloop/boundary_serial.h:182-197). The phase sweep reads T as it stands, the
 * concentration sweep after the copies. T enters c(mu) and Mu(c) through Beq + dBbdT (T - Teq) and the free
 * energy through B(T), C(T), which init_propertymatrices(T) sets per cell (GPs/functions/functionF_01.h:5-230).
 * Here T is one value per allocated row (trow), and each step forms the coefficients of every row for the two
 * sweeps (k_thermo); the kernels read them through View<..., TG = true>.
 *
 * Shift = 1 (GPs/microsim_gp.c:115-126): every 100 steps the largest y at which the liquid crosses 1/2 upwards
 * (check_SHIFT, GPs/functions/shift.h:50-66, taken over the rows of the concentration sweep) is compared with
 * Shiftj, and when it is larger every allocated column moves down by the difference (apply_shiftY, shift.h:4-49):
 * phi, mu and T of row y take those of row y + s, and the top s rows become liquid at Mu(cfill, Tfill) (their T
 * stays).
 */
#pragma once

#include "gp_common.h"
#include "gp_model.h"

namespace gpk {

/* what k_thermo needs of the parameters (FP64, passed by value) */
struct ThermoParams {
  int32_t N, NC;
  int32_t hy[6], srcT[6]; /* halo columns and the columns their T copies (itself when that side copies nothing) */
  double Teq;
  double Beq[P][NCM], dBbdT[P][NCM], A[P][NCM][NCM];
  double ceq_sol[P][NCM], ceq_liq[P][NCM], delta_c[P][NCM], delta_T_sol[P], delta_T_liq[P];
};

/* the coefficients of one row at temperature T, with the reference's operations: Beq + dBbdT (T - Teq)
 * (functionF_01.h:19-37), function_B and function_C (functionF_01.h:62-104) */
template <typename R>
__device__ __forceinline__ void thermo_row(const ThermoParams &tp, double T, RowThermo<R> &o) {
  const int N = tp.N, NC = tp.NC, L = N - 1;
  for (int a = 0; a < N; a++)
    for (int i = 0; i < NC; i++) o.bmu[a][i] = (R)(tp.Beq[a][i] + tp.dBbdT[a][i] * (T - tp.Teq));
  for (int a = 0; a < N; a++) {
    double cl[NCM], cs[NCM];
    if (a != L)
      for (int k = 0; k < NC; k++) {
        cl[k] = tp.ceq_liq[a][k] - (tp.delta_c[a][k]) * (tp.Teq - T) / (tp.delta_T_liq[a]);
        cs[k] = tp.ceq_sol[a][k] - (tp.delta_c[a][k]) * (tp.Teq - T) / (tp.delta_T_sol[a]);
      }
    for (int i = 0; i < NC; i++) {
      double Bai = 0.0;
      if (a != L) {
        double sc = 0.0;
        for (int k = 0; k < NC; k++)
          if (k != i) sc += tp.A[L][k][i] * cl[k] - tp.A[a][k][i] * cs[k];
        Bai = (2.0 * (tp.A[L][i][i] * cl[i] - tp.A[a][i][i] * cs[i]) + sc);
      }
      o.b[a][i] = (R)Bai;
    }
    double Cv = 0.0;
    if (a != L)
      for (int i = 0; i < NC; i++)
        for (int j = 0; j < NC; j++)
          if (i <= j) Cv += (tp.A[a][i][j] * cs[i] * cs[j] - tp.A[L][i][j] * cl[i] * cl[j]);
    o.c[a] = (R)Cv;
  }
  o.T = (R)T;
}

/* coefficients of every allocated row for the phase sweep (T as it stands) and the concentration sweep (T after
 * the boundary copies) */
template <typename R>
__global__ __launch_bounds__(256) void k_thermo(ThermoParams tp, const double *__restrict__ trow,
                                                RowThermo<R> *__restrict__ thp, RowThermo<R> *__restrict__ thc,
                                                int64_t ry) {
  const int64_t y = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (y >= ry) return;
  int64_t ys = y;
  for (int k = 0; k < 6; k++)
    if (tp.hy[k] == y) ys = tp.srcT[k];
  RowThermo<R> o;
  thermo_row<R>(tp, trow[y], o);
  thp[y] = o;
  thermo_row<R>(tp, trow[ys], o);
  thc[y] = o;
}

/* T(y) = temp_bottom + y GRADIENT on every allocated row */
static __global__ __launch_bounds__(256) void k_trow_set(double *__restrict__ trow, int64_t ry, double bottom,
                                                         double grad) {
  const int64_t y = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (y < ry) trow[y] = bottom + (double)y * grad;
}

/* largest allocated y in [1, rows_y - 1] of the interior rows x where phi_L(y-1) - (1 - phi_L(y-1)) < 0 and
 * phi_L(y) - (1 - phi_L(y)) > 0 (check_SHIFT), into *pos (0 when there is none; *pos starts at 0) */
template <typename R>
__global__ __launch_bounds__(256) void k_front(const R *__restrict__ phiL, int *pos, int64_t ps, int64_t mx,
                                               int64_t ry) {
  const int64_t y = 1 + (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
  const int64_t x = 3 + (int64_t)blockIdx.y * 4 + (threadIdx.x >> 6);
  bool hit = false;
  if (x <= mx + 2 && y < ry) {
    const int64_t c = x * ps + y + kYOff;
    const R a = phiL[c - 1], b = phiL[c];
    hit = (a - (R(1.0) - a) < R(0.0)) && (b - (R(1.0) - b) > R(0.0));
  }
  if (hit) atomicMax(pos, (int)y);
}

/* values of the rows a shift adds, per component */
struct FillValues {
  double v[P];
};

/* the moving frame on ncomp planes of one field: dst(x, y) = src(x, y + s) for y <= rows_y - 1 - s, the fill value
 * of that component above (apply_shiftY over every allocated cell) */
template <typename R>
__global__ __launch_bounds__(256) void k_shift(const R *__restrict__ src, R *__restrict__ dst, int ncomp,
                                               FillValues fill, int64_t s, int64_t fs, int64_t ps, int64_t rx,
                                               int64_t ry) {
  const int64_t y = (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
  const int64_t x = (int64_t)blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x >= rx || y >= ry) return;
  const int64_t c = x * ps + y + kYOff;
  for (int a = 0; a < ncomp; a++) dst[a * fs + c] = y + s <= ry - 1 ? src[a * fs + c + s] : (R)fill.v[a];
}

}  // namespace gpk
