/*
 * gp_common.h -- device layout constants and helpers shared by the step kernels.
 */
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gpk {

constexpr int kYOff = 61; /* element offset of y = 0 inside a row: y = 3 lands on a 64-element boundary */

/* halo index maps of the two fields (0 phi, 1 mu): source x/y of allocated x/y = 0,1,2, m+3, m+4, m+5 */
struct Maps {
  int32_t sx[2][6], sy[2][6];
  int32_t hx[6], hy[6];
};

/* An interior cell is copied to every halo cell whose proxy it is (apply_boundary_conditions as a scatter):
 * X in {x} u {halo x with source x}, Y likewise, all (X, Y) but (x, y). */
template <typename R>
__device__ __forceinline__ void write_images(R *__restrict__ base, int ncomp, const R *vals, int f, const Maps &mp,
                                             int64_t x, int64_t y, int64_t fs, int64_t ps) {
  int64_t X[3], Y[3];
  int nx = 1, ny = 1;
  X[0] = x;
  Y[0] = y;
  for (int k = 0; k < 6; k++) {
    if (mp.sx[f][k] == x && nx < 3) X[nx++] = mp.hx[k];
    if (mp.sy[f][k] == y && ny < 3) Y[ny++] = mp.hy[k];
  }
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++) {
      if (i == 0 && j == 0) continue;
      const int64_t c = X[i] * ps + Y[j] + kYOff;
      for (int a = 0; a < ncomp; a++) base[a * fs + c] = vals[a];
    }
}

/* Wavefront reductions of a double over the 64 lanes, returned to every lane. DPP steps keep the data in the
 * VALU (no LDS round trips): quad_perm [1,0,3,2] and [2,3,0,1], row_half_mirror and row_mirror reduce each
 * 16-lane row; row_bcast:15 (rows 1, 3) and row_bcast:31 (rows 2, 3) combine the rows, leaving the result in
 * lane 63. Lanes that a masked step does not write receive the operation's identity. */
namespace dpp {
constexpr int kQuad1032 = 0xB1, kQuad2301 = 0x4E, kRowHalfMirror = 0x141, kRowMirror = 0x140, kRowBcast15 = 0x142,
              kRowBcast31 = 0x143;
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double mov(double v, double identity) {
  const long long b = __double_as_longlong(v), id = __double_as_longlong(identity);
  const int lo = __builtin_amdgcn_update_dpp((int)(id & 0xffffffffll), (int)(b & 0xffffffffll), CTRL, ROWMASK, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(id >> 32), (int)(b >> 32), CTRL, ROWMASK, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double lane63(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), 63), hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <int CTRL, int ROWMASK>
__device__ __forceinline__ float mov(float v, float identity) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(__float_as_int(identity), __float_as_int(v), CTRL, ROWMASK, 0xf, false));
}
__device__ __forceinline__ float lane63(float v) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}
template <class T, class Op>
__device__ __forceinline__ T reduce(T v, T identity, Op op) {
  v = op(v, mov<kQuad1032, 0xf>(v, identity));
  v = op(v, mov<kQuad2301, 0xf>(v, identity));
  v = op(v, mov<kRowHalfMirror, 0xf>(v, identity));
  v = op(v, mov<kRowMirror, 0xf>(v, identity));
  v = op(v, mov<kRowBcast15, 0xa>(v, identity));
  v = op(v, mov<kRowBcast31, 0xc>(v, identity));
  return lane63(v);
}
}  // namespace dpp

__device__ __forceinline__ double wave_max(double v) {
  return dpp::reduce<double>(v, -INFINITY, [](double a, double b) { return fmax(a, b); });
}
__device__ __forceinline__ double wave_min(double v) {
  return dpp::reduce<double>(v, INFINITY, [](double a, double b) { return fmin(a, b); });
}
__device__ __forceinline__ double wave_sum(double v) {
  return dpp::reduce<double>(v, 0.0, [](double a, double b) { return a + b; });
}
/* extrema are exact in any precision: FP32 values take the single-register form */
__device__ __forceinline__ float wave_max(float v) {
  return dpp::reduce<float>(v, -INFINITY, [](float a, float b) { return fmaxf(a, b); });
}
__device__ __forceinline__ float wave_min(float v) {
  return dpp::reduce<float>(v, INFINITY, [](float a, float b) { return fminf(a, b); });
}

}  // namespace gpk
