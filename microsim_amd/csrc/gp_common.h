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

template <typename T>
__device__ __forceinline__ T wave_max(T v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}
template <typename T>
__device__ __forceinline__ T wave_min(T v) {
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o));
  return v;
}
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

}  // namespace gpk
