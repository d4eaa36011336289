// Device helpers of the learned proposal: bf16 conversion, input
// encoding and the von Mises-Fisher mixture (density and sampling).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "guide.h"

namespace sgr {

#define GD_PI 3.14159265358979323846f

// fp32 -> bf16, round to nearest even (inputs are finite).
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Feature F of the encoding of inputs U[6]: F < 6 the raw input; then
// for input i and frequency k, sin and cos of 2^k * pi * u_i; 78, 79
// are zero padding.
__device__ __forceinline__ float gd_feature(const float *u, int f) {
  if (f < GD_IN) return u[f];
  if (f >= GD_IN + GD_IN * GD_NFREQ * 2) return 0.0f;
  const int e = f - GD_IN, i = e / (GD_NFREQ * 2), r = e - i * (GD_NFREQ * 2);
  const int k = r >> 1;
  const float x = u[i] * (float)(1 << k) * GD_PI;
  return (r & 1) ? cosf(x) : sinf(x);
}

// The same feature with the hardware sine / cosine (argument in
// revolutions: 2^k pi u = 2^(k-1) u turns, |.| <= 32 turns for |u| <= 1);
// differs from gd_feature by about 1e-6, far below the bf16 rounding
// the features then get.
__device__ __forceinline__ float gd_feature_fast(const float *u, int f) {
  if (f < GD_IN) return u[f];
  if (f >= GD_IN + GD_IN * GD_NFREQ * 2) return 0.0f;
  const int e = f - GD_IN, i = e / (GD_NFREQ * 2), r = e - i * (GD_NFREQ * 2);
  const int k = r >> 1;
  const float turns = u[i] * (float)(1 << k) * 0.5f;
  return (r & 1) ? __builtin_amdgcn_cosf(turns) : __builtin_amdgcn_sinf(turns);
}

// vMF density on the unit sphere with mean MU and concentration KAPPA
// at cosine C = mu . w: kappa / (2 pi (1 - e^{-2 kappa})) e^{kappa (c - 1)}.
__device__ __forceinline__ float vmf_pdf(float kappa, float c) {
  if (kappa < 1e-4f) return 1.0f / (4.0f * GD_PI);
  return kappa / (2.0f * GD_PI * (1.0f - expf(-2.0f * kappa))) * expf(kappa * (c - 1.0f));
}

// Mixture density at world direction W of the 8 lobes at L (mu xyz,
// kappa, weight per lobe).
__device__ __forceinline__ float mix_pdf(const float *L, float wx, float wy, float wz) {
  float q = 0.0f;
  for (int k = 0; k < GD_LOBES; k++) {
    const float *o = L + 5 * k;
    const float c = o[0] * wx + o[1] * wy + o[2] * wz;
    q += o[4] * vmf_pdf(o[3], c);
  }
  return q;
}

// Sample the mixture: lobe by U0 over the weights, then the lobe with
// (U1, U2) by the stable inversion w = 1 + log(u1 + (1 - u1) e^{-2 kappa}) / kappa
// (Jakob, "Numerically stable sampling of the von Mises-Fisher
// distribution on S^2", 2012).
__device__ __forceinline__ void mix_sample(const float *L, float u0, float u1, float u2, float &wx,
                                           float &wy, float &wz) {
  int k = 0;
  float acc = 0.0f;
  for (; k < GD_LOBES - 1; k++) {
    acc += L[5 * k + 4];
    if (u0 < acc) break;
  }
  const float *o = L + 5 * k;
  const float kappa = o[3];
  float w;
  if (kappa < 1e-4f) w = 1.0f - 2.0f * u1;
  else w = 1.0f + logf(u1 + (1.0f - u1) * expf(-2.0f * kappa)) / kappa;
  w = fminf(fmaxf(w, -1.0f), 1.0f);
  const float sr = sqrtf(fmaxf(0.0f, 1.0f - w * w));
  const float ph = 2.0f * GD_PI * u2;
  const float lx = sr * cosf(ph), ly = sr * sinf(ph);
  // orthonormal basis around mu (Duff et al., "Building an orthonormal
  // basis, revisited", 2017)
  const float nx = o[0], ny = o[1], nz = o[2];
  const float sg = nz >= 0.0f ? 1.0f : -1.0f;
  const float a = -1.0f / (sg + nz), b = nx * ny * a;
  const float tx = 1.0f + sg * nx * nx * a, ty = sg * b, tz = -sg * nx;
  const float bx = b, by = sg + ny * ny * a, bz = -ny;
  wx = lx * tx + ly * bx + w * nx;
  wy = lx * ty + ly * by + w * ny;
  wz = lx * tz + ly * bz + w * nz;
}

}  // namespace sgr
