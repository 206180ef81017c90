// Device math shared by the Debye kernels (debye.hip, symgrad.hip): FP32
// sin/cos after an FP64 reduction, the two-float recurrence coefficient
// 2 cos(dtheta), pair geometry and the FP64 lane reduction through LDS.
#pragma once

#include "internal.h"

namespace pyiid {
namespace {

// sin and cos of |rf| <= pi/4: Taylor series, sin through r^9 and cos
// through r^10 (truncation < 2e-9).
__device__ __forceinline__ void sincos_poly(float rf, float& s, float& c) {
  const float r2 = rf * rf;
  float ps = fmaf(r2, 1.0f / 362880.0f, -1.0f / 5040.0f);
  ps = fmaf(r2, ps, 1.0f / 120.0f);
  ps = fmaf(r2, ps, -1.0f / 6.0f);
  s = fmaf(rf * r2, ps, rf);
  float pc = fmaf(r2, -1.0f / 3628800.0f, 1.0f / 40320.0f);
  pc = fmaf(r2, pc, -1.0f / 720.0f);
  pc = fmaf(r2, pc, 1.0f / 24.0f);
  pc = fmaf(r2, pc, -0.5f);
  c = fmaf(r2, pc, 1.0f);
}

// sin and cos of x (radians, x >= 0, |x| < 1e6).  Cody-Waite reduction by
// pi/2 in FP64 (two-term split, FMA), then Taylor series on |r| <= pi/4:
// sin through r^9 and cos through r^10 (truncation < 2e-9).
__device__ __forceinline__ void sincos_red(double x, float& s, float& c) {
  constexpr double kTwoOverPi = 0.63661977236758134308;
  constexpr double kPio2Hi = 1.5707963267948966192;
  constexpr double kPio2Lo = 6.123233995736766036e-17;
  const double n = rint(x * kTwoOverPi);
  double r = fma(-n, kPio2Hi, x);
  r = fma(-n, kPio2Lo, r);
  float sp, cp;
  sincos_poly(static_cast<float>(r), sp, cp);
  const int q = static_cast<int>(n) & 3;
  const bool swap = q & 1;
  float so = swap ? cp : sp;
  float co = swap ? sp : cp;
  s = (q & 2) ? -so : so;
  c = ((q + 1) & 2) ? -co : co;
}

// dtheta = k pi + d with k the nearest integer (|d| <= pi/2): the three-term
// recurrence y' = 2 cos(dtheta) y - y_prev amplifies rounding by about
// 1 / |sin(dtheta)| = 1 / sin|d|, so its safe range is set by |d|, near
// every multiple of pi, not only near 0.
// sin x for |x| <= pi/4 in FP64 (Taylor series through x^17, truncation
// below 1e-19).
__device__ __forceinline__ double sin_poly_f64(double x) {
  const double x2 = x * x;
  double p = fma(x2, -1.0 / 355687428096000.0, 1.0 / 1307674368000.0);
  p = fma(x2, p, -1.0 / 6227020800.0);
  p = fma(x2, p, 1.0 / 39916800.0);
  p = fma(x2, p, -1.0 / 362880.0);
  p = fma(x2, p, 1.0 / 5040.0);
  p = fma(x2, p, -1.0 / 120.0);
  p = fma(x2, p, 1.0 / 6.0);
  return fma(-x * x2, p, x);
}

__device__ __forceinline__ double pi_reduce(double dth, int& k_odd) {
  constexpr double kInvPi = 0.31830988618379067154;
  constexpr double kPiHi = 3.141592653589793116;
  constexpr double kPiLo = 1.2246467991473532072e-16;
  const double k = rint(dth * kInvPi);
  k_odd = static_cast<int>(static_cast<long long>(k) & 1);
  return fma(-k, kPiLo, fma(-k, kPiHi, dth));
}

// 2 cos(dtheta) as hi + lo (FP32 each, from FP64): the three-term
// recurrence y' = (hi + lo) y - y_prev then runs at the exact frequency.
// With hi alone the frequency error ~ulp / |sin(dtheta)| becomes a phase
// drift that grows with the step count.
// Also sin(dtheta) in FP32 from the same reduction: (-1)^k 2 sin(d/2) cos(d/2).
__device__ __forceinline__ void two_cos(double dth, float& hi, float& lo, float& s1) {
  int k_odd;
  const double d = pi_reduce(dth, k_odd);
  const double shd = sin_poly_f64(0.5 * d);
  const double k2 = (k_odd ? -1.0 : 1.0) * fma(-4.0 * shd, shd, 2.0);
  hi = static_cast<float>(k2);
  lo = static_cast<float>(k2 - static_cast<double>(hi));
  const float shf = static_cast<float>(shd);
  const float s = 2.0f * shf * __builtin_amdgcn_sqrtf(fmaf(-shf, shf, 1.0f));
  s1 = k_odd ? -s : s;
}

struct PairGeom {
  float inv_r;   // 1/r (0 for excluded pairs)
  float dpx, dpy, dpz;  // (q_i - q_j) / r^2
  double r;      // |q_i - q_j| (0 for excluded pairs)
};

// Geometry of the pair (pi, pj).  `keep` false zeroes the pair's weights so
// it contributes nothing (self pair, padding lanes, j >= i in F-only).
__device__ __forceinline__ PairGeom pair_geom(const AtomRec& pi, const AtomRec& pj, bool keep) {
  const double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
  const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
  keep = keep && (r2 > 0.0);
  // FP32 square root refined by one Newton step in FP64 (relative error ~1e-14).
  const float rf = __builtin_amdgcn_sqrtf(static_cast<float>(r2));
  const double rf64 = static_cast<double>(rf);
  const double rn = rf64 + fma(-rf64, rf64, r2) * static_cast<double>(0.5f * __builtin_amdgcn_rcpf(rf));
  PairGeom g;
  g.r = keep ? rn : 0.0;
  const float inv_r = keep ? __builtin_amdgcn_rcpf(static_cast<float>(rn)) : 0.0f;
  const float inv_r2 = inv_r * inv_r;
  g.inv_r = inv_r;
  g.dpx = static_cast<float>(dx) * inv_r2;
  g.dpy = static_cast<float>(dy) * inv_r2;
  g.dpz = static_cast<float>(dz) * inv_r2;
  return g;
}

// Transpose buffer of the FP64 lane reduction: 64 lanes + 1 pad (bank
// spread) per row, kRedM Q bins per pass.
constexpr int kRedRow = 65;
constexpr int kRedM = 16;

// Sum acc[OFF + n], n < N <= 64, over the wave's 64 lanes in FP64 (through
// LDS); returns, in lane L < N, the total of bin OFF + L.  Must be reached by
// every wave of the block.
template <int M, int OFF = 0, int N = M>
__device__ __forceinline__ double lane_reduce_f64(const float (&acc)[M], unsigned char* smem,
                                                  int wave, int lane) {
  static_assert(N <= 64 && OFF + N <= M, "bins per pass");
  float* red = reinterpret_cast<float*>(smem) + wave * kRedM * kRedRow;
  double mine = 0.0;
#pragma unroll
  for (int m0 = 0; m0 < N; m0 += kRedM) {
    constexpr int kLast = M - 1;
    __syncthreads();
#pragma unroll
    for (int mm = 0; mm < kRedM; ++mm)
      if (m0 + mm < N) red[mm * kRedRow + lane] = acc[min(OFF + m0 + mm, kLast)];
    __syncthreads();
    if (lane >= m0 && lane < min(m0 + kRedM, N)) {
      const float* row = red + (lane - m0) * kRedRow;
      double t = 0.0;
      for (int l = 0; l < 64; ++l) t += static_cast<double>(row[l]);
      mine = t;
    }
  }
  return mine;
}

}  // namespace
}  // namespace pyiid
