// Debye scattering kernels for gfx950 (CDNA4, wave64).
//
//  * fused_grad_kernel: ordered pairs (i, j), one lane per atom i, a wave
//    covers 64 atoms x m consecutive Q bins; j atoms are staged in LDS and
//    read as wave-wide broadcasts, and the per-pair setup is shared by the
//    block's waves through LDS.  Per (i, j) the geometry, the phase
//    theta0 = Q0 r and the step dtheta = dq r are formed in FP64 and reduced
//    with a Cody-Waite split of pi/2; sin/cos come from Taylor polynomials in
//    FP32.  Along Q, sin(Q r)/r and cos(Q r) follow the three-term
//    recurrence with 2 cos(dtheta) as a two-float constant (FP64 setup, the
//    low part applied every eighth bin), so the inner loop is FMA only.
//    Each lane accumulates F and the three gradient components for its atom;
//    no atomics, no K x Q intermediates.
//  * fq_kernel: unordered pairs (j < i) for F(Q) alone; omega_m = sin(Q_m r)/r
//    follows the three-term recurrence w_{m+1} = 2 cos(dtheta) w_m - w_{m-1}
//    with 2 cos(dtheta) as a two-float constant, the low part applied every
//    eighth bin (2.125 VALU ops per pair and Q bin), when every pair of the
//    wave has its step at least 0.2 away from a multiple of pi, else the
//    Reinsch form
//    w_{m+1} = w_m + D_{m+1}, D_{m+2} = D_{m+1} - eps w_{m+1}, eps =
//    4 sin^2(d/2) from a half-angle sincos of d = dtheta - k pi, on the
//    sequence (-1)^(k m) w_m.  With several elements the items run over the
//    atoms in element order, each j range of one element, so the loop is the
//    same and the sums take the weight f_i(Q) f_j(Q) once, at the end.
//  * partial sums leave each wave in FP64 and are reduced across tiles in
//    FP64 by small deterministic reduction kernels.
//
// Reference semantics (pyIID): cpu_flat.py:77-153, cpu_experimental.py:8-15,
// flat_multi_cpu_wrap.py:22-102, gpu_flat.py:11-299.
#include <type_traits>

#include "debye_math.h"
#include "internal.h"

namespace pyiid {
namespace {

// Smallest distance |d| = 0.2 of the per-bin phase step from a multiple of pi
// for the three-term recurrence of the F-only kernel (64 bins per anchor);
// closer steps take the Reinsch form.  The kernel tests
// eps(d) = 4 sin^2(d / 2) >= 4 sin^2(0.1).
constexpr float kChebEpsMin = 0.039866844f;

// Shared memory of the Debye kernels: the j tile during the pair loop, then
// reused as the transpose buffer of the end-of-kernel lane reduction.
constexpr int cmax3(int a, int b, int c) { return a > b ? (a > c ? a : c) : (b > c ? b : c); }
// Fused kernel shapes: a block covers M W Q bins of one 64-atom tile as W
// waves of M consecutive bins: M = 32 in 8 waves (2 per SIMD, 237 VGPRs),
// M = 16 in 16 waves (4 per SIMD within 128 VGPRs), or, with compensated
// sums, M = 16 in 8 waves.  Per round of GJ j atoms the block's waves (same
// atoms, different Q chunks) share the per-pair setup through LDS (FP32,
// [field][k][lane]): 1/r, (q_i - q_j)/r^2, sin(dtheta), 2 cos(dtheta) as hi
// and 8 lo, and the anchor sin, cos(Q0 r) of every wave's Q chunk.
// hi + lo -> hi, with the rounding error of that addition left in lo
// (TwoSum: exact for any magnitudes, no branches).
__device__ __forceinline__ void two_sum_fold(float& hi, float& lo) {
  const float s = hi + lo;
  const float bb = s - hi;
  lo = (hi - (s - bb)) + (lo - bb);
  hi = s;
}

// Launch shapes of the fused kernel: M bins per wave, WAVES waves per block
// (M WAVES bins per block).  Plain sums: 32 x 8 and 16 x 16 (256 bins);
// compensated sums: 16 x 8 (128 bins), which leaves the registers for a
// second set of accumulators.
template <int M, int WAVES>
struct FusedShape {
  static constexpr int W = WAVES;                 // waves per block
  static constexpr int kBins = M * W;             // bins per block
  static constexpr int GJ = W == 8 ? 16 : 8;      // j atoms per setup round
  static constexpr int H = W / 8;                 // setup tasks per pair (8 anchors each)
  static constexpr int kFields = 8 + 2 * W;
  static constexpr int kRow = M + 4;              // store-transpose row (16-byte aligned)
  static constexpr int kSmem = cmax3(W * kRedM * kRedRow * 4, W * 64 * kRow * 4,
                                     kJTile * 32 + GJ * 64 * kFields * 4);
};
// F-only kernel: rounds of kFqGeoJ j atoms share the per-pair setup (FP32,
// [field][k][lane]): 1/r, sin(dtheta), 2 cos(dtheta) as hi and 8 lo, eps of d,
// (-1)^k (dtheta = k pi + d), and the anchor sin, cos(Q0 r) of each of the
// block's kWaves Q chunks.  Rounds of 8 keep the block at 36 KB of LDS.
constexpr int kFqGeoJ = 8;
constexpr int kFqGeoFields = 6 + 2 * kWaves;
constexpr int kFqGeoBytes = kFqGeoJ * 64 * kFqGeoFields * 4;
constexpr int kFqSmemBytes = kWaves * kRedM * kRedRow * 4 > kJTile * 32 + kFqGeoBytes
                                 ? kWaves * kRedM * kRedRow * 4
                                 : kJTile * 32 + kFqGeoBytes;

// Coalesced store of a wave's 64 (atoms) x M (bins) tile: lane l holds row l.
// Rows go through LDS and leave as contiguous M-float row segments.
// dst(row) = base + row * row_stride; rows >= n_rows and bins >= n_valid are
// not written.  Must be reached by every wave of the block.
template <int M>
__device__ __forceinline__ void store_tile(const float (&acc)[M], unsigned char* smem, int wave,
                                           int lane, float* base, size_t row_stride, int n_rows,
                                           int n_valid) {
  static_assert(64 % M == 0 && M <= kGradM, "tile width");
  constexpr int kStoreRow = M + 4;
  float* buf = reinterpret_cast<float*>(smem) + wave * 64 * kStoreRow;
  __syncthreads();
#pragma unroll
  for (int m = 0; m < M; ++m) buf[lane * kStoreRow + m] = acc[m];
  __syncthreads();
  // dword stores: lanes sub = 0..M-1 of a row write M consecutive floats
  // (the row stride is only 4-byte aligned in general).
  constexpr int kRowsPer = 64 / M;  // rows per store instruction
  const int sub = lane % M;
#pragma unroll
  for (int t = 0; t < 64 / kRowsPer; ++t) {
    const int row = t * kRowsPer + lane / M;
    if (row < n_rows && sub < n_valid) base[row * row_stride + sub] = buf[row * kStoreRow + sub];
  }
}

// ---------------------------------------------------------------------------
// Fused F + gradient kernel (ordered pairs).
//   grid: x = work item (i tile x j range), y = Q group (W waves x M bins)
//   gpart[split][i][c][q]  float  sum_j t_ij(q) (q_i - q_j)_c / r^2 (ff-weighted)
//   fpart[item][q]         double sum_{i in tile} sum_{j < i} ff sin(Q r)/r (each
//                          unordered pair once, from the item that holds its
//                          j < i half)
// Weighted: the item's j indices are positions in jperm (the atoms in
// element order) and its j range is of one element, item.pad; the loop is
// the single-element one and f_i(Q) f_j(Q) scales the lane's sums at the end.
// ---------------------------------------------------------------------------
// COMP: compensated sums.  Every accumulator has a second FP32 register
// that takes the terms of the current j tile; at the end of the tile an
// error-free TwoSum moves it into the first and leaves the rounding error of
// that addition in the second (Knuth, TAOCP vol. 2, 4.2.2).  The FP32
// rounding then acts on sums of at most kJTile terms plus the carried error,
// not on a running sum over the whole j range.
// WITH_F: the item holds j < i pairs, so it adds to F (each unordered pair
// enters F once, from its j < i half, with the pair's weight 0 or 1); without,
// the F sums are skipped and the item's F partial is written as 0.
template <int M, int WAVES, bool WEIGHTED, bool COMP, bool WITH_F>
__device__ __forceinline__ void fused_grad_body(
    const AtomRec* __restrict__ pos, const int32_t* __restrict__ jperm, int n_total,
    const float* __restrict__ ftab, int nq, double dq, const Item& it,
    float* __restrict__ gpart, double* __restrict__ fpart, unsigned char* smem) {
  using S = FusedShape<M, WAVES>;
  constexpr int W = S::W, GJ = S::GJ, H = S::H;
  AtomRec* tile = reinterpret_cast<AtomRec*>(smem);
  float* geo_f = reinterpret_cast<float*>(smem + kJTile * 32);
  constexpr int kGS = GJ * 64;  // field stride in geo_f

  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int i = it.i_base + lane;
  const int qchunk = blockIdx.y * W + wave;
  const int q0 = qchunk * M;
  const int j_begin = it.jbeg;
  const int j_end = it.jend;
  const bool lane_ok = i < it.i_end;
  const AtomRec pi = pos[lane_ok ? i : it.i_base];

  // Q values stay in VGPRs: an FMA with a scalar-register operand issues
  // more slowly than an all-VGPR one on gfx950.
  float qv[M];
#pragma unroll
  for (int m = 0; m < M; ++m) {
    qv[m] = static_cast<float>(static_cast<double>(q0 + m) * dq);
    asm volatile("" : "+v"(qv[m]));
  }

  float aF[M], ax[M], ay[M], az[M];
  float bF[M], bx[M], by[M], bz[M];  // the tile's terms (COMP only)
#pragma unroll
  for (int m = 0; m < M; ++m) aF[m] = ax[m] = ay[m] = az[m] = bF[m] = bx[m] = by[m] = bz[m] = 0.0f;
  float (&sF)[M] = COMP ? bF : aF;
  float (&sx)[M] = COMP ? bx : ax;
  float (&sy)[M] = COMP ? by : ay;
  float (&sz)[M] = COMP ? bz : az;

  // first bin of the block (its waves take consecutive chunks of M bins)
  const double qbase = static_cast<double>(blockIdx.y * S::kBins);
  for (int jt = j_begin; jt < j_end; jt += kJTile) {
    __syncthreads();
    if ((int)threadIdx.x < kJTile && jt + (int)threadIdx.x < j_end) {
      const int ja = WEIGHTED ? jperm[jt + threadIdx.x] : jt + threadIdx.x;
      AtomRec rec = pos[ja];
      rec.w = static_cast<double>(ja);  // the j atom's index (its element is item.pad)
      tile[threadIdx.x] = rec;
    }
    __syncthreads();
    const int cnt = min(kJTile, j_end - jt);
    for (int js = 0; js < cnt; js += GJ) {
      const int kk = min(GJ, cnt - js);
      // shared setup: task t = (pair k, anchor part) of the round, wave w
      // takes tasks w, w + W, ...; part p writes the anchors of chunks
      // 8 p .. 8 p + 7 (part 0 also the pair's geometry fields)
      for (int t = wave; t < kk * H; t += W) {
        const int k = t / H, part = t - k * H;
        const AtomRec pj = tile[js + k];
        const PairGeom g = pair_geom(pi, pj, lane_ok);
        const double dth = dq * g.r;
        const int o = k * 64 + lane;
        if (part == 0) {
          float s1, k2hi, k2lo;
          two_cos(dth, k2hi, k2lo, s1);
          geo_f[o] = g.inv_r;
          geo_f[o + kGS] = g.dpx;
          geo_f[o + 2 * kGS] = g.dpy;
          geo_f[o + 3 * kGS] = g.dpz;
          geo_f[o + 4 * kGS] = s1;
          geo_f[o + 5 * kGS] = k2hi;
          geo_f[o + 6 * kGS] = 8.0f * k2lo;
          geo_f[o + 7 * kGS] = static_cast<int>(pj.w) < i ? 1.0f : 0.0f;  // F weight
        }
        // the anchors sin, cos(Q0 r) of the part's 8 Q chunks: the first
        // reduced in FP64, the others by FP32 rotations by M dtheta
        float as, ac, rs, rc;
        sincos_red((qbase + static_cast<double>(8 * part * M)) * dth, as, ac);
        sincos_red(static_cast<double>(M) * dth, rs, rc);
#pragma unroll
        for (int w = 0; w < 8; ++w) {
          geo_f[o + (8 + 2 * (8 * part + w)) * kGS] = as;
          geo_f[o + (9 + 2 * (8 * part + w)) * kGS] = ac;
          const float as2 = fmaf(as, rc, ac * rs);
          ac = fmaf(ac, rc, -as * rs);
          as = as2;
        }
      }
      __syncthreads();
      if (q0 < nq) {  // wave-uniform
        // Two register sets: the next atom's setup is read from LDS while the
        // current atom's bins run.  With two waves per SIMD a wave stalled on
        // LDS leaves the SIMD single-issuing; the scheduling barriers keep the
        // reads ahead of the bins (about 3% faster than reading at first use).
        float ga[10], gb[10];
        auto load = [&](float (&g)[10], int k) {
          const int o = k * 64 + lane;
#pragma unroll
          for (int f = 0; f < 8; ++f) g[f] = geo_f[o + f * kGS];
          g[8] = geo_f[o + (8 + 2 * wave) * kGS];
          g[9] = geo_f[o + (9 + 2 * wave) * kGS];
        };
        auto bins = [&](const float (&g)[10], auto with_f) {
          constexpr bool kWithF = decltype(with_f)::value;
          const float inv_r = g[0], dpx = g[1], dpy = g[2], dpz = g[3], s1 = g[4], k2 = g[5],
                      lo8 = g[6], fw = g[7];
          float sn = g[8], cs = g[9];
          // Along Q, w = sin(Q r) / r and cs = cos(Q r) follow the three-term
          // recurrence y' = 2 cos(dtheta) y - y_prev (one FMA each per bin)
          // with 2 cos(dtheta) = k2 + lo: the low part is added every eighth
          // bin as 8 lo, which keeps the frequency exact on average; the
          // remaining error is bounded (about 75 ulp after 32 bins for steps
          // near k pi) instead of the drift of an FP32-rounded coefficient.
          const float c1 = 0.5f * k2;
          float w = sn * inv_r;
          float wp = fmaf(sn, c1, -cs * s1) * inv_r;  // sin(Q0 r - dtheta) / r
          float cp = fmaf(cs, c1, sn * s1);           // cos(Q0 r - dtheta)
#pragma unroll
          for (int m = 0; m < M; ++m) {
            const float t = fmaf(qv[m], cs, -w);
            if constexpr (kWithF) sF[m] = fmaf(w, fw, sF[m]);
            sx[m] = fmaf(t, dpx, sx[m]);
            sy[m] = fmaf(t, dpy, sy[m]);
            sz[m] = fmaf(t, dpz, sz[m]);
            float wn = fmaf(k2, w, -wp);
            float cn = fmaf(k2, cs, -cp);
            if ((m & 7) == 7) {
              wn = fmaf(lo8, w, wn);
              cn = fmaf(lo8, cs, cn);
            }
            wp = w;
            cp = cs;
            w = wn;
            cs = cn;
          }
        };
        const std::integral_constant<bool, WITH_F> with_f;
        load(ga, 0);
        int k = 0;
        for (; k + 1 < kk; k += 2) {
          load(gb, k + 1);
          __builtin_amdgcn_sched_barrier(0);
          bins(ga, with_f);
          load(ga, k + 2 < kk ? k + 2 : k + 1);
          __builtin_amdgcn_sched_barrier(0);
          bins(gb, with_f);
        }
        if (k < kk) bins(ga, with_f);
      }
      __syncthreads();
    }
    if constexpr (COMP) {
#pragma unroll
      for (int m = 0; m < M; ++m) {
        two_sum_fold(aF[m], bF[m]);
        two_sum_fold(ax[m], bx[m]);
        two_sum_fold(ay[m], by[m]);
        two_sum_fold(az[m], bz[m]);
      }
    }
  }
  if constexpr (COMP) {
#pragma unroll
    for (int m = 0; m < M; ++m) {
      aF[m] += bF[m];
      ax[m] += bx[m];
      ay[m] += by[m];
      az[m] += bz[m];
    }
  }

  // Weighted: the item's j range is of one element, item.pad, so the pair
  // factor f_i(Q) f_j(Q) is the same for every term of the lane's sums.
  if constexpr (WEIGHTED) {
    const float* fti = ftab + static_cast<size_t>(pi.w) * nq;
    const float* ftj = ftab + static_cast<size_t>(it.pad) * nq;
#pragma unroll
    for (int m = 0; m < M; ++m) {
      const float f = (q0 + m < nq) ? fti[q0 + m] * ftj[q0 + m] : 0.0f;
      aF[m] *= f;
      ax[m] *= f;
      ay[m] *= f;
      az[m] *= f;
    }
  }

  // F partial of this wave (FP64 lane reduction), lane m keeps bin q0 + m.
  const double fmine = lane_reduce_f64<M>(aF, smem, wave, lane);
  if (lane < M && q0 + lane < nq) fpart[(size_t)blockIdx.x * nq + q0 + lane] = fmine;

  // gpart[split][i][c][q0 + m]: rows of the tile are atoms, stride 3 nq floats.
  const int i0 = it.i_base;
  const int n_rows = it.i_end - it.i_base;
  const int n_valid = max(0, min(M, nq - q0));
  float* base = gpart + (((size_t)it.split * n_total + i0) * 3) * nq + (q0 < nq ? q0 : 0);
  const size_t stride = (size_t)3 * nq;
  store_tile<M>(ax, smem, wave, lane, base, stride, n_rows, n_valid);
  store_tile<M>(ay, smem, wave, lane, base + nq, stride, n_rows, n_valid);
  store_tile<M>(az, smem, wave, lane, base + 2 * (size_t)nq, stride, n_rows, n_valid);
}

// An item whose j atoms are all at or above every i of its tile holds no
// j < i pair and skips the F sums.  Within an item the j atom index grows
// with the position (the unweighted j range is in atom order; a weighted one
// lies in one element's run of the stable element order), so its first j
// atom decides.  The two bodies share no live registers, so the branch costs
// nothing in register allocation.
template <int M, int WAVES, bool WEIGHTED, bool COMP>
__global__ __launch_bounds__(64 * WAVES) void fused_grad_kernel(
    const AtomRec* __restrict__ pos, const int32_t* __restrict__ jperm, int n_total,
    const float* __restrict__ ftab, int nq, double dq, const Item* __restrict__ items,
    float* __restrict__ gpart, double* __restrict__ fpart) {
  __shared__ __attribute__((aligned(32))) unsigned char smem[FusedShape<M, WAVES>::kSmem];
  const Item it = items[blockIdx.x];
  const int j_first = it.jbeg < it.jend ? (WEIGHTED ? jperm[it.jbeg] : it.jbeg) : it.i_end;
  if (j_first < it.i_end - 1)
    fused_grad_body<M, WAVES, WEIGHTED, COMP, true>(pos, jperm, n_total, ftab, nq, dq, it, gpart,
                                                    fpart, smem);
  else
    fused_grad_body<M, WAVES, WEIGHTED, COMP, false>(pos, jperm, n_total, ftab, nq, dq, it,
                                                     gpart, fpart, smem);
}

// ---------------------------------------------------------------------------
// F-only kernel (unordered pairs j < i), Chebyshev recurrence on omega.
//   grid: x = work item (i tile, j range of the same chain), y = Q group
//   fpart[item][q] double
// Item indices are positions in perm (the atoms in Z-order, after element
// order for a weighted table).  Weighted: the item's i tile is of one element,
// item.split, and its j range of one element, item.pad, so that
// F = sum_{e,e'} f_e f_e' S_ee' holds per item: the loop sums plain
// sin(Q r)/r and each bin's total is weighted by f_i(Q) f_j(Q) at the end.
// ---------------------------------------------------------------------------
// F-only inner loops over NP j atoms of a round at once (their LDS offsets
// o[p]), adding sin(Q_m r)/r of each pair to acc[m] for the wave's M bins.
// Three-term form: w_{m+2} = 2 cos(dtheta) w_{m+1} - w_m with 2 cos(dtheta)
// as hi + lo, the low part added every eighth bin as 8 lo.
template <int M, int NP>
__device__ __forceinline__ void fq_three_term(float (&acc)[M], const float* geo_f, int kGS,
                                              const int (&o)[NP], int chunk) {
  float w0[NP], w1[NP], k2[NP], lo8[NP];
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    const float inv_r = geo_f[o[p]], s1 = geo_f[o[p] + kGS];
    k2[p] = geo_f[o[p] + 2 * kGS];
    lo8[p] = geo_f[o[p] + 3 * kGS];
    const float sn = geo_f[o[p] + (6 + 2 * chunk) * kGS], cs = geo_f[o[p] + (7 + 2 * chunk) * kGS];
    w0[p] = sn * inv_r;
    w1[p] = fmaf(sn, 0.5f * k2[p], cs * s1) * inv_r;
  }
#pragma unroll
  for (int m = 0; m < M; ++m) {
    float t = w0[0];
#pragma unroll
    for (int p = 1; p < NP; ++p) t += w0[p];
    acc[m] += t;
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      float w2 = fmaf(k2[p], w1[p], -w0[p]);
      if ((m & 7) == 7) w2 = fmaf(lo8[p], w1[p], w2);
      w0[p] = w1[p];
      w1[p] = w2;
    }
  }
}

// Reinsch difference form on y_m = (-1)^(k m) w_m = sin(Q0 r + m d) / r
// (dtheta = k pi + d):  y_{m+1} = y_m + D_{m+1},  D_{m+2} = D_{m+1} - eps y_{m+1},
// eps = 4 sin^2(d/2); odd bins take the sign (-1)^k back.
template <int M, int NP>
__device__ __forceinline__ void fq_reinsch(float (&acc)[M], const float* geo_f, int kGS,
                                           const int (&o)[NP], int chunk) {
  float y[NP], dd[NP], eps[NP], sgn[NP];
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    const float inv_r = geo_f[o[p]], s1 = geo_f[o[p] + kGS];
    eps[p] = geo_f[o[p] + 4 * kGS];
    sgn[p] = geo_f[o[p] + 5 * kGS];
    const float sn = geo_f[o[p] + (6 + 2 * chunk) * kGS], cs = geo_f[o[p] + (7 + 2 * chunk) * kGS];
    // D_1 = y_1 - y_0 = (sin(t + d) - sin t) / r
    //     = 2 sin(d/2) (cos t cos(d/2) - sin t sin(d/2)) / r,
    // from sin(d/2)^2 = eps / 4 and sin(d/2) cos(d/2) = sgn s1 / 2
    // (no cancellation for small d).
    dd[p] = fmaf(cs, sgn[p] * s1, -2.0f * sn * (0.25f * eps[p])) * inv_r;
    y[p] = sn * inv_r;
  }
#pragma unroll
  for (int m = 0; m < M; m += 2) {
    float t = y[0];
#pragma unroll
    for (int p = 1; p < NP; ++p) t += y[p];
    acc[m] += t;
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      y[p] += dd[p];
      dd[p] = fmaf(-eps[p], y[p], dd[p]);
    }
    float u = acc[m + 1];
#pragma unroll
    for (int p = 0; p < NP; ++p) u = fmaf(sgn[p], y[p], u);
    acc[m + 1] = u;
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      y[p] += dd[p];
      dd[p] = fmaf(-eps[p], y[p], dd[p]);
    }
  }
}

// M bins per wave (a multiple of 8, at most 128); at least 3 waves per SIMD
// (2 with COMP, whose second accumulator set takes M more registers).
// COMP: compensated sums as in fused_grad_kernel, the terms of each j tile
// gathered in a second accumulator and folded in by TwoSum.
template <int M, bool WEIGHTED, bool COMP>
__global__ __launch_bounds__(64 * kWaves) __attribute__((amdgpu_waves_per_eu(COMP ? 2 : 3))) void
fq_kernel(const AtomRec* __restrict__ pos, const int32_t* __restrict__ perm,
          const float* __restrict__ ftab, int nq, double dq, const Item* __restrict__ items,
          double* __restrict__ fpart, unsigned long long* __restrict__ forms) {
  static_assert(M % 8 == 0 && M <= 128, "bins per wave");
  __shared__ __attribute__((aligned(32))) unsigned char smem[kFqSmemBytes];
  AtomRec* tile = reinterpret_cast<AtomRec*>(smem);
  float* geo_f = reinterpret_cast<float*>(smem + kJTile * 32);
  constexpr int kGS = kFqGeoJ * 64;

  const Item it = items[blockIdx.x];
  // an empty item (padding of a list sized on the host) writes nothing
  if (it.jbeg >= it.jend) return;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int i = it.i_base + lane;
  // (wave, j atom) steps run in the three-term and in the Reinsch form
  // (wave-uniform counts, added to forms[0], forms[1] at the end when given)
  unsigned n_three = 0, n_reinsch = 0;
  // The block's kWaves Q chunks go to its waves in an order rotated by a hash
  // of the block index: in a Q group with fewer chunks than waves (the last,
  // when kWaves M does not divide nq) the working waves of the blocks that
  // share a CU then fall on different SIMDs (wave w of a block runs on SIMD w).
  const int chunk = (wave + static_cast<int>((blockIdx.x * 0x9E3779B1u) >> 30)) & (kWaves - 1);
  const int q0 = (blockIdx.y * kWaves + chunk) * M;
  const bool lane_ok = i < it.i_end;
  const AtomRec pi = pos[perm[lane_ok ? i : it.i_base]];

  float acc[M], lo[M];
#pragma unroll
  for (int m = 0; m < M; ++m) acc[m] = lo[m] = 0.0f;
  float (&sum)[M] = COMP ? lo : acc;

  // first bin of the block (its waves take consecutive chunks of M bins)
  const double qbase = static_cast<double>(blockIdx.y * kWaves * M);
  for (int jt = it.jbeg; jt < it.jend; jt += kJTile) {
    __syncthreads();
    if ((int)threadIdx.x < kJTile && jt + (int)threadIdx.x < it.jend)
      tile[threadIdx.x] = pos[perm[jt + threadIdx.x]];
    __syncthreads();
    const int cnt = min(kJTile, it.jend - jt);
    for (int js = 0; js < cnt; js += kFqGeoJ) {
      const int kk = min(kFqGeoJ, cnt - js);
      // setup shared by the block's waves (same atoms, different Q chunks)
      for (int k = wave; k < kk; k += kWaves) {
        const PairGeom g = pair_geom(pi, tile[js + k], lane_ok && (jt + js + k < i));
        const double dth = dq * g.r;
        int k_odd;
        const double d = pi_reduce(dth, k_odd);
        float sh, ch;
        sincos_poly(static_cast<float>(0.5 * d), sh, ch);
        const float sgn = k_odd ? -1.0f : 1.0f;
        const int o = k * 64 + lane;
        geo_f[o] = g.inv_r;
        // sin(dtheta) = (-1)^k sin d; 2 cos(dtheta) = (-1)^k (2 - 4 sin^2(d/2))
        // in FP64, kept as hi + lo so that the three-term recurrence runs at
        // the exact frequency; eps = 4 sin^2(d/2), rounded once from FP64,
        // keeps its relative accuracy for every |d| (a Reinsch wave runs all
        // its lanes in that form, whatever their step).
        const double shd = sin_poly_f64(0.5 * d);
        const double k2 = (k_odd ? -1.0 : 1.0) * fma(-4.0 * shd, shd, 2.0);
        const float k2hi = static_cast<float>(k2);
        geo_f[o + kGS] = sgn * 2.0f * sh * ch;
        geo_f[o + 2 * kGS] = k2hi;
        geo_f[o + 3 * kGS] = 8.0f * static_cast<float>(k2 - static_cast<double>(k2hi));
        geo_f[o + 4 * kGS] = static_cast<float>(4.0 * shd * shd);
        geo_f[o + 5 * kGS] = sgn;
        // the anchors sin, cos(Q0 r) of the block's kWaves Q chunks: the
        // first reduced in FP64, the others by FP32 rotations by M dtheta
        float as, ac, rs, rc;
        sincos_red(qbase * dth, as, ac);
        sincos_red(static_cast<double>(M) * dth, rs, rc);
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
          geo_f[o + (6 + 2 * w) * kGS] = as;
          geo_f[o + (7 + 2 * w) * kGS] = ac;
          const float as2 = fmaf(as, rc, ac * rs);
          ac = fmaf(ac, rc, -as * rs);
          as = as2;
        }
      }
      __syncthreads();
      if (q0 < nq) {  // wave-uniform
        // When every kept pair of the wave has |d| >= 0.2 the three-term form
        // with the two-float coefficient (exact frequency on average,
        // rounding amplified by at most 1/sin|d| < 5) is accurate enough;
        // otherwise the wave runs the Reinsch form, whose rounding stays O(1)
        // in the step count.  The choice is wave-uniform: bit k of `three`
        // marks the round's j atoms of the three-term form, and each form
        // runs as a loop of its own over its atoms (one loop with a branch per
        // atom makes the compiler route the sums through copies).
        uint32_t three = 0;
        for (int k = 0; k < kk; ++k) {
          const int o = k * 64 + lane;
          if (__all(geo_f[o] == 0.0f || geo_f[o + 4 * kGS] >= kChebEpsMin)) three |= 1u << k;
        }
        n_three += __builtin_popcount(three);
        n_reinsch += kk - __builtin_popcount(three);
        // Two j atoms at a time: two independent recurrences per bin keep the
        // wave issuing while one waits on its previous step.  An odd atom
        // left over runs alone after the loop.
        uint32_t left = three;
        for (; __builtin_popcount(left) >= 2; left &= left - 1, left &= left - 1) {
          const int ka = __builtin_ctz(left), kb = __builtin_ctz(left & (left - 1));
          fq_three_term<M, 2>(sum, geo_f, kGS, {ka * 64 + lane, kb * 64 + lane}, chunk);
        }
        if (left != 0) fq_three_term<M, 1>(sum, geo_f, kGS, {__builtin_ctz(left) * 64 + lane}, chunk);
        left = ~three & ((1u << kk) - 1);
        for (; __builtin_popcount(left) >= 2; left &= left - 1, left &= left - 1) {
          const int ka = __builtin_ctz(left), kb = __builtin_ctz(left & (left - 1));
          fq_reinsch<M, 2>(sum, geo_f, kGS, {ka * 64 + lane, kb * 64 + lane}, chunk);
        }
        if (left != 0) fq_reinsch<M, 1>(sum, geo_f, kGS, {__builtin_ctz(left) * 64 + lane}, chunk);
      }
      __syncthreads();
    }
    if constexpr (COMP) {
#pragma unroll
      for (int m = 0; m < M; ++m) two_sum_fold(acc[m], lo[m]);
    }
  }
  if constexpr (COMP) {
#pragma unroll
    for (int m = 0; m < M; ++m) acc[m] += lo[m];
  }

  // lane l keeps bin q0 + l, and bin q0 + 64 + l when M > 64; weighted: the
  // item's atoms are of one element on each side (item.split for i, item.pad
  // for j), so each bin's total takes f_i(Q) f_j(Q) once
  auto weight = [&](int q) -> double {
    if constexpr (WEIGHTED)
      return static_cast<double>(ftab[static_cast<size_t>(it.split) * nq + q] *
                                 ftab[static_cast<size_t>(it.pad) * nq + q]);
    return 1.0;
  };
  constexpr int kLo = M < 64 ? M : 64;
  const double mine = lane_reduce_f64<M, 0, kLo>(acc, smem, wave, lane);
  if (lane < kLo && q0 + lane < nq)
    fpart[(size_t)blockIdx.x * nq + q0 + lane] = mine * weight(q0 + lane);
  if constexpr (M > 64) {
    const double hi = lane_reduce_f64<M, 64, M - 64>(acc, smem, wave, lane);
    if (lane < M - 64 && q0 + 64 + lane < nq)
      fpart[(size_t)blockIdx.x * nq + q0 + 64 + lane] = hi * weight(q0 + 64 + lane);
  }
  if (forms != nullptr && lane == 0) {
    atomicAdd(&forms[0], static_cast<unsigned long long>(n_three));
    atomicAdd(&forms[1], static_cast<unsigned long long>(n_reinsch));
  }
}

// ---------------------------------------------------------------------------
// Contracted forces over unordered pairs.  An item is a pair of tiles (I, J),
// J <= I, of at most 64 positions of one chain in perm (the atom order of the
// work list; for a weighted table each tile is of one element); a wave takes
// one item and one Q group.  Lane l
// holds atom I_l; the J tile's atoms and their force accumulators turn round
// the wave one lane per step, so lane l meets J_((l+s) mod 64) at step s and
// every unordered pair is met once (a diagonal item, I = J, runs steps 1..32
// and drops the duplicate half of step 32).  Per pair the Q sum collapses to
// two scalars
//   A = sum_q u_q Q_q ff cos(Q r),   B = sum_q u_q ff sin(Q r),
// and the pair adds S (q_i - q_j) / r^2 to atom i and the negative to atom j
// with S = A - B / r: the same quantity as sum_q u_q grad(i, c, q) of the
// fused kernel (t = Q cos - sin / r there), without per-bin accumulators.
//   part[work][side][64][3]  (side 0: I tile, 1: J tile; work = item * groups + g,
//                             rows are tile positions)
// The recurrences restart every kPfChunk bins from an anchor: an FP64-reduced
// sincos every kPfAnchor chunks, an FP32 rotation by kPfChunk dtheta between.
// ---------------------------------------------------------------------------
constexpr int kPfChunk = 32;
constexpr int kPfAnchor = 4;  // chunks per FP64-reduced anchor

// S = A - B / r of NP pairs (geometry g[p]) over the wave's `span` bins from
// q_lo, with uw[k] = (u, u Q) of bin q_lo + k:
//   A = sum u Q cos(Q r),  B = sum u sin(Q r).
// sin and cos follow the three-term recurrence (one FMA each per bin) with
// 2 cos(dtheta) = k2 + lo, the low part added every eighth bin: exact
// frequency on average, bounded error for every dtheta (close pairs and steps
// near multiples of pi included), no per-wave branch.  The recurrences restart
// every kPfChunk bins from an anchor: reduced in FP64 every kPfAnchor chunks,
// rotated in FP32 by kPfChunk dtheta in between.  The NP pairs share each
// bin's u reads.
template <int NP>
__device__ __forceinline__ void pf_sums(const float2* uw, int span, int q_lo, double dq,
                                        const PairGeom (&g)[NP], float (&S)[NP]) {
  double dth[NP];
  float k2[NP], lo8[NP], c1[NP], s1[NP], rs[NP], rc[NP], an_s[NP], an_c[NP];
  float a0[NP], a1[NP], b0[NP], b1[NP];
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    dth[p] = dq * g[p].r;
    float lo;
    two_cos(dth[p], k2[p], lo, s1[p]);
    c1[p] = 0.5f * k2[p];
    lo8[p] = 8.0f * lo;
    rs[p] = 0.f;
    rc[p] = 1.f;
    if (span > kPfChunk) sincos_red(static_cast<double>(kPfChunk) * dth[p], rs[p], rc[p]);
    an_s[p] = 0.f;
    an_c[p] = 1.f;
    a0[p] = a1[p] = b0[p] = b1[p] = 0.f;
  }
  for (int c0 = 0; c0 < span; c0 += kPfChunk) {
    float sn[NP], cs[NP], sp[NP], cp[NP];
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      if ((c0 / kPfChunk) % kPfAnchor == 0) {
        sincos_red(static_cast<double>(q_lo + c0) * dth[p], an_s[p], an_c[p]);
      } else {
        const float s2 = fmaf(an_s[p], rc[p], an_c[p] * rs[p]);
        an_c[p] = fmaf(an_c[p], rc[p], -an_s[p] * rs[p]);
        an_s[p] = s2;
      }
      sn[p] = an_s[p];
      cs[p] = an_c[p];
      sp[p] = fmaf(sn[p], c1[p], -cs[p] * s1[p]);
      cp[p] = fmaf(cs[p], c1[p], sn[p] * s1[p]);
    }
#pragma unroll
    for (int m = 0; m < kPfChunk; ++m) {
      const float2 w = uw[c0 + m];
#pragma unroll
      for (int p = 0; p < NP; ++p) {
        if (m & 1) {
          a1[p] = fmaf(w.y, cs[p], a1[p]);
          b1[p] = fmaf(w.x, sn[p], b1[p]);
        } else {
          a0[p] = fmaf(w.y, cs[p], a0[p]);
          b0[p] = fmaf(w.x, sn[p], b0[p]);
        }
        float sn2 = fmaf(k2[p], sn[p], -sp[p]);
        float cs2 = fmaf(k2[p], cs[p], -cp[p]);
        if ((m & 7) == 7) {
          sn2 = fmaf(lo8[p], sn[p], sn2);
          cs2 = fmaf(lo8[p], cs[p], cs2);
        }
        sp[p] = sn[p];
        cp[p] = cs[p];
        sn[p] = sn2;
        cs[p] = cs2;
      }
    }
  }
#pragma unroll
  for (int p = 0; p < NP; ++p) S[p] = (a0[p] + a1[p]) - (b0[p] + b1[p]) * g[p].inv_r;
}

template <bool WEIGHTED>
__global__ __launch_bounds__(256) void pair_force_kernel(
    const AtomRec* __restrict__ pos, const int32_t* __restrict__ perm,
    const float* __restrict__ ftab, int nq, double dq, const Item* __restrict__ items,
    int n_work, int n_groups, int span, const float* __restrict__ u,
    const float* __restrict__ uq, float* __restrict__ part) {
  extern __shared__ __attribute__((aligned(16))) float lds[];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int work = blockIdx.x * kWaves + wave;
  const bool active = work < n_work;
  const int item = active ? work / n_groups : 0;
  const int g = active ? work - item * n_groups : 0;
  const Item it = items[item];
  const int q_lo = g * span;
  // u and u Q of the wave's Q group; weighted: times f_I(Q) f_J(Q), the
  // item's tiles being of one element each (item.split = e_I | e_J << 16)
  float2* uw = reinterpret_cast<float2*>(lds) + static_cast<size_t>(wave) * span;
  const float* uc = u + static_cast<size_t>(it.chain) * nq;
  const float* uqc = uq + static_cast<size_t>(it.chain) * nq;
  const float* fI = ftab + static_cast<size_t>(it.split & 0xffff) * nq;
  const float* fJ = ftab + static_cast<size_t>(it.split >> 16) * nq;
  for (int k = lane; k < span; k += 64) {
    const int q = q_lo + k;
    float2 v = make_float2(0.f, 0.f);
    if (active && q < nq) {
      const float ff = WEIGHTED ? fI[q] * fJ[q] : 1.0f;
      v = make_float2(uc[q] * ff, uqc[q] * ff);
    }
    uw[k] = v;
  }
  __syncthreads();
  if (!active) return;

  const bool diag = it.jbeg == it.i_base;
  const int i = it.i_base + lane;
  const bool i_ok = i < it.i_end;
  const AtomRec pi = pos[perm[i_ok ? i : it.i_base]];
  int jl = lane;  // slot (position jbeg + jl) whose data and accumulators this lane holds
  AtomRec pj = pos[perm[it.jbeg + lane < it.jend ? it.jbeg + lane : it.jbeg]];
  float fix = 0.f, fiy = 0.f, fiz = 0.f, fjx = 0.f, fjy = 0.f, fjz = 0.f;
  const int src = (lane + 1) & 63;
  auto turn = [&]() {
    pj.x = __shfl(pj.x, src, 64);
    pj.y = __shfl(pj.y, src, 64);
    pj.z = __shfl(pj.z, src, 64);
    jl = __shfl(jl, src, 64);
    fjx = __shfl(fjx, src, 64);
    fjy = __shfl(fjy, src, 64);
    fjz = __shfl(fjz, src, 64);
  };
  if (diag) {
    // I = J: offsets 1..32 cover every unordered pair once (the upper half of
    // offset 32 repeats the lower half and is dropped)
    turn();
    for (int s = 1; s <= 32; ++s) {
      const bool keep = i_ok && it.jbeg + jl < it.jend && !(s == 32 && lane >= 32);
      const PairGeom pg = pair_geom(pi, pj, keep);
      float S[1];
      pf_sums<1>(uw, span, q_lo, dq, {pg}, S);
      fix = fmaf(S[0], pg.dpx, fix);
      fiy = fmaf(S[0], pg.dpy, fiy);
      fiz = fmaf(S[0], pg.dpz, fiz);
      fjx = fmaf(-S[0], pg.dpx, fjx);
      fjy = fmaf(-S[0], pg.dpy, fjy);
      fjz = fmaf(-S[0], pg.dpz, fjz);
      turn();
    }
  } else {
    // two pairs per lane and step, sharing the step's u reads from LDS: the
    // lane's J slot (offset s) and the slot of lane ^ 32 (offset s + 32);
    // 32 steps cover the 64 offsets.  The second pair's J-side force goes back
    // to the lane that holds that slot.
    for (int s = 0; s < 32; ++s) {
      AtomRec pk = pj;
      pk.x = __shfl_xor(pj.x, 32, 64);
      pk.y = __shfl_xor(pj.y, 32, 64);
      pk.z = __shfl_xor(pj.z, 32, 64);
      const int kl = __shfl_xor(jl, 32, 64);
      const PairGeom g1 = pair_geom(pi, pj, i_ok && it.jbeg + jl < it.jend);
      const PairGeom g2 = pair_geom(pi, pk, i_ok && it.jbeg + kl < it.jend);
      float S[2];
      pf_sums<2>(uw, span, q_lo, dq, {g1, g2}, S);
      fix = fmaf(S[0], g1.dpx, fmaf(S[1], g2.dpx, fix));
      fiy = fmaf(S[0], g1.dpy, fmaf(S[1], g2.dpy, fiy));
      fiz = fmaf(S[0], g1.dpz, fmaf(S[1], g2.dpz, fiz));
      fjx = fmaf(-S[0], g1.dpx, fjx) + __shfl_xor(-S[1] * g2.dpx, 32, 64);
      fjy = fmaf(-S[0], g1.dpy, fjy) + __shfl_xor(-S[1] * g2.dpy, 32, 64);
      fjz = fmaf(-S[0], g1.dpz, fjz) + __shfl_xor(-S[1] * g2.dpz, 32, 64);
      turn();
    }
  }
  float* o = part + static_cast<size_t>(work) * 384;
  o[lane * 3 + 0] = fix;
  o[lane * 3 + 1] = fiy;
  o[lane * 3 + 2] = fiz;
  o[192 + jl * 3 + 0] = fjx;
  o[192 + jl * 3 + 1] = fjy;
  o[192 + jl * 3 + 2] = fjz;
}

// force[i][c] = conv * sum over the tile's (work, side) entries (FP64, fixed
// order); one thread per (tile, lane, component).
__global__ void reduce_pair_force_kernel(const float* __restrict__ part,
                                         const int* __restrict__ entries,
                                         const Range* __restrict__ tiles, int n_tiles,
                                         const int32_t* __restrict__ perm, double conv,
                                         double* __restrict__ out) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= (int64_t)n_tiles * 192) return;
  const int tile = static_cast<int>(t / 192), rem = static_cast<int>(t % 192);
  const Range tr = tiles[tile];
  if (tr.i_base + rem / 3 >= tr.i_end) return;
  double acc = 0.0;
  for (int e = tr.begin; e < tr.end; ++e)
    acc += static_cast<double>(part[static_cast<size_t>(entries[e]) * 192 + rem]);
  out[static_cast<size_t>(perm[tr.i_base + rem / 3]) * 3 + rem % 3] = acc * conv;
}

// uq[b][q] = u[b][q] * Q_q, Q_q = float32(q * dq) as in the Debye kernels
__global__ void uq_kernel(const float* __restrict__ u, int64_t total, int nq, double dq,
                          float* __restrict__ uq) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e < total) uq[e] = u[e] * static_cast<float>(static_cast<double>(e % nq) * dq);
}

// out[r][q] = scale[r][q] * sum_{p in ranges[r]} part[p][q] in FP64.  A block
// covers 64 bins of one row with 16 slices of the partials (p = begin + s,
// begin + s + 16, ...); the slices are added in a fixed order (deterministic).
__global__ __launch_bounds__(1024) void reduce_ranges_kernel(const double* __restrict__ part,
                                                             const Range* __restrict__ ranges,
                                                             int n_rows, int nq,
                                                             const double* __restrict__ scale,
                                                             double* __restrict__ out64,
                                                             float* __restrict__ out32) {
  __shared__ double sh[16][65];
  const int qo = threadIdx.x & 63, sl = threadIdx.x >> 6;
  const int q = blockIdx.x * 64 + qo;
  const int r = blockIdx.y;
  const Range rg = ranges[r];
  double acc = 0.0;
  if (q < nq)
    for (int p = rg.begin + sl; p < rg.end; p += 16) acc += part[(size_t)p * nq + q];
  sh[sl][qo] = acc;
  __syncthreads();
  if (sl != 0 || q >= nq || r >= n_rows) return;
  double t = 0.0;
#pragma unroll
  for (int k = 0; k < 16; ++k) t += sh[k][qo];
  const double v = t * scale[(size_t)r * nq + q];
  if (out64) out64[(size_t)r * nq + q] = v;
  if (out32) out32[(size_t)r * nq + q] = static_cast<float>(v);
}

// out[e] = gscale[e mod nq] * sum_s gpart[s][e] in FP64 over the flat
// element index e of the rows x nq planes (rows = (atom, component)); the
// planes are summed in plane order.  Grid-strided; PAIR: two elements per
// thread with 8-byte loads (the plane size is even, so every plane stays
// 8-byte aligned).  The loads of four planes are issued before their adds,
// so a thread keeps several HBM reads in flight.
template <typename T, bool PAIR>
__global__ __launch_bounds__(256) void reduce_grad_kernel(const float* __restrict__ gpart,
                                                          int n_split, int64_t total, int nq,
                                                          const double* __restrict__ gscale,
                                                          T* __restrict__ out) {
  constexpr int kPer = PAIR ? 2 : 1;
  using V = typename std::conditional<PAIR, float2, float>::type;
  const int64_t n_vec = total / kPer;
  const V* planes = reinterpret_cast<const V*>(gpart);
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n_vec;
       v += (int64_t)gridDim.x * blockDim.x) {
    double acc[kPer] = {};
    auto add = [&](const V& x) {
      if constexpr (PAIR) {
        acc[0] += static_cast<double>(x.x);
        acc[1] += static_cast<double>(x.y);
      } else {
        acc[0] += static_cast<double>(x);
      }
    };
    int s = 0;
    for (; s + 4 <= n_split; s += 4) {
      const V x0 = planes[(s + 0) * n_vec + v], x1 = planes[(s + 1) * n_vec + v];
      const V x2 = planes[(s + 2) * n_vec + v], x3 = planes[(s + 3) * n_vec + v];
      add(x0);
      add(x1);
      add(x2);
      add(x3);
    }
    for (; s < n_split; ++s) add(planes[s * n_vec + v]);
    const int64_t e = v * kPer;
    int q = static_cast<int>(e % nq);
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      out[e + k] = static_cast<T>(acc[k] * gscale[q]);
      if (++q == nq) q = 0;
    }
  }
}

// na(q) = N * mean_{i>j} f_i f_j = ((sum_e n_e f_e)^2 - sum_e n_e f_e^2) / (N - 1),
// per chain (counts[chain][e], n_atoms[chain]).  Unweighted kernels (one
// element) accumulate without f_i f_j = f^2, so their scales carry w = f^2;
// weighted kernels carry w = 1.  0/0 -> 0 (nan_to_num).
__global__ void norm_kernel(const float* __restrict__ ftab, const int64_t* __restrict__ counts,
                            int n_elem, int nq, const int64_t* __restrict__ n_atoms, int n_chain,
                            int weighted, double* na, double* fsu, double* fso, double* gs) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (q >= nq || b >= n_chain) return;
  const int64_t n = n_atoms[b];
  double s1 = 0.0, s2 = 0.0;
  for (int e = 0; e < n_elem; ++e) {
    const double f = ftab[(size_t)e * nq + q];
    const double c = static_cast<double>(counts[(size_t)b * n_elem + e]);
    s1 += c * f;
    s2 += c * f * f;
  }
  double nav = 0.0, w = 0.0;
  if (n >= 2) {
    nav = (s1 * s1 - s2) / static_cast<double>(n - 1);
    const double f0 = ftab[q];
    w = weighted ? 1.0 : f0 * f0;
  }
  const bool ok = (n >= 2) && (nav != 0.0);
  const size_t o = (size_t)b * nq + q;
  if (na) na[o] = (n >= 2) ? nav : __builtin_nan("");
  fsu[o] = ok ? 2.0 * w / nav : 0.0;
  fso[o] = ok ? w / nav : 0.0;
  gs[o] = ok ? -w / nav : 0.0;
}

// Flat pair k = j + i(i-1)/2 (i > j): d = q_i - q_j, r = |d| (float32 out).
__global__ void pair_geometry_kernel(const AtomRec* __restrict__ pos, int64_t k0, int64_t kc,
                                     float* __restrict__ d, float* __restrict__ r) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < kc;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t k = k0 + t;
    int64_t i = static_cast<int64_t>((1.0 + sqrt(1.0 + 8.0 * static_cast<double>(k))) * 0.5);
    if (i * (i - 1) / 2 > k) --i;
    if ((i + 1) * i / 2 <= k) ++i;
    const int64_t j = k - i * (i - 1) / 2;
    const AtomRec a = pos[i], b = pos[j];
    const double dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    const float fx = static_cast<float>(dx), fy = static_cast<float>(dy),
                fz = static_cast<float>(dz);
    if (d) {
      d[3 * t + 0] = fx;
      d[3 * t + 1] = fy;
      d[3 * t + 2] = fz;
    }
    if (r) r[t] = static_cast<float>(sqrt(fma(dx, dx, fma(dy, dy, dz * dz))));
  }
}

int grid_for(int64_t total, int block) {
  int64_t g = (total + block - 1) / block;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return static_cast<int>(g);
}

}  // namespace

template <int M, int WAVES, bool COMP>
hipError_t launch_fused_m(hipStream_t st, const AtomRec* pos, const int32_t* jperm, int n_total,
                          const float* ftab, int nq, double dq, const Item* items, int n_items,
                          bool weighted, float* gpart, double* fpart) {
  constexpr int kBins = FusedShape<M, WAVES>::kBins;
  dim3 grid(n_items, (nq + kBins - 1) / kBins);
  dim3 block(64 * WAVES);
  if (weighted)
    hipLaunchKernelGGL((fused_grad_kernel<M, WAVES, true, COMP>), grid, block, 0, st, pos, jperm,
                       n_total, ftab, nq, dq, items, gpart, fpart);
  else
    hipLaunchKernelGGL((fused_grad_kernel<M, WAVES, false, COMP>), grid, block, 0, st, pos, jperm,
                       n_total, ftab, nq, dq, items, gpart, fpart);
  return hipGetLastError();
}

hipError_t launch_fused_grad(hipStream_t st, int m, bool compensated, const AtomRec* pos,
                             const int32_t* jperm, int n_total, const float* ftab, int nq,
                             double dq, const Item* items, int n_items, bool weighted,
                             float* gpart, double* fpart) {
  if (n_items <= 0) return hipSuccess;
  if (compensated) {
    if (m != 16) return hipErrorInvalidValue;
    return launch_fused_m<16, 8, true>(st, pos, jperm, n_total, ftab, nq, dq, items, n_items,
                                       weighted, gpart, fpart);
  }
  if (m == 16)
    return launch_fused_m<16, 16, false>(st, pos, jperm, n_total, ftab, nq, dq, items, n_items,
                                         weighted, gpart, fpart);
  if (m == 32)
    return launch_fused_m<32, 8, false>(st, pos, jperm, n_total, ftab, nq, dq, items, n_items,
                                        weighted, gpart, fpart);
  return hipErrorInvalidValue;
}

hipError_t launch_pair_force(hipStream_t st, const AtomRec* pos, const int32_t* perm,
                             const float* ftab, int nq, double dq, const Item* items, int n_items,
                             int n_groups, int span, bool weighted, const float* u,
                             const float* uq, float* part) {
  const int n_work = n_items * n_groups;
  if (n_work <= 0) return hipSuccess;
  const size_t lds = sizeof(float) * 2 * kWaves * static_cast<size_t>(span);
  if (span % kPfChunk != 0 || lds > kPairForceLds) return hipErrorInvalidValue;
  dim3 grid((n_work + kWaves - 1) / kWaves);
  if (weighted)
    hipLaunchKernelGGL(pair_force_kernel<true>, grid, dim3(256), lds, st, pos, perm, ftab, nq, dq,
                       items, n_work, n_groups, span, u, uq, part);
  else
    hipLaunchKernelGGL(pair_force_kernel<false>, grid, dim3(256), lds, st, pos, perm, ftab, nq,
                       dq, items, n_work, n_groups, span, u, uq, part);
  return hipGetLastError();
}

hipError_t launch_reduce_pair_force(hipStream_t st, const float* part, const int* entries,
                                    const Range* tiles, int n_tiles, const int32_t* perm,
                                    double conv, double* out) {
  const int64_t total = (int64_t)n_tiles * 192;
  if (total <= 0) return hipSuccess;
  hipLaunchKernelGGL(reduce_pair_force_kernel, dim3(static_cast<unsigned>((total + 255) / 256)),
                     dim3(256), 0, st, part, entries, tiles, n_tiles, perm, conv, out);
  return hipGetLastError();
}

hipError_t launch_uq(hipStream_t st, const float* u, int n_chain, int nq, double dq, float* uq) {
  const int64_t total = (int64_t)n_chain * nq;
  if (total <= 0) return hipSuccess;
  hipLaunchKernelGGL(uq_kernel, dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0,
                     st, u, total, nq, dq, uq);
  return hipGetLastError();
}

template <int M, bool WEIGHTED, bool COMP = false>
static void launch_fq_m(hipStream_t st, const AtomRec* pos, const int32_t* perm,
                        const float* ftab, int nq, double dq, const Item* items, int n_items,
                        double* fpart, unsigned long long* forms) {
  dim3 grid(n_items, (nq + kWaves * M - 1) / (kWaves * M));
  hipLaunchKernelGGL((fq_kernel<M, WEIGHTED, COMP>), grid, dim3(64 * kWaves), 0, st, pos, perm,
                     ftab, nq, dq, items, fpart, forms);
}

hipError_t launch_fq(hipStream_t st, int m, bool compensated, const AtomRec* pos,
                     const int32_t* perm, const float* ftab, int nq, double dq, const Item* items,
                     int n_items, bool weighted, double* fpart, unsigned long long* forms) {
  if (n_items <= 0) return hipSuccess;
  if (compensated) {
    if (m != 64) return hipErrorInvalidValue;
    if (weighted)
      launch_fq_m<64, true, true>(st, pos, perm, ftab, nq, dq, items, n_items, fpart, forms);
    else
      launch_fq_m<64, false, true>(st, pos, perm, ftab, nq, dq, items, n_items, fpart, forms);
    return hipGetLastError();
  }
  if (weighted) {
    if (m == 64)
      launch_fq_m<64, true>(st, pos, perm, ftab, nq, dq, items, n_items, fpart, forms);
    else if (m == 72)
      launch_fq_m<72, true>(st, pos, perm, ftab, nq, dq, items, n_items, fpart, forms);
    else if (m == 80)
      launch_fq_m<80, true>(st, pos, perm, ftab, nq, dq, items, n_items, fpart, forms);
    else if (m == 88)
      launch_fq_m<88, true>(st, pos, perm, ftab, nq, dq, items, n_items, fpart, forms);
    else if (m == 96)
      launch_fq_m<96, true>(st, pos, perm, ftab, nq, dq, items, n_items, fpart, forms);
    else if (m == 112)
      launch_fq_m<112, true>(st, pos, perm, ftab, nq, dq, items, n_items, fpart, forms);
    else
      return hipErrorInvalidValue;
  } else {
    if (m == 64)
      launch_fq_m<64, false>(st, pos, perm, ftab, nq, dq, items, n_items, fpart, forms);
    else if (m == 72)
      launch_fq_m<72, false>(st, pos, perm, ftab, nq, dq, items, n_items, fpart, forms);
    else if (m == 80)
      launch_fq_m<80, false>(st, pos, perm, ftab, nq, dq, items, n_items, fpart, forms);
    else if (m == 88)
      launch_fq_m<88, false>(st, pos, perm, ftab, nq, dq, items, n_items, fpart, forms);
    else if (m == 96)
      launch_fq_m<96, false>(st, pos, perm, ftab, nq, dq, items, n_items, fpart, forms);
    else if (m == 112)
      launch_fq_m<112, false>(st, pos, perm, ftab, nq, dq, items, n_items, fpart, forms);
    else
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_reduce_ranges(hipStream_t st, const double* part, const Range* ranges,
                                int n_rows, int nq, const double* scale, double* out_f64,
                                float* out_f32) {
  if (n_rows <= 0) return hipSuccess;
  hipLaunchKernelGGL(reduce_ranges_kernel, dim3((nq + 63) / 64, n_rows), dim3(1024), 0, st, part,
                     ranges, n_rows, nq, scale, out_f64, out_f32);
  return hipGetLastError();
}

hipError_t launch_reduce_grad(hipStream_t st, const float* gpart, int n_split, int64_t n, int nq,
                              const double* gscale, void* out, bool out_f64) {
  const int64_t total = n * 3 * nq;
  if (total <= 0) return hipSuccess;
  const bool pair = total % 2 == 0 && reinterpret_cast<uintptr_t>(gpart) % 8 == 0;
  const int64_t n_vec = pair ? total / 2 : total;
  // grid-strided over at most 32 blocks per CU (256 CUs)
  const unsigned blocks = static_cast<unsigned>(std::min<int64_t>((n_vec + 255) / 256, 256 * 32));
  if (pair) {
    if (out_f64)
      hipLaunchKernelGGL((reduce_grad_kernel<double, true>), dim3(blocks), dim3(256), 0, st, gpart,
                         n_split, total, nq, gscale, static_cast<double*>(out));
    else
      hipLaunchKernelGGL((reduce_grad_kernel<float, true>), dim3(blocks), dim3(256), 0, st, gpart,
                         n_split, total, nq, gscale, static_cast<float*>(out));
  } else {
    if (out_f64)
      hipLaunchKernelGGL((reduce_grad_kernel<double, false>), dim3(blocks), dim3(256), 0, st,
                         gpart, n_split, total, nq, gscale, static_cast<double*>(out));
    else
      hipLaunchKernelGGL((reduce_grad_kernel<float, false>), dim3(blocks), dim3(256), 0, st,
                         gpart, n_split, total, nq, gscale, static_cast<float*>(out));
  }
  return hipGetLastError();
}

hipError_t launch_norm(hipStream_t st, const float* ftab, const int64_t* counts, int n_elem,
                       int nq, const int64_t* n_atoms, int n_chain, bool weighted, double* na,
                       double* fsu, double* fso, double* gs) {
  hipLaunchKernelGGL(norm_kernel, dim3((nq + 255) / 256, n_chain), dim3(256), 0, st, ftab, counts,
                     n_elem, nq, n_atoms, n_chain, weighted ? 1 : 0, na, fsu, fso, gs);
  return hipGetLastError();
}

hipError_t launch_pair_geometry(hipStream_t st, const AtomRec* pos, int64_t k_offset,
                                int64_t k_count, float* d, float* r) {
  hipLaunchKernelGGL(pair_geometry_kernel, dim3(grid_for(k_count, 256)), dim3(256), 0, st, pos,
                     k_offset, k_count, d, r);
  return hipGetLastError();
}

}  // namespace pyiid
