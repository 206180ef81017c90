// F(Q) -> G(r) transform on the FP32 matrix cores, and the FP64 potential
// (Rw / chi^2) reductions with their adjoint vectors.
//
// The reference transform (kernels/master_kernel.py:39-203) is an odd-extended
// zero-padded inverse FFT followed by linear interpolation onto the r grid;
// it is exactly G = T . F with a dense R x Q matrix T (built on the host in
// FP64, stored here as FP32).  grad G = grad F . T^T is the only dense
// contraction of the path and runs on v_mfma_f32_16x16x4_f32.
#include "internal.h"

namespace pyiid {
namespace {

using v4f = float __attribute__((ext_vector_type(4)));

// C[m][n] = sum_k A[m][k] * B[n][k]; block tile 64 x 64, 4 waves of 32 x 32,
// each wave 2 x 2 MFMA 16x16x4 tiles; K staged through LDS 16 at a time.
// MFMA f32 16x16x4 operand layout (lane l): A[l%16][l/16], B[l/16][l%16];
// result register v holds C[4*(l/16)+v][l%16].
template <bool A64>
__global__ __launch_bounds__(256) void gemm_nt_kernel(const void* __restrict__ a_raw, int64_t m,
                                                      int k, const float* __restrict__ b, int n,
                                                      double* __restrict__ c64,
                                                      float* __restrict__ c32) {
  constexpr int KS = 16;
  __shared__ float as[64][KS + 1];
  __shared__ float bs[64][KS + 1];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int64_t bm = static_cast<int64_t>(blockIdx.y) * 64;
  const int bn = blockIdx.x * 64;
  v4f acc[2][2];
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y) acc[x][y] = v4f{0.f, 0.f, 0.f, 0.f};

  for (int k0 = 0; k0 < k; k0 += KS) {
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int idx = threadIdx.x + e * 256;
      const int row = idx / KS, kk = idx % KS;
      const int64_t ar = bm + row;
      float av = 0.0f;
      if (ar < m && k0 + kk < k) {
        if constexpr (A64)
          av = static_cast<float>(static_cast<const double*>(a_raw)[ar * k + k0 + kk]);
        else
          av = static_cast<const float*>(a_raw)[ar * k + k0 + kk];
      }
      as[row][kk] = av;
      const int br = bn + row;
      bs[row][kk] = (br < n && k0 + kk < k) ? b[(int64_t)br * k + k0 + kk] : 0.0f;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < KS; kk += 4) {
      float af[2], bf[2];
#pragma unroll
      for (int x = 0; x < 2; ++x) af[x] = as[wm * 32 + x * 16 + (lane & 15)][kk + (lane >> 4)];
#pragma unroll
      for (int y = 0; y < 2; ++y) bf[y] = bs[wn * 32 + y * 16 + (lane & 15)][kk + (lane >> 4)];
#pragma unroll
      for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
          acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[x], bf[y], acc[x][y], 0, 0, 0);
    }
  }
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t row = bm + wm * 32 + x * 16 + 4 * (lane >> 4) + v;
        const int col = bn + wn * 32 + y * 16 + (lane & 15);
        if (row < m && col < n) {
          if (c64) c64[row * n + col] = static_cast<double>(acc[x][y][v]);
          if (c32) c32[row * n + col] = acc[x][y][v];
        }
      }
}

// The same product for tall A (the gradient rows): block tile 128 x 128,
// 4 waves of 64 x 64 (4 x 4 MFMA 16x16x4 tiles), K in chunks of 16 through
// double-buffered LDS; the next chunk's global loads are in flight during the
// current chunk's MFMAs, one barrier per chunk.  Blocks are dispatched round
// robin to the 8 XCDs (block id mod 8); the block -> tile map gives each XCD
// its own column tiles, so the rows of B (the transform) that it reads stay
// in that XCD's L2 while it sweeps the row tiles.
constexpr int kG2 = 128, kG2K = 16, kG2Ld = kG2K + 1;

template <bool A64>
__global__ __launch_bounds__(256) void gemm_nt128_kernel(const void* __restrict__ a_raw,
                                                         int64_t m, int k,
                                                         const float* __restrict__ b, int n,
                                                         int n_ct, int64_t n_rt,
                                                         double* __restrict__ c64,
                                                         float* __restrict__ c32) {
  __shared__ float as[2][kG2 * kG2Ld];
  __shared__ float bs[2][kG2 * kG2Ld];
  const int64_t bid = blockIdx.x;
  const int ct_per = (n_ct + 7) / 8;
  const int ct = static_cast<int>(bid % 8) + 8 * static_cast<int>((bid / 8) % ct_per);
  const int64_t rt = (bid / 8) / ct_per;
  if (ct >= n_ct || rt >= n_rt) return;  // padding block (uniform, before any barrier)
  const int64_t bm = rt * kG2;
  const int bn = ct * kG2;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wm = wave >> 1, wn = wave & 1;

  float ra[8], rb[8];
  auto load = [&](int k0) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int idx = threadIdx.x + e * 256;
      const int row = idx / kG2K, kk = idx % kG2K;
      const int64_t ar = bm + row;
      float av = 0.0f;
      if (ar < m && k0 + kk < k) {
        if constexpr (A64)
          av = static_cast<float>(static_cast<const double*>(a_raw)[ar * k + k0 + kk]);
        else
          av = static_cast<const float*>(a_raw)[ar * k + k0 + kk];
      }
      ra[e] = av;
      const int br = bn + row;
      rb[e] = (br < n && k0 + kk < k) ? b[static_cast<int64_t>(br) * k + k0 + kk] : 0.0f;
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int idx = threadIdx.x + e * 256;
      const int row = idx / kG2K, kk = idx % kG2K;
      as[buf][row * kG2Ld + kk] = ra[e];
      bs[buf][row * kG2Ld + kk] = rb[e];
    }
  };

  v4f acc[4][4];
#pragma unroll
  for (int x = 0; x < 4; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y) acc[x][y] = v4f{0.f, 0.f, 0.f, 0.f};

  const int nk = (k + kG2K - 1) / kG2K;
  load(0);
  store(0);
  __syncthreads();
  for (int c = 0; c < nk; ++c) {
    const int cur = c & 1;
    if (c + 1 < nk) load((c + 1) * kG2K);
    const float* a_s = as[cur];
    const float* b_s = bs[cur];
#pragma unroll
    for (int kk = 0; kk < kG2K; kk += 4) {
      float af[4], bf[4];
#pragma unroll
      for (int x = 0; x < 4; ++x)
        af[x] = a_s[(wm * 64 + x * 16 + (lane & 15)) * kG2Ld + kk + (lane >> 4)];
#pragma unroll
      for (int y = 0; y < 4; ++y)
        bf[y] = b_s[(wn * 64 + y * 16 + (lane & 15)) * kG2Ld + kk + (lane >> 4)];
#pragma unroll
      for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y)
          acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[x], bf[y], acc[x][y], 0, 0, 0);
    }
    if (c + 1 < nk) store(cur ^ 1);
    __syncthreads();
  }
#pragma unroll
  for (int x = 0; x < 4; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t row = bm + wm * 64 + x * 16 + 4 * (lane >> 4) + v;
        const int col = bn + wn * 64 + y * 16 + (lane & 15);
        if (row < m && col < n) {
          if (c64) c64[row * n + col] = static_cast<double>(acc[x][y][v]);
          if (c32) c32[row * n + col] = acc[x][y][v];
        }
      }
}

__device__ double block_sum(double v, double* sh) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  double t = 0.0;
  const int nw = blockDim.x >> 6;
  for (int w = 0; w < nw; ++w) t += sh[w];
  return t;
}

// One block per chain.  Rw: master_kernel.py:206-236; chi^2: :239-266; the adjoint
// w is d(potential)/d(gcalc) as get_grad_rw / get_grad_chi_sq apply it
// (:293-375): grad = gradG . w with
//   w = c * (s * res + (g.res)/(g.g) * (o - 2 a g)),
//   c = -value/|res|^2 (Rw) or -2 (chi^2), res = o - s g, a = optimal scale.
__global__ __launch_bounds__(1024) void potential_kernel(int potential,
                                                         const double* __restrict__ o_all,
                                                         const double* __restrict__ g_all, int n,
                                                         double* __restrict__ out5_all,
                                                         double* __restrict__ w_all) {
  __shared__ double sh[16];
  // one block per chain
  const double* o = o_all + (size_t)blockIdx.x * n;
  const double* g = g_all + (size_t)blockIdx.x * n;
  double* out5 = out5_all + (size_t)blockIdx.x * 5;
  double* w = w_all ? w_all + (size_t)blockIdx.x * n : nullptr;
  double go = 0.0, gg = 0.0, oo = 0.0;
  for (int t = threadIdx.x; t < n; t += blockDim.x) {
    go += g[t] * o[t];
    gg += g[t] * g[t];
    oo += o[t] * o[t];
  }
  go = block_sum(go, sh);
  gg = block_sum(gg, sh);
  oo = block_sum(oo, sh);
  const double a = go / gg;
  const double s = (a <= 0.0) ? 1.0 : a;
  double rr = 0.0, gr = 0.0;
  for (int t = threadIdx.x; t < n; t += blockDim.x) {
    const double res = o[t] - s * g[t];
    rr += res * res;
    gr += g[t] * res;
  }
  rr = block_sum(rr, sh);
  gr = block_sum(gr, sh);
  double value, coef;
  if (potential == 0) {
    value = (a <= 0.0) ? 1.0 : sqrt(rr / oo);
    coef = -value / rr;
  } else {
    value = rr;
    coef = -2.0;
  }
  if (w) {
    const double gfac = gr / gg;
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
      const double res = o[t] - s * g[t];
      w[t] = coef * (s * res + gfac * (o[t] - 2.0 * a * g[t]));
    }
  }
  if (threadIdx.x == 0) {
    out5[0] = value;
    out5[1] = s;
    out5[2] = a;
    out5[3] = rr;
    out5[4] = gr;
  }
}

__global__ void gemv_rows_kernel(const double* __restrict__ x, int64_t rows, int n,
                                 const double* __restrict__ w, double* __restrict__ y) {
  const int lane = threadIdx.x & 63;
  const int64_t row = blockIdx.x * (int64_t)(blockDim.x >> 6) + (threadIdx.x >> 6);
  if (row >= rows) return;
  double acc = 0.0;
  for (int t = lane; t < n; t += 64) acc += x[row * n + t] * w[t];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if (lane == 0) y[row] = acc;
}

// u[b][q] = gscale[b][q] * sum_t T[t][q] w[b][t]  (FP64 accumulation; T float).
// A block takes 64 bins of one chain (lane = bin, coalesced rows of T) with
// kAdjWaves waves, each summing a contiguous slice of the rows into four
// independent accumulators so that its loads overlap; the waves' partials
// are added in wave order through LDS (deterministic).
constexpr int kAdjWaves = 16;
__global__ __launch_bounds__(64 * kAdjWaves) void adjoint_kernel(
    const float* __restrict__ t, int n_r, int nq, const double* __restrict__ w,
    const double* __restrict__ gscale, float* __restrict__ u) {
  __shared__ double part[kAdjWaves][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int q = blockIdx.x * 64 + lane;
  const int b = blockIdx.y;
  const double* wb = w + (size_t)b * n_r;
  const int per = (n_r + kAdjWaves - 1) / kAdjWaves;
  const int r0 = wave * per, r1 = min(n_r, r0 + per);
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  if (q < nq) {
    int r = r0;
    for (; r + 4 <= r1; r += 4) {
      a0 += static_cast<double>(t[(size_t)r * nq + q]) * wb[r];
      a1 += static_cast<double>(t[(size_t)(r + 1) * nq + q]) * wb[r + 1];
      a2 += static_cast<double>(t[(size_t)(r + 2) * nq + q]) * wb[r + 2];
      a3 += static_cast<double>(t[(size_t)(r + 3) * nq + q]) * wb[r + 3];
    }
    for (; r < r1; ++r) a0 += static_cast<double>(t[(size_t)r * nq + q]) * wb[r];
  }
  part[wave][lane] = (a0 + a1) + (a2 + a3);
  __syncthreads();
  if (wave != 0 || q >= nq) return;
  double acc = 0.0;
#pragma unroll
  for (int k = 0; k < kAdjWaves; ++k) acc += part[k][lane];
  u[(size_t)b * nq + q] = static_cast<float>(acc * gscale[(size_t)b * nq + q]);
}

__global__ void adjoint_fq_kernel(int nq, int q_first, const double* __restrict__ w,
                                  const double* __restrict__ gscale, float* __restrict__ u) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq) return;
  u[q] = (q >= q_first) ? static_cast<float>(w[q - q_first] * gscale[q]) : 0.0f;
}

__global__ void f64_to_f32_kernel(const double* __restrict__ x, int64_t n, float* __restrict__ y) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x)
    y[e] = static_cast<float>(x[e]);
}

__global__ void f32_to_f64_kernel(const float* __restrict__ x, int64_t n, double* __restrict__ y) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x)
    y[e] = static_cast<double>(x[e]);
}

int blocks_for(int64_t total, int block) {
  int64_t g = (total + block - 1) / block;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return static_cast<int>(g);
}

}  // namespace

hipError_t launch_gemm_nt(hipStream_t st, const void* a, bool a_f64, int64_t m, int k,
                          const float* b, int n, double* c_f64, float* c_f32) {
  if (m >= 1024) {
    const int n_ct = (n + kG2 - 1) / kG2;
    const int64_t n_rt = (m + kG2 - 1) / kG2;
    const int64_t blocks = 8LL * ((n_ct + 7) / 8) * n_rt;
    if (blocks > 2147483647LL) return hipErrorInvalidValue;
    const dim3 grid(static_cast<unsigned>(blocks));
    if (a_f64)
      hipLaunchKernelGGL(gemm_nt128_kernel<true>, grid, dim3(256), 0, st, a, m, k, b, n, n_ct,
                         n_rt, c_f64, c_f32);
    else
      hipLaunchKernelGGL(gemm_nt128_kernel<false>, grid, dim3(256), 0, st, a, m, k, b, n, n_ct,
                         n_rt, c_f64, c_f32);
    return hipGetLastError();
  }
  const int64_t gy = (m + 63) / 64;
  if (gy > 2147483647LL) return hipErrorInvalidValue;
  dim3 grid((n + 63) / 64, static_cast<unsigned>(gy));
  if (a_f64)
    hipLaunchKernelGGL(gemm_nt_kernel<true>, grid, dim3(256), 0, st, a, m, k, b, n, c_f64, c_f32);
  else
    hipLaunchKernelGGL(gemm_nt_kernel<false>, grid, dim3(256), 0, st, a, m, k, b, n, c_f64, c_f32);
  return hipGetLastError();
}

hipError_t launch_potential(hipStream_t st, int potential, const double* gobs,
                            const double* gcalc, int n, int n_chain, double* out5, double* w) {
  hipLaunchKernelGGL(potential_kernel, dim3(n_chain), dim3(1024), 0, st, potential, gobs, gcalc, n,
                     out5, w);
  return hipGetLastError();
}

hipError_t launch_gemv_rows(hipStream_t st, const double* x, int64_t rows, int n, const double* w,
                            double* y) {
  const int64_t blocks = (rows + 3) / 4;
  hipLaunchKernelGGL(gemv_rows_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, st, x,
                     rows, n, w, y);
  return hipGetLastError();
}

hipError_t launch_adjoint(hipStream_t st, const float* t, int n_r, int nq, const double* w,
                          const double* gscale, int n_chain, float* u) {
  hipLaunchKernelGGL(adjoint_kernel, dim3((nq + 63) / 64, n_chain), dim3(64 * kAdjWaves), 0, st, t,
                     n_r, nq, w, gscale, u);
  return hipGetLastError();
}

hipError_t launch_adjoint_fq(hipStream_t st, int nq, int q_first, const double* w,
                             const double* gscale, float* u) {
  hipLaunchKernelGGL(adjoint_fq_kernel, dim3((nq + 255) / 256), dim3(256), 0, st, nq, q_first, w,
                     gscale, u);
  return hipGetLastError();
}

hipError_t launch_f64_to_f32(hipStream_t st, const double* x, int64_t n, float* y) {
  hipLaunchKernelGGL(f64_to_f32_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, st, x, n, y);
  return hipGetLastError();
}

hipError_t launch_f32_to_f64(hipStream_t st, const float* x, int64_t n, double* y) {
  hipLaunchKernelGGL(f32_to_f64_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, st, x, n, y);
  return hipGetLastError();
}

}  // namespace pyiid
