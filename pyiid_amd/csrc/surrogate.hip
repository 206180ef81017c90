// Learned HMC integrator for the batched sampler: a per-atom MLP on radial
// neighbour descriptors whose summed output is a surrogate energy; its
// gradient drives the leapfrog while the exact Debye Rw decides every
// Metropolis step.  No counterpart exists in pyIID (pyiid/sim/hmc.py drives
// the leapfrog with exact forces only).
//
// Descriptors: D[i][m] = sum_{j != i, r_ij < rc} exp(-eta (r_ij - mu_m)^2) fc(r_ij),
// m < 16, fc(r) = (cos(pi r / rc) + 1) / 2, mu_m = mu0 + m dmu.  Neighbour lists
// are built per chain with radius rc + skin from the proposal's start.
// Network (per atom): x = (D - mean) * inv_scale -> 64 SiLU -> 64 SiLU -> 1,
// BF16 weights and activations into v_mfma_f32_16x16x16_bf16, FP32
// accumulation; the output weights w3 are per chain.  Energy E = sum_i e_i.
// Force on atom a: F_a = -sum_{j in nbr(a)} sum_m (c_a,m + c_j,m) g'_m(r_aj) u_aj,
// c = dE/dD, u_aj = (x_a - x_j) / r_aj.
#include "internal.h"

namespace pyiid {
namespace {

using bf16x4 = short __attribute__((ext_vector_type(4)));
using v4f = float __attribute__((ext_vector_type(4)));

constexpr float kPiF = 3.14159265358979323846f;

__device__ __forceinline__ unsigned short f2bf(float x) {
  // round to nearest even (inputs are finite)
  unsigned u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return static_cast<unsigned short>(u >> 16);
}

__device__ __forceinline__ float bf2f(unsigned short h) {
  return __uint_as_float(static_cast<unsigned>(h) << 16);
}

__device__ __forceinline__ bf16x4 pack4(float a, float b, float c, float d) {
  bf16x4 v;
  v[0] = static_cast<short>(f2bf(a));
  v[1] = static_cast<short>(f2bf(b));
  v[2] = static_cast<short>(f2bf(c));
  v[3] = static_cast<short>(f2bf(d));
  return v;
}

__device__ __forceinline__ bf16x4 load4(const unsigned short* p) {
  bf16x4 v;
  v[0] = static_cast<short>(p[0]);
  v[1] = static_cast<short>(p[1]);
  v[2] = static_cast<short>(p[2]);
  v[3] = static_cast<short>(p[3]);
  return v;
}

// Candidates within rl of atom i among the atoms of its chain (brute force
// over LDS-staged tiles; O(n_b^2) per chain, a fraction of one Debye pass).
// Neighbour k of atom i is nbr[k * n + i] (coalesced across atoms).
__global__ __launch_bounds__(256) void nbr_build_kernel(const AtomRec* __restrict__ pos, int n,
                                                        const Item* __restrict__ tiles, float rl2,
                                                        int max_nbr, int* __restrict__ nbr,
                                                        int* __restrict__ cnt,
                                                        int* __restrict__ viol) {
  __shared__ float4 sj[256];
  const Item t = tiles[blockIdx.x];
  const int i = t.i_base + static_cast<int>(threadIdx.x);
  const bool valid = i < t.i_end;
  float xi = 0.f, yi = 0.f, zi = 0.f;
  if (valid) {
    const AtomRec a = pos[i];
    xi = static_cast<float>(a.x);
    yi = static_cast<float>(a.y);
    zi = static_cast<float>(a.z);
  }
  int c = 0;
  for (int j0 = t.jbeg; j0 < t.jend; j0 += 256) {
    __syncthreads();
    const int j = j0 + static_cast<int>(threadIdx.x);
    if (j < t.jend) {
      const AtomRec b = pos[j];
      sj[threadIdx.x] = make_float4(static_cast<float>(b.x), static_cast<float>(b.y),
                                    static_cast<float>(b.z), 0.f);
    }
    __syncthreads();
    const int m = min(256, t.jend - j0);
    if (valid) {
      for (int jj = 0; jj < m; ++jj) {
        const float4 p = sj[jj];
        const float dx = xi - p.x, dy = yi - p.y, dz = zi - p.z;
        const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
        if (r2 < rl2 && j0 + jj != i) {
          if (c < max_nbr) nbr[static_cast<size_t>(c) * n + i] = j0 + jj;
          ++c;
        }
      }
    }
  }
  if (valid) {
    cnt[i] = min(c, max_nbr);
    if (c > max_nbr) viol[t.chain] = 1;
  }
}

struct Rbf {
  float rc, mu0, dmu, eta;
};

__global__ __launch_bounds__(256) void descriptor_kernel(const AtomRec* __restrict__ pos, int n,
                                                         const int* __restrict__ nbr,
                                                         const int* __restrict__ cnt, Rbf g,
                                                         float* __restrict__ d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const AtomRec a = pos[i];
  float acc[16];
#pragma unroll
  for (int m = 0; m < 16; ++m) acc[m] = 0.f;
  const int nn = cnt[i];
  for (int k = 0; k < nn; ++k) {
    const AtomRec b = pos[nbr[static_cast<size_t>(k) * n + i]];
    const float dx = static_cast<float>(a.x - b.x);
    const float dy = static_cast<float>(a.y - b.y);
    const float dz = static_cast<float>(a.z - b.z);
    const float r = sqrtf(fmaf(dx, dx, fmaf(dy, dy, dz * dz)));
    if (r < g.rc) {
      const float fc = 0.5f * (cosf(kPiF * r / g.rc) + 1.0f);
#pragma unroll
      for (int m = 0; m < 16; ++m) {
        const float x = r - fmaf(static_cast<float>(m), g.dmu, g.mu0);
        acc[m] = fmaf(expf(-g.eta * x * x), fc, acc[m]);
      }
    }
  }
  float4* out = reinterpret_cast<float4*>(d + static_cast<size_t>(i) * 16);
#pragma unroll
  for (int q = 0; q < 4; ++q)
    out[q] = make_float4(acc[4 * q], acc[4 * q + 1], acc[4 * q + 2], acc[4 * q + 3]);
}

// Network parameters in device memory.
struct Net {
  unsigned short w1[16 * 64];  // [m][j]
  unsigned short w2[64 * 64];  // [j][k]
  float b1[64], b2[64];
  float mean[16], inv_scale[16];
};

__device__ __forceinline__ void silu(float z, float& h, float& dh) {
  const float s = 1.0f / (1.0f + expf(-z));
  h = z * s;
  dh = s * (1.0f + z * (1.0f - s));
}

// One wave per 16 atoms of a 64-atom tile (4 waves).  v_mfma_f32_16x16x16_bf16
// operand layout (lane l, element e): A[l%16][4(l/16)+e], B[4(l/16)+e][l%16];
// result register v holds C[4(l/16)+v][l%16].  Activations change from the
// result layout to the A layout through a per-wave LDS tile.
// Output: c[i][m] = dE/dD[i][m]; e_atom[i] (nullable).  w3 row of chain b at
// w3 + b * w3_stride (stride 0: one row for all chains).
__global__ __launch_bounds__(256) void mlp_kernel(const float* __restrict__ d,
                                                  const Item* __restrict__ tiles,
                                                  const Net* __restrict__ net,
                                                  const unsigned short* __restrict__ w3,
                                                  int w3_stride, float* __restrict__ c_out,
                                                  float* __restrict__ e_atom) {
  constexpr int kRow = 64 + 8;  // bf16 per LDS row (16-byte padding)
  __shared__ __align__(16) unsigned short sh[4][16][kRow];
  const Item t = tiles[blockIdx.x];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r16 = lane & 15, g4 = lane >> 4;
  unsigned short(*buf)[kRow] = sh[wave];
  const unsigned short* w3c = w3 + static_cast<size_t>(t.chain) * w3_stride;

  // input rows: atom base + r16, features 4 g4 .. 4 g4 + 3
  const int atom_a = t.i_base + 16 * wave + r16;
  float xin[4] = {0.f, 0.f, 0.f, 0.f};
  if (atom_a < t.i_end) {
    const float4 v = *reinterpret_cast<const float4*>(d + static_cast<size_t>(atom_a) * 16 + 4 * g4);
    xin[0] = v.x;
    xin[1] = v.y;
    xin[2] = v.z;
    xin[3] = v.w;
  }
#pragma unroll
  for (int e = 0; e < 4; ++e) xin[e] = (xin[e] - net->mean[4 * g4 + e]) * net->inv_scale[4 * g4 + e];
  const bf16x4 a1 = pack4(xin[0], xin[1], xin[2], xin[3]);

  // layer 1: Z1[16 x 64] = X[16 x 16] W1[16 x 64]
  float s1[4][4];
#pragma unroll
  for (int tt = 0; tt < 4; ++tt) {
    bf16x4 b;
#pragma unroll
    for (int e = 0; e < 4; ++e) b[e] = static_cast<short>(net->w1[(4 * g4 + e) * 64 + 16 * tt + r16]);
    const v4f z = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1, b, v4f{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
    const float bias = net->b1[16 * tt + r16];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      float h;
      silu(z[v] + bias, h, s1[tt][v]);
      buf[4 * g4 + v][16 * tt + r16] = f2bf(h);
    }
  }
  __syncthreads();

  // layer 2: Z2 = H1 W2
  v4f z2[4];
#pragma unroll
  for (int tt = 0; tt < 4; ++tt) z2[tt] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const bf16x4 a = load4(&buf[r16][16 * s + 4 * g4]);
#pragma unroll
    for (int tt = 0; tt < 4; ++tt) {
      bf16x4 b;
#pragma unroll
      for (int e = 0; e < 4; ++e)
        b[e] = static_cast<short>(net->w2[(16 * s + 4 * g4 + e) * 64 + 16 * tt + r16]);
      z2[tt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, z2[tt], 0, 0, 0);
    }
  }
  __syncthreads();

  // output layer and its gradient seed g2 = w3 * silu'(Z2)
  float ep[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int tt = 0; tt < 4; ++tt) {
    const float bias = net->b2[16 * tt + r16];
    const float w = bf2f(w3c[16 * tt + r16]);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      float h, dh;
      silu(z2[tt][v] + bias, h, dh);
      ep[v] = fmaf(h, w, ep[v]);
      buf[4 * g4 + v][16 * tt + r16] = f2bf(w * dh);
    }
  }
  if (e_atom) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      float x = ep[v];
      x += __shfl_xor(x, 1, 64);
      x += __shfl_xor(x, 2, 64);
      x += __shfl_xor(x, 4, 64);
      x += __shfl_xor(x, 8, 64);
      const int atom = t.i_base + 16 * wave + 4 * g4 + v;
      if (r16 == 0 && atom < t.i_end) e_atom[atom] = x;
    }
  }
  __syncthreads();

  // dH1 = G2 W2^T, G1 = dH1 * silu'(Z1)
  v4f dh1[4];
#pragma unroll
  for (int tt = 0; tt < 4; ++tt) dh1[tt] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const bf16x4 a = load4(&buf[r16][16 * s + 4 * g4]);
#pragma unroll
    for (int tt = 0; tt < 4; ++tt) {
      bf16x4 b;
#pragma unroll
      for (int e = 0; e < 4; ++e)
        b[e] = static_cast<short>(net->w2[(16 * tt + r16) * 64 + 16 * s + 4 * g4 + e]);
      dh1[tt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, dh1[tt], 0, 0, 0);
    }
  }
  __syncthreads();
#pragma unroll
  for (int tt = 0; tt < 4; ++tt)
#pragma unroll
    for (int v = 0; v < 4; ++v) buf[4 * g4 + v][16 * tt + r16] = f2bf(dh1[tt][v] * s1[tt][v]);
  __syncthreads();

  // dX = G1 W1^T, c = dX * inv_scale
  v4f dx = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const bf16x4 a = load4(&buf[r16][16 * s + 4 * g4]);
    bf16x4 b;
#pragma unroll
    for (int e = 0; e < 4; ++e) b[e] = static_cast<short>(net->w1[r16 * 64 + 16 * s + 4 * g4 + e]);
    dx = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, dx, 0, 0, 0);
  }
  const float isc = net->inv_scale[r16];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int atom = t.i_base + 16 * wave + 4 * g4 + v;
    if (atom < t.i_end) c_out[static_cast<size_t>(atom) * 16 + r16] = dx[v] * isc;
  }
}

// Forces from c = dE/dD (force_out, double n x 3) or, with phi_out set, the
// same sum written as float (a feature force column of the fit).
__global__ __launch_bounds__(256) void sur_force_kernel(const AtomRec* __restrict__ pos, int n,
                                                        const int* __restrict__ nbr,
                                                        const int* __restrict__ cnt, Rbf g,
                                                        const float* __restrict__ c,
                                                        double* __restrict__ force_out,
                                                        float* __restrict__ phi_out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const AtomRec a = pos[i];
  float ca[16];
  {
    const float4* p = reinterpret_cast<const float4*>(c + static_cast<size_t>(i) * 16);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float4 v = p[q];
      ca[4 * q] = v.x;
      ca[4 * q + 1] = v.y;
      ca[4 * q + 2] = v.z;
      ca[4 * q + 3] = v.w;
    }
  }
  float fx = 0.f, fy = 0.f, fz = 0.f;
  const float k_arg = kPiF / g.rc;
  const int nn = cnt[i];
  for (int k = 0; k < nn; ++k) {
    const int j = nbr[static_cast<size_t>(k) * n + i];
    const AtomRec b = pos[j];
    const float dx = static_cast<float>(a.x - b.x);
    const float dy = static_cast<float>(a.y - b.y);
    const float dz = static_cast<float>(a.z - b.z);
    const float r = sqrtf(fmaf(dx, dx, fmaf(dy, dy, dz * dz)));
    if (r >= g.rc) continue;
    float sn, cs;
    sincosf(k_arg * r, &sn, &cs);
    const float fc = 0.5f * (cs + 1.0f);
    const float dfc = -0.5f * k_arg * sn;
    const float4* cj = reinterpret_cast<const float4*>(c + static_cast<size_t>(j) * 16);
    float s = 0.f;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float4 v = cj[q];
      const float cjv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int m = 4 * q + e;
        const float x = r - fmaf(static_cast<float>(m), g.dmu, g.mu0);
        const float ex = expf(-g.eta * x * x);
        const float dg = ex * fmaf(-2.0f * g.eta * x, fc, dfc);
        s = fmaf(ca[m] + cjv[e], dg, s);
      }
    }
    const float sr = s / r;
    fx = fmaf(-sr, dx, fx);
    fy = fmaf(-sr, dy, fy);
    fz = fmaf(-sr, dz, fz);
  }
  if (phi_out) {
    phi_out[3 * static_cast<size_t>(i) + 0] = fx;
    phi_out[3 * static_cast<size_t>(i) + 1] = fy;
    phi_out[3 * static_cast<size_t>(i) + 2] = fz;
  } else {
    force_out[3 * static_cast<size_t>(i) + 0] = fx;
    force_out[3 * static_cast<size_t>(i) + 1] = fy;
    force_out[3 * static_cast<size_t>(i) + 2] = fz;
  }
}

// Partial normal equations of the output-layer fit over one column chunk
// [jbeg, jend) of chain item.chain: part[split] = {Phi Phi^T (64 x 64),
// Phi y (64), y.y}.  Phi is 64 x n3 (row stride n3), y the target forces.
constexpr int kGramStride = 64 * 64 + 64 + 1;

__global__ __launch_bounds__(256) void gram_kernel(const float* __restrict__ phi, int64_t n3,
                                                   const double* __restrict__ y,
                                                   const Item* __restrict__ chunks,
                                                   float* __restrict__ part) {
  constexpr int KC = 32;
  __shared__ float sp[64][KC + 1];
  __shared__ float sy[KC];
  const Item t = chunks[blockIdx.x];
  const int k = threadIdx.x >> 2, l0 = (threadIdx.x & 3) * 16;
  float acc[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) acc[j] = 0.f;
  float racc = 0.f, yy = 0.f;
  for (int c0 = t.jbeg; c0 < t.jend; c0 += KC) {
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int idx = threadIdx.x + 256 * e;
      const int row = idx / KC, cc = idx % KC;
      sp[row][cc] = (c0 + cc < t.jend) ? phi[static_cast<int64_t>(row) * n3 + c0 + cc] : 0.f;
    }
    if (threadIdx.x < KC)
      sy[threadIdx.x] = (c0 + static_cast<int>(threadIdx.x) < t.jend)
                            ? static_cast<float>(y[c0 + threadIdx.x])
                            : 0.f;
    __syncthreads();
#pragma unroll 4
    for (int cc = 0; cc < KC; ++cc) {
      const float a = sp[k][cc];
#pragma unroll
      for (int j = 0; j < 16; ++j) acc[j] = fmaf(a, sp[l0 + j][cc], acc[j]);
      if (threadIdx.x < 64) racc = fmaf(sp[threadIdx.x][cc], sy[cc], racc);
      if (threadIdx.x == 64) yy = fmaf(sy[cc], sy[cc], yy);
    }
  }
  float* out = part + static_cast<size_t>(t.split) * kGramStride;
#pragma unroll
  for (int j = 0; j < 16; ++j) out[k * 64 + l0 + j] = acc[j];
  if (threadIdx.x < 64) out[4096 + threadIdx.x] = racc;
  if (threadIdx.x == 64) out[4160] = yy;
}

// sum of a chain's partials in FP64: out[b] = sum_{s in ranges[b]} part[s]
__global__ __launch_bounds__(256) void gram_reduce_kernel(const float* __restrict__ part,
                                                          const Range* __restrict__ ranges,
                                                          double* __restrict__ out) {
  const Range r = ranges[blockIdx.x];
  for (int e = threadIdx.x; e < kGramStride; e += blockDim.x) {
    double s = 0.0;
    for (int p = r.begin; p < r.end; ++p) s += static_cast<double>(part[static_cast<size_t>(p) * kGramStride + e]);
    out[static_cast<size_t>(blockIdx.x) * kGramStride + e] = s;
  }
}

// Feature energies of the fit: out[b][k] = sum over chain b's atoms of
// e_atom[k][a] (the per-atom outputs of the network with w3 = e_k), in FP64
// and a fixed order.  Block (b, k); 256 threads.
__global__ __launch_bounds__(256) void chain_feature_energy_kernel(const float* __restrict__ e_atom,
                                                                   int64_t n,
                                                                   const int64_t* __restrict__ off,
                                                                   double* __restrict__ out) {
  __shared__ double red[256];
  const int b = blockIdx.x, k = blockIdx.y;
  const float* row = e_atom + static_cast<size_t>(k) * n;
  double sum = 0.0;
  for (int64_t a = off[b] + threadIdx.x; a < off[b + 1]; a += 256) sum += static_cast<double>(row[a]);
  red[threadIdx.x] = sum;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[static_cast<size_t>(b) * 64 + k] = red[0];
}

// Flags chains with an atom farther than sqrt(lim2) from its proposal start
// (the neighbour list would no longer hold every pair inside the cutoff).
__global__ void disp_check_kernel(const AtomRec* __restrict__ pos,
                                  const double* __restrict__ pos_save,
                                  const int* __restrict__ atom_chain, int64_t n, double lim2,
                                  int* __restrict__ viol) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const AtomRec a = pos[i];
  const double dx = a.x - pos_save[3 * i], dy = a.y - pos_save[3 * i + 1],
               dz = a.z - pos_save[3 * i + 2];
  if (dx * dx + dy * dy + dz * dz > lim2) viol[atom_chain[i]] = 1;
}

}  // namespace

static_assert(sizeof(Net) == kSurNetBytes, "Net layout");
static_assert(kGramStride == kSurGram, "Gram layout");

hipError_t launch_nbr_build(hipStream_t st, const AtomRec* pos, int n, const Item* tiles,
                            int n_tiles, float rl, int max_nbr, int* nbr, int* cnt, int* viol) {
  if (n_tiles == 0) return hipSuccess;
  hipLaunchKernelGGL(nbr_build_kernel, dim3(n_tiles), dim3(256), 0, st, pos, n, tiles, rl * rl,
                     max_nbr, nbr, cnt, viol);
  return hipGetLastError();
}

hipError_t launch_descriptors(hipStream_t st, const AtomRec* pos, int n, const int* nbr,
                              const int* cnt, const float* rbf, float* d) {
  if (n == 0) return hipSuccess;
  const Rbf g{rbf[0], rbf[1], rbf[2], rbf[3]};
  hipLaunchKernelGGL(descriptor_kernel, dim3((n + 255) / 256), dim3(256), 0, st, pos, n, nbr, cnt,
                     g, d);
  return hipGetLastError();
}

hipError_t launch_mlp(hipStream_t st, const float* d, int, const Item* tiles, int n_tiles,
                      const void* net, const unsigned short* w3, int w3_stride, float* c,
                      float* e_atom) {
  if (n_tiles == 0) return hipSuccess;
  hipLaunchKernelGGL(mlp_kernel, dim3(n_tiles), dim3(256), 0, st, d, tiles,
                     static_cast<const Net*>(net), w3, w3_stride, c, e_atom);
  return hipGetLastError();
}

hipError_t launch_sur_force(hipStream_t st, const AtomRec* pos, int n, const int* nbr,
                            const int* cnt, const float* rbf, const float* c, double* force,
                            float* phi) {
  if (n == 0) return hipSuccess;
  const Rbf g{rbf[0], rbf[1], rbf[2], rbf[3]};
  hipLaunchKernelGGL(sur_force_kernel, dim3((n + 255) / 256), dim3(256), 0, st, pos, n, nbr, cnt,
                     g, c, force, phi);
  return hipGetLastError();
}

hipError_t launch_gram(hipStream_t st, const float* phi, int64_t n3, const double* y,
                       const Item* chunks, int n_chunks, float* part, const Range* ranges,
                       int n_chain, double* out) {
  if (n_chunks > 0) {
    hipLaunchKernelGGL(gram_kernel, dim3(n_chunks), dim3(256), 0, st, phi, n3, y, chunks, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(gram_reduce_kernel, dim3(n_chain), dim3(256), 0, st, part, ranges, out);
  return hipGetLastError();
}

hipError_t launch_chain_feature_energy(hipStream_t st, const float* e_atom, int64_t n,
                                       const int64_t* off, int n_chain, double* out) {
  if (n_chain == 0) return hipSuccess;
  hipLaunchKernelGGL(chain_feature_energy_kernel, dim3(n_chain, 64), dim3(256), 0, st, e_atom, n,
                     off, out);
  return hipGetLastError();
}

hipError_t launch_disp_check(hipStream_t st, const AtomRec* pos, const double* pos_save,
                             const int* atom_chain, int64_t n, double lim, int* viol) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(disp_check_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256),
                     0, st, pos, pos_save, atom_chain, n, lim * lim, viol);
  return hipGetLastError();
}

}  // namespace pyiid
