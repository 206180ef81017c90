// Device-resident Hamiltonian Monte Carlo for a batch of independent chains.
//
// One proposal = hmc_begin (save positions, draw momenta, K0, half kick),
// [drift, force evaluation, kick] x L, hmc_end (K1, Metropolis, accept or
// restore, record).  Velocity-Verlet ordering follows pyiid/sim/__init__.py:
// 10-38 (p += eps/2 F; x += eps p/m; p += eps/2 F(x')); acceptance follows
// pyiid/sim/hmc.py:66-92 (accept if u < exp(beta (H0 - H1)), beta being the
// reference's `temp` argument).  Every decision is taken on the device, so a
// proposal is a fixed sequence of launches (capturable in a HIP graph).
//
// Random numbers: splitmix64 (Steele, Lea, Flood 2014) hashes of (seed,
// proposal counter, atom, component) -> uniform doubles -> Box-Muller normals.
#include "internal.h"

namespace pyiid {
namespace {

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// uniform in (0, 1]
__device__ __forceinline__ double uniform_key(uint64_t seed, int64_t counter, uint64_t stream,
                                              uint64_t idx) {
  const uint64_t k = mix64(mix64(seed ^ mix64(static_cast<uint64_t>(counter) * 4 + stream)) + idx);
  return (static_cast<double>(k >> 11) + 1.0) * (1.0 / 9007199254740992.0);
}

__device__ __forceinline__ double normal_key(uint64_t seed, int64_t counter, uint64_t idx) {
  const double u1 = uniform_key(seed, counter, 0, idx);
  const double u2 = uniform_key(seed, counter, 1, idx);
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586477 * u2);
}

__device__ double block_sum256(double v, double* sh) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += sh[w];
  return t;
}

__global__ __launch_bounds__(256) void hmc_begin_kernel(
    AtomRec* __restrict__ pos, double* __restrict__ pos_save, double* __restrict__ mom,
    const double* __restrict__ force, const double* __restrict__ mass,
    const HmcChain* __restrict__ chains, double beta, double eps, uint64_t seed,
    const int64_t* __restrict__ counter, double* __restrict__ kin0) {
  __shared__ double sh[8];
  const HmcChain ch = chains[blockIdx.x];
  const int64_t cnt = *counter;
  double k = 0.0;
  for (int i = ch.atom_begin + threadIdx.x; i < ch.atom_end; i += blockDim.x) {
    const AtomRec a = pos[i];
    pos_save[3 * (size_t)i + 0] = a.x;
    pos_save[3 * (size_t)i + 1] = a.y;
    pos_save[3 * (size_t)i + 2] = a.z;
    const double m = mass[i];
    const double sd = sqrt(m / beta);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double p = sd * normal_key(seed, cnt, 3 * static_cast<uint64_t>(i) + c);
      k += p * p / (2.0 * m);
      mom[3 * (size_t)i + c] = p + 0.5 * eps * force[3 * (size_t)i + c];
    }
  }
  k = block_sum256(k, sh);
  if (threadIdx.x == 0) kin0[blockIdx.x] = k;
}

__global__ void hmc_drift_kernel(AtomRec* __restrict__ pos, const double* __restrict__ mom,
                                 const double* __restrict__ mass, int64_t n, double eps) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double s = eps / mass[i];
  pos[i].x += s * mom[3 * i + 0];
  pos[i].y += s * mom[3 * i + 1];
  pos[i].z += s * mom[3 * i + 2];
}

__global__ void hmc_kick_kernel(double* __restrict__ mom, const double* __restrict__ force,
                                int64_t n3, double h) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e < n3) mom[e] += h * force[e];
}

__global__ __launch_bounds__(256) void hmc_end_kernel(
    AtomRec* __restrict__ pos, const double* __restrict__ pos_save,
    const double* __restrict__ mom, const double* __restrict__ mass,
    double* __restrict__ force_cur, const double* __restrict__ force_new,
    double* __restrict__ e_cur, const double* __restrict__ out5_new, double conv,
    const HmcChain* __restrict__ chains, int n_chain, double beta, uint64_t seed,
    const int64_t* __restrict__ counter, const double* __restrict__ kin0,
    double* __restrict__ e_hist, int32_t* __restrict__ acc_hist, int64_t hist_cap,
    int* __restrict__ viol, int* __restrict__ viol_count) {
  __shared__ double sh[8];
  __shared__ int accept_sh;
  const int b = blockIdx.x;
  const HmcChain ch = chains[b];
  const int64_t cnt = *counter;
  double k = 0.0;
  for (int i = ch.atom_begin + threadIdx.x; i < ch.atom_end; i += blockDim.x)
    for (int c = 0; c < 3; ++c) {
      const double p = mom[3 * (size_t)i + c];
      k += p * p / (2.0 * mass[i]);
    }
  k = block_sum256(k, sh);
  const double e_new = conv * out5_new[5 * (size_t)b];
  if (threadIdx.x == 0) {
    const double h0 = e_cur[b] + kin0[b];
    const double h1 = e_new + k;
    const double u = uniform_key(seed, cnt, 2, static_cast<uint64_t>(b));
    int acc = (u < exp(beta * (h0 - h1))) ? 1 : 0;
    if (viol && viol[b]) {  // the proposal left its neighbour list's range
      acc = 0;
      viol[b] = 0;
      viol_count[b] += 1;
    }
    accept_sh = acc;
    if (acc) e_cur[b] = e_new;
    if (cnt < hist_cap) {
      e_hist[cnt * n_chain + b] = acc ? e_new : e_cur[b];
      acc_hist[cnt * n_chain + b] = acc;
    }
  }
  __syncthreads();
  const int acc = accept_sh;
  for (int i = ch.atom_begin + threadIdx.x; i < ch.atom_end; i += blockDim.x) {
    if (acc) {
      if (force_cur)
        for (int c = 0; c < 3; ++c) force_cur[3 * (size_t)i + c] = force_new[3 * (size_t)i + c];
    } else {
      pos[i].x = pos_save[3 * (size_t)i + 0];
      pos[i].y = pos_save[3 * (size_t)i + 1];
      pos[i].z = pos_save[3 * (size_t)i + 2];
    }
  }
}

__global__ void bump_kernel(int64_t* counter) { *counter += 1; }

__global__ void energy_kernel(const double* __restrict__ out5, int n_chain, double conv,
                              double* __restrict__ e) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < n_chain) e[b] = conv * out5[5 * (size_t)b];
}

}  // namespace

hipError_t launch_hmc_begin(hipStream_t st, AtomRec* pos, double* pos_save, double* mom,
                            const double* force, const double* mass, const HmcChain* chains,
                            int n_chain, double beta, double eps, uint64_t seed,
                            const int64_t* counter, double* kin0) {
  hipLaunchKernelGGL(hmc_begin_kernel, dim3(n_chain), dim3(256), 0, st, pos, pos_save, mom, force,
                     mass, chains, beta, eps, seed, counter, kin0);
  return hipGetLastError();
}

hipError_t launch_hmc_drift(hipStream_t st, AtomRec* pos, const double* mom, const double* mass,
                            int64_t n, double eps) {
  hipLaunchKernelGGL(hmc_drift_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0,
                     st, pos, mom, mass, n, eps);
  return hipGetLastError();
}

hipError_t launch_hmc_kick(hipStream_t st, double* mom, const double* force, int64_t n, double h) {
  const int64_t n3 = 3 * n;
  hipLaunchKernelGGL(hmc_kick_kernel, dim3(static_cast<unsigned>((n3 + 255) / 256)), dim3(256), 0,
                     st, mom, force, n3, h);
  return hipGetLastError();
}

hipError_t launch_hmc_end(hipStream_t st, AtomRec* pos, const double* pos_save, const double* mom,
                          const double* mass, double* force_cur, const double* force_new,
                          double* e_cur, const double* out5_new, double conv,
                          const HmcChain* chains, int n_chain, double beta, uint64_t seed,
                          int64_t* counter, const double* kin0, double* e_hist,
                          int32_t* acc_hist, int64_t hist_cap, int* viol, int* viol_count) {
  hipLaunchKernelGGL(hmc_end_kernel, dim3(n_chain), dim3(256), 0, st, pos, pos_save, mom, mass,
                     force_cur, force_new, e_cur, out5_new, conv, chains, n_chain, beta, seed,
                     counter, kin0, e_hist, acc_hist, hist_cap, viol, viol_count);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(bump_kernel, dim3(1), dim3(1), 0, st, counter);
  return hipGetLastError();
}

hipError_t launch_energy_from_out5(hipStream_t st, const double* out5, int n_chain, double conv,
                                   double* e) {
  hipLaunchKernelGGL(energy_kernel, dim3((n_chain + 255) / 256), dim3(256), 0, st, out5, n_chain,
                     conv, e);
  return hipGetLastError();
}

}  // namespace pyiid
