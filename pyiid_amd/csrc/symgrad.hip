// Symmetric F + gradient kernel for gfx950 (CDNA4, wave64): every unordered
// pair of atoms is formed once and feeds both of its atoms' gradient rows.
//
//  * Positions are in a permutation of the atoms (element order, then
//    Z-order) cut into tiles of at most 64 positions (of one element).  A
//    work item is a rectangle of the lower triangle of tile pairs: an I
//    group [g0, g1) of tiles against a J range [r0, r1) of tiles of one
//    element, pairs (I, J) with J <= I.
//  * A block is W waves of M consecutive Q bins each (16 x 8 or 12 x 12,
//    128 or 144 bins); every wave works on the same pairs.  Lane l = 16 R + p (row R of 16 lanes, place p) holds
//    atom I_l.  A tile pair (I, J) runs in 4 phases d of 16 steps s: lane
//    (R, p) forms the pair (I_l, J_b) with b = 16 ((R + d) mod 4) +
//    ((p + s) mod 16), so the 64 steps cover the 64 x 64 pairs once.  A
//    diagonal tile (J = I) runs the same steps and keeps the pairs with
//    l < b, which covers its 64 x 63 / 2 pairs once.
//  * Per step and bin the lane forms t = Q cos(Q r) - sin(Q r)/r from the
//    same three-term recurrences as the ordered kernel and adds t (q_i -
//    q_j)/r^2 to its own I row.  In phase d, lane (R, p) also holds the J
//    sums of J atom 16 ((R + d) mod 4) + p; its partner in step s is lane
//    (R, (p - s) mod 16), whose t it reads through the DPP row rotation
//    row_ror:s, an operand modifier of its FMAs: the recurrence and t are
//    computed once per pair and handing t over costs no instruction.  The
//    partner's (q_i - q_j)/r^2 comes from the shared setup.  Between phases
//    the J sums move up one row (ds_bpermute, 48 values).
//  * The per-pair setup (FP64 geometry, 2 cos(dtheta) as hi + lo, the W
//    waves' anchors sin, cos(Q0 r)) is computed once per block, one phase at a
//    time, and shared through LDS, as in the ordered kernel.
//  * Outputs go to slots of 64 rows x 3 x nq floats ([c][q][row], coalesced
//    per lane = row): one slot per (item, I) for the I rows (one FP32 run
//    over the J range, written once) and one per (item, J) for the J rows,
//    which the item's I tiles add to in turn (the first writes, the others
//    add with no-return atomics).  The item's block is the slot's only
//    writer, the same thread writes a given float every time, and the FP64
//    lane reduction at the end of each I tile passes a full barrier (which
//    waits for the thread's stores and atomics), so a slot's first store has
//    completed before the next I tile adds to it and the adds land in I
//    order: the sums repeat bit for bit.  reduce_slots_kernel sums every
//    tile's slots in FP64.
//
// Reference semantics (pyIID): the gradient of F(Q), cpu_flat.py:77-153 and
// gpu_flat.py:11-299 (get_grad_fq).
#include <type_traits>

#include "debye_math.h"
#include "internal.h"

namespace pyiid {
namespace {

constexpr int kSymS = 16;  // steps per phase (one setup round)

// Launch shapes: M Q bins per wave, W waves per block (M W bins per block):
// 16 x 8 (2 waves per SIMD) and 12 x 12 (3 per SIMD, within 168 VGPRs).
// LDS: the setup as FP32 fields [field][step][lane] (1/r, (q_i - q_j)/r^2,
// sin dtheta, hi, 8 lo, then sin, cos(Q0 r) of each wave's chunk), then the
// J tile's atoms.  The FP64 lane reduction reuses the start.
template <int M, int W>
struct SymShape {
  static constexpr int kBins = M * W;
  static constexpr int kFields = 7 + 2 * W;
  static constexpr int kGS = kSymS * 64;  // field stride (floats)
  static constexpr int kTile = kFields * kGS * 4;
  static constexpr int kEnd = kTile + 64 * 32;
  static constexpr int kRed = W * kRedM * kRedRow * 4;
  static constexpr int kSmem = kEnd > kRed ? kEnd : kRed;
};

// v from lane (R, (p - S) mod 16) of the caller's row (DPP row_ror:S)
template <int S>
__device__ __forceinline__ float row_ror(float v) {
  if constexpr (S == 0)
    return v;
  else
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x120 + S, 0xF, 0xF,
                                                      true));
}

// a += tp x, b += tp y, c += tp z with tp = t from lane (R, (p - S) mod 16)
// of the caller's row: the DPP row rotation as an operand modifier of the
// FMAs (the compiler keeps a separate v_mov_b32_dpp for an FMA's operand).
// A DPP read of a VGPR needs two wait states after the VALU write of it: the
// caller issues these after a scheduling barrier and an s_nop 1 that follow
// every write of t (volatile asm keeps them in order behind the s_nop).
template <int S>
__device__ __forceinline__ void fma3_ror(float& a, float& b, float& c, float t, float x, float y,
                                         float z) {
  if constexpr (S == 0) {
    a = fmaf(t, x, a);
    b = fmaf(t, y, b);
    c = fmaf(t, z, c);
  } else {
    asm volatile(
        "v_fmac_f32_dpp %0, %3, %4 row_ror:%7 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
        "v_fmac_f32_dpp %1, %3, %5 row_ror:%7 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
        "v_fmac_f32_dpp %2, %3, %6 row_ror:%7 row_mask:0xf bank_mask:0xf bound_ctrl:1"
        : "+v"(a), "+v"(b), "+v"(c)
        : "v"(t), "v"(x), "v"(y), "v"(z), "i"(S));
  }
}

template <int S>
using IC = std::integral_constant<int, S>;

// Block barrier that orders LDS only: the slot stores and atomics in flight
// need no wait here (every slot float has one writer thread, and the FP64
// lane reduction of each I tile, which uses a full barrier, orders them).
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// grid: n_items x n_qg blocks (block b: item b / n_qg, Q group b mod n_qg)
//   slots[slot][c][q][row]  float  gradient sums of the slot's 64 rows
//   fpart[item][q]          double sum over the item's pairs of sin(Q r)/r
// Weighted: every slot and F sum takes f_I(Q) f_J(Q) of its tile pair's
// elements (tile_el) when it is written; unweighted sums are plain.
template <int M, int W, bool WEIGHTED>
__global__ __launch_bounds__(64 * W) void sym_grad_kernel(
    const AtomRec* __restrict__ pos, const int32_t* __restrict__ perm,
    const Range* __restrict__ tiles, const int* __restrict__ tile_el,
    const float* __restrict__ ftab, int nq, double dq, const SymItem* __restrict__ items,
    int n_qg, float* __restrict__ slots, double* __restrict__ fpart) {
  using Sh = SymShape<M, W>;
  constexpr int kSymBins = Sh::kBins;
  __shared__ __attribute__((aligned(32))) unsigned char smem[Sh::kSmem];
  constexpr int GS = Sh::kGS;
  float* geo = reinterpret_cast<float*>(smem);
  AtomRec* jt = reinterpret_cast<AtomRec*>(smem + Sh::kTile);

  const int item = blockIdx.x / n_qg;
  const int qg = blockIdx.x - item * n_qg;
  const SymItem it = items[item];
  const int lane = threadIdx.x & 63;
  const int row = lane >> 4, place = lane & 15;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int q0 = qg * kSymBins + wave * M;
  const bool active = q0 < nq;  // wave-uniform
  const double qbase = static_cast<double>(qg * kSymBins);
  const size_t slot_floats = (size_t)3 * nq * 64;
  const int rot_addr = ((lane + 16) & 63) << 2;  // lane + 16: the next row

  // Q values stay in VGPRs (an FMA with a scalar operand issues more slowly)
  float qv[M];
#pragma unroll
  for (int m = 0; m < M; ++m) {
    qv[m] = static_cast<float>(static_cast<double>(q0 + m) * dq);
    asm volatile("" : "+v"(qv[m]));
  }
  // f_e(Q) of this wave's bins (weighted), 0 beyond nq
  auto fval = [&](int e, int m) {
    return q0 + m < nq ? ftab[(size_t)e * nq + q0 + m] : 0.0f;
  };

  double fsum = 0.0;  // lane m < M: the item's F partial of bin q0 + m
  const int i_first = max(it.g0, it.r0);
  for (int I = i_first; I < it.g1; ++I) {
    const Range ti = tiles[I];
    const int p_i = ti.i_base + lane;
    const bool lane_ok = p_i < ti.i_end;
    const AtomRec pi = pos[perm[lane_ok ? p_i : ti.i_base]];
    const int j_end = min(it.r1, I + 1);
    float aF[M], ix[M], iy[M], iz[M];
#pragma unroll
    for (int m = 0; m < M; ++m) aF[m] = ix[m] = iy[m] = iz[m] = 0.0f;

    for (int J = it.r0; J < j_end; ++J) {
      const Range tj = tiles[J];
      const int nj = tj.i_end - tj.i_base;
      const bool diag = J == I;
      // (the J tile is read by the setup only, which the last phase of the
      // previous J tile has finished behind its barrier)
      if ((int)threadIdx.x < nj) jt[threadIdx.x] = pos[perm[tj.i_base + threadIdx.x]];
      float jx[M], jy[M], jz[M];
#pragma unroll
      for (int m = 0; m < M; ++m) jx[m] = jy[m] = jz[m] = 0.0f;

      for (int d = 0; d < 4; ++d) {
        lds_barrier();
        if (d > 0) {
          // phase d - 1 -> d: the J sums move up one row (their ds_bpermute
          // latency runs under the setup below)
#pragma unroll
          for (int m = 0; m < M; ++m) {
            jx[m] = __int_as_float(__builtin_amdgcn_ds_bpermute(rot_addr, __float_as_int(jx[m])));
            jy[m] = __int_as_float(__builtin_amdgcn_ds_bpermute(rot_addr, __float_as_int(jy[m])));
            jz[m] = __int_as_float(__builtin_amdgcn_ds_bpermute(rot_addr, __float_as_int(jz[m])));
          }
        }
        // shared setup of the phase: wave w forms steps w, w + W, ...
        const int jrow = 16 * ((row + d) & 3);
        for (int k = wave; k < kSymS; k += W) {
          const int jl = jrow + ((place + k) & 15);
          const bool keep = lane_ok && jl < nj && (!diag || lane < jl);
          const AtomRec pj = jt[jl < nj ? jl : 0];
          const PairGeom g = pair_geom(pi, pj, keep);
          const double dth = dq * g.r;
          float s1, k2hi, k2lo;
          two_cos(dth, k2hi, k2lo, s1);
          const int o = k * 64 + lane;
          geo[o] = g.inv_r;
          geo[o + GS] = g.dpx;
          geo[o + 2 * GS] = g.dpy;
          geo[o + 3 * GS] = g.dpz;
          geo[o + 4 * GS] = s1;
          geo[o + 5 * GS] = k2hi;
          geo[o + 6 * GS] = 8.0f * k2lo;
          // anchors sin, cos(Q0 r) of the W waves' chunks: the first reduced
          // in FP64, the others by FP32 rotations by M dtheta
          float as, ac, rsn, rcs;
          sincos_red(qbase * dth, as, ac);
          sincos_red(static_cast<double>(M) * dth, rsn, rcs);
#pragma unroll
          for (int w = 0; w < W; ++w) {
            geo[o + (7 + 2 * w) * GS] = as;
            geo[o + (8 + 2 * w) * GS] = ac;
            const float as2 = fmaf(as, rcs, ac * rsn);
            ac = fmaf(ac, rcs, -as * rsn);
            as = as2;
          }
        }
        lds_barrier();
        // Every wave runs the steps (a wave past nq sums bins that are never
        // written): a branch around them would split the accumulators'
        // registers between two paths.
        //
        // g: 1/r, dpx, dpy, dpz, sin dtheta, hi, 8 lo, anchor sin, cos
        struct G {
          float f[9];
        };
        auto load = [&](G& g, int k) {
          const int o = k * 64 + lane;
#pragma unroll
          for (int f = 0; f < 7; ++f) g.f[f] = geo[o + f * GS];
          g.f[7] = geo[o + (7 + 2 * wave) * GS];
          g.f[8] = geo[o + (8 + 2 * wave) * GS];
        };
        // One step: per bin, t from the recurrences; the I sums take t dp,
        // the J sums the partner's t (row_ror:S) times the partner's dp (its
        // lane's dp through the same rotation; the sign is applied when the
        // slot is written).
        auto step = [&](auto sc, const G& g) {
          constexpr int S = decltype(sc)::value;
          const float inv_r = g.f[0], dpx = g.f[1], dpy = g.f[2], dpz = g.f[3];
          const float s1 = g.f[4], k2 = g.f[5], lo8 = g.f[6];
          const float px = row_ror<S>(dpx), py = row_ror<S>(dpy), pz = row_ror<S>(dpz);
          const float sn = g.f[7];
          float cs = g.f[8];
          const float c1 = 0.5f * k2;
          float w = sn * inv_r;
          float wp = fmaf(sn, c1, -cs * s1) * inv_r;  // sin(Q0 r - dtheta) / r
          float cp = fmaf(cs, c1, sn * s1);           // cos(Q0 r - dtheta)
          // the bins' t are kept, so that the J side of one bin does not wait
          // on the I side of the next
          float t[M];
#pragma unroll
          for (int m = 0; m < M; ++m) {
            t[m] = fmaf(qv[m], cs, -w);
            aF[m] += w;
            ix[m] = fmaf(t[m], dpx, ix[m]);
            iy[m] = fmaf(t[m], dpy, iy[m]);
            iz[m] = fmaf(t[m], dpz, iz[m]);
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
          if constexpr (S != 0) {
            __builtin_amdgcn_sched_barrier(0);
            asm volatile("s_nop 1");
          }
#pragma unroll
          for (int m = 0; m < M; ++m) fma3_ror<S>(jx[m], jy[m], jz[m], t[m], px, py, pz);
        };
        // Two register sets: the next step's setup is read from LDS while
        // the current step runs.
        G ga, gb;
        auto run = [&](auto sc) {
          constexpr int S = decltype(sc)::value;
          G& cur = (S & 1) ? gb : ga;
          G& nxt = (S & 1) ? ga : gb;
          if constexpr (S + 1 < kSymS) load(nxt, S + 1);
          __builtin_amdgcn_sched_barrier(0);
          step(sc, cur);
        };
        load(ga, 0);
        run(IC<0>{});
        run(IC<1>{});
        run(IC<2>{});
        run(IC<3>{});
        run(IC<4>{});
        run(IC<5>{});
        run(IC<6>{});
        run(IC<7>{});
        run(IC<8>{});
        run(IC<9>{});
        run(IC<10>{});
        run(IC<11>{});
        run(IC<12>{});
        run(IC<13>{});
        run(IC<14>{});
        run(IC<15>{});
      }

      // J rows (after phase 3 lane (R, p) holds J atom 16 ((R + 3) mod 4) +
      // p): the first I tile of the item to meet J writes its slot, the later
      // ones add to it
      const int jr = 16 * ((row + 3) & 3) + place;
      if (active && jr < nj) {
        const int eI = WEIGHTED ? tile_el[I] : 0, eJ = WEIGHTED ? tile_el[J] : 0;
        const bool first = I == max(it.g0, J);
        float* base = slots + (size_t)(it.jslot0 + (J - it.r0)) * slot_floats + (size_t)q0 * 64 + jr;
        // (fully unrolled with a guard per bin, so the sums stay in registers)
#pragma unroll
        for (int m = 0; m < M; ++m) {
          if (q0 + m < nq) {
            const float sc = WEIGHTED ? -(fval(eI, m) * fval(eJ, m)) : -1.0f;
            float* p = base + (size_t)m * 64;
            const float vx = jx[m] * sc, vy = jy[m] * sc, vz = jz[m] * sc;
            if (first) {
              p[0] = vx;
              p[(size_t)nq * 64] = vy;
              p[(size_t)2 * nq * 64] = vz;
            } else {
              unsafeAtomicAdd(p, vx);
              unsafeAtomicAdd(p + (size_t)nq * 64, vy);
              unsafeAtomicAdd(p + (size_t)2 * nq * 64, vz);
            }
          }
        }
      }
    }

    // I rows: one slot per (item, I), written once
    const int eI = WEIGHTED ? tile_el[I] : 0, eR = WEIGHTED ? tile_el[it.r0] : 0;
    if (active && lane_ok) {
      float* base = slots + (size_t)(it.islot0 + (I - i_first)) * slot_floats + (size_t)q0 * 64 + lane;
#pragma unroll
      for (int m = 0; m < M; ++m) {
        if (q0 + m < nq) {
          const float sc = WEIGHTED ? fval(eI, m) * fval(eR, m) : 1.0f;
          float* p = base + (size_t)m * 64;
          p[0] = ix[m] * sc;
          p[(size_t)nq * 64] = iy[m] * sc;
          p[(size_t)2 * nq * 64] = iz[m] * sc;
        }
      }
    }
    if constexpr (WEIGHTED) {
#pragma unroll
      for (int m = 0; m < M; ++m) aF[m] *= fval(eI, m) * fval(eR, m);
    }
    fsum += lane_reduce_f64<M>(aF, smem, wave, lane);
  }
  if (lane < M && q0 + lane < nq) fpart[(size_t)item * nq + q0 + lane] = fsum;
}

// out[perm[p]][c][q] = gscale[q] * sum over the tile's slots of
// slots[s][c][q][p - tile.i_base], in FP64 in list order.  Block: one tile
// x 16 bins, 4 waves over the 48 (component, bin) rows, lane = tile row; the
// sums are transposed through LDS into row-major output.
template <typename T>
__global__ __launch_bounds__(256) void reduce_slots_kernel(
    const float* __restrict__ slots, const int* __restrict__ entries,
    const Range* __restrict__ tiles, const int32_t* __restrict__ perm, int nq,
    const double* __restrict__ gscale, T* __restrict__ out) {
  __shared__ double sh[48][65];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const Range tr = tiles[blockIdx.y];
  const int qc = blockIdx.x * 16;
  const size_t slot_floats = (size_t)3 * nq * 64;
  for (int r = wave; r < 48; r += 4) {
    const int c = r >> 4, q = qc + (r & 15);
    double acc = 0.0;
    if (q < nq) {
      const float* src = slots + ((size_t)c * nq + q) * 64 + lane;
      int e = tr.begin;
      for (; e + 4 <= tr.end; e += 4) {
        const float x0 = src[(size_t)entries[e] * slot_floats];
        const float x1 = src[(size_t)entries[e + 1] * slot_floats];
        const float x2 = src[(size_t)entries[e + 2] * slot_floats];
        const float x3 = src[(size_t)entries[e + 3] * slot_floats];
        acc += static_cast<double>(x0);
        acc += static_cast<double>(x1);
        acc += static_cast<double>(x2);
        acc += static_cast<double>(x3);
      }
      for (; e < tr.end; ++e) acc += static_cast<double>(src[(size_t)entries[e] * slot_floats]);
      acc *= gscale[q];
    }
    sh[r][lane] = acc;
  }
  __syncthreads();
  // thread (row, part): row = tile row, part picks every fourth of the 48
  const int row = threadIdx.x >> 2, part = threadIdx.x & 3;
  if (tr.i_base + row >= tr.i_end) return;
  T* dst = out + (size_t)perm[tr.i_base + row] * 3 * nq;
  for (int r = part; r < 48; r += 4) {
    const int c = r >> 4, q = qc + (r & 15);
    if (q < nq) dst[(size_t)c * nq + q] = static_cast<T>(sh[r][row]);
  }
}

}  // namespace

template <int M, int W>
hipError_t launch_sym_m(hipStream_t st, const AtomRec* pos, const int32_t* perm,
                        const Range* tiles, const int* tile_el, const float* ftab, int nq,
                        double dq, const SymItem* items, int n_items, bool weighted, float* slots,
                        double* fpart) {
  constexpr int kBins = SymShape<M, W>::kBins;
  const int n_qg = (nq + kBins - 1) / kBins;
  if ((int64_t)n_items * n_qg > INT32_MAX) return hipErrorInvalidValue;
  dim3 grid(static_cast<unsigned>(n_items * n_qg));
  if (weighted)
    hipLaunchKernelGGL((sym_grad_kernel<M, W, true>), grid, dim3(64 * W), 0, st, pos, perm, tiles,
                       tile_el, ftab, nq, dq, items, n_qg, slots, fpart);
  else
    hipLaunchKernelGGL((sym_grad_kernel<M, W, false>), grid, dim3(64 * W), 0, st, pos, perm,
                       tiles, tile_el, ftab, nq, dq, items, n_qg, slots, fpart);
  return hipGetLastError();
}

hipError_t launch_sym_grad(hipStream_t st, int m, const AtomRec* pos, const int32_t* perm,
                           const Range* tiles, const int* tile_el, const float* ftab, int nq,
                           double dq, const SymItem* items, int n_items, bool weighted,
                           float* slots, double* fpart) {
  if (n_items <= 0 || nq <= 0) return hipSuccess;
  if (m == 16)
    return launch_sym_m<16, 8>(st, pos, perm, tiles, tile_el, ftab, nq, dq, items, n_items,
                               weighted, slots, fpart);
  if (m == 12)
    return launch_sym_m<12, 12>(st, pos, perm, tiles, tile_el, ftab, nq, dq, items, n_items,
                                weighted, slots, fpart);
  return hipErrorInvalidValue;
}

hipError_t launch_reduce_slots(hipStream_t st, const float* slots, const int* entries,
                               const Range* tiles, int n_tiles, const int32_t* perm, int nq,
                               const double* gscale, void* out, bool out_f64) {
  if (n_tiles <= 0 || nq <= 0) return hipSuccess;
  dim3 grid((nq + 15) / 16, n_tiles);
  if (out_f64)
    hipLaunchKernelGGL(reduce_slots_kernel<double>, grid, dim3(256), 0, st, slots, entries, tiles,
                       perm, nq, gscale, static_cast<double*>(out));
  else
    hipLaunchKernelGGL(reduce_slots_kernel<float>, grid, dim3(256), 0, st, slots, entries, tiles,
                       perm, nq, gscale, static_cast<float*>(out));
  return hipGetLastError();
}

}  // namespace pyiid
