// Internal declarations shared by the gfx950 kernels and the C ABI layer.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace pyiid {

// Atom record in HBM: float32 input positions widened exactly to double, and
// the element index stored in w.  32 B per atom, one vector load.
struct alignas(32) AtomRec {
  double x, y, z, w;
};

// Q bins per wave in the fused F + gradient kernel and in the F-only kernel.
constexpr int kGradM = 32;
constexpr int kFqM = 64;
constexpr int kWaves = 4;      // waves per block in the Debye kernels
constexpr int kJTile = 256;    // atoms staged in LDS per j tile
constexpr int kMaxElem = 8;    // element-table rows staged in LDS (weighted kernels)

// Work item of the F-only (unordered pair) kernel: i tile and j range.
struct FqItem {
  int itile, jbeg, jend, pad;
};

// ---- launch wrappers (debye.hip) ----
hipError_t launch_fused_grad(hipStream_t st, const AtomRec* pos, int n, const float* ftab,
                             int n_elem, int nq, double dq, int n_split, int j_chunk,
                             bool weighted, float* gpart, double* fpart, const float* u,
                             float* force_part);
hipError_t launch_fq(hipStream_t st, const AtomRec* pos, int n, const float* ftab, int n_elem,
                     int nq, double dq, const FqItem* items, int n_items, bool weighted,
                     double* fpart);
// out[q] = scale[q] * sum_p part[p * nq + q]   (FP64 cross-tile reduction)
hipError_t launch_reduce_rows(hipStream_t st, const double* part, int n_part, int nq,
                              const double* scale, double* out_f64, float* out_f32);
// grad[i][c][q] = gscale[q] * sum_s gpart[s][i][c][q]
hipError_t launch_reduce_grad(hipStream_t st, const float* gpart, int n_split, int64_t n,
                              int nq, const double* gscale, void* out, bool out_f64);
// force[i][c] = conv * sum_p fpart[p][i][c]
hipError_t launch_reduce_force(hipStream_t st, const float* fpart, int n_part, int64_t n,
                               double conv, double* out);
// na[q], F scale 2w/na (unordered), F scale w/na (ordered), grad scale -w/na
hipError_t launch_norm(hipStream_t st, const float* ftab, const int64_t* counts, int n_elem,
                       int nq, int64_t n, bool weighted, double* na, double* fscale_unordered,
                       double* fscale_ordered, double* gscale);
hipError_t launch_pair_geometry(hipStream_t st, const AtomRec* pos, int64_t k_offset,
                                int64_t k_count, float* d, float* r);

// ---- transform / potential (transform.hip) ----
// C[m][ldc] (+)= A[m][K] . B[n][K]^T in FP32 MFMA; A float or double.
hipError_t launch_gemm_nt(hipStream_t st, const void* a, bool a_f64, int64_t m, int k,
                          const float* b, int n, double* c_f64, float* c_f32);
// potential value/scale and adjoint vector w (all FP64), single block.
// out5 = {value, scale, a_opt, res_dot_res, g_dot_res}
hipError_t launch_potential(hipStream_t st, int potential, const double* gobs,
                            const double* gcalc, int n, double* out5, double* w);
// y[row] = sum_t X[row][t] * w[t]  (FP64), one wave per row
hipError_t launch_gemv_rows(hipStream_t st, const double* x, int64_t rows, int n,
                            const double* w, double* y);
// u[q] = gscale[q] * sum_t T[t][q] * w[t]  for q in [0,nq); w has n_r entries
hipError_t launch_adjoint(hipStream_t st, const float* t, int n_r, int nq, const double* w,
                          const double* gscale, float* u);
// u[q] = gscale[q] * w[q - q_first] (0 below q_first)
hipError_t launch_adjoint_fq(hipStream_t st, int nq, int q_first, const double* w,
                             const double* gscale, float* u);
hipError_t launch_f64_to_f32(hipStream_t st, const double* x, int64_t n, float* y);
hipError_t launch_f32_to_f64(hipStream_t st, const float* x, int64_t n, double* y);

}  // namespace pyiid
