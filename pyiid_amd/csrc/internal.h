// Internal declarations shared by the gfx950 kernels and the C ABI layer.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace pyiid {

// Atom record in HBM: float32 input positions widened exactly to double (the
// sampler integrates them in double), and the element index stored in w.
// 32 B per atom, one vector load.
struct alignas(32) AtomRec {
  double x, y, z, w;
};

// Q bins per wave of the fused F + gradient kernel (the larger of its two shapes).
constexpr int kGradM = 32;
constexpr int kFqM = 64;  // default Q bins per wave of the F-only kernel
constexpr int kWaves = 4;      // waves per block in the contracted-force kernel
constexpr int kJTile = 256;    // atoms staged in LDS per j tile

// Work item of the Debye kernels.  Structures ("chains") are stored back to
// back; an item is a 64-atom i tile [i_base, i_end) of one chain against the
// j range [jbeg, jend) of the same chain.  `split` indexes the gradient
// partial buffer, `tile` the chain-global i tile (for force reduction).
struct Item {
  int i_base, i_end, jbeg, jend;
  int split, chain, tile, pad;
};

// Items [item_begin, item_end) of one i tile (or of one chain, for F).
struct Range {
  int begin, end, i_base, i_end;
};

// ---- launch wrappers (debye.hip) ----
// Ordered-pair F + gradient, m = 16 or 32 Q bins per wave (compensated
// sums: m = 16 only):
// gpart[item.split][i][c][q] (float) for i in the item's tile and
// fpart[item][q] (F partial, double).  Weighted: the j indices of an item are
// positions in jperm and item.pad is the element of the item's j range;
// unweighted: atom indices.
hipError_t launch_fused_grad(hipStream_t st, int m, bool compensated, const AtomRec* pos,
                             const int32_t* jperm, int n_total, const float* ftab, int nq,
                             double dq, const Item* items, int n_items, bool weighted,
                             float* gpart, double* fpart);
// Work item of the symmetric F + gradient kernel: 64-position tiles [g0, g1)
// against tiles [r0, r1) of one element, pairs (I, J) with J <= I.  The I
// rows of tile I go to slot islot0 + (I - max(g0, r0)), the J rows of tile J
// to slot jslot0 + (J - r0).
struct SymItem {
  int g0, g1, r0, r1, islot0, jslot0, pad0, pad1;
};
// Unordered-pair F + gradient: slots[slot][c][q][64] (float; tiles[t] =
// {slot list begin, end, position begin, end}, positions in perm) and
// fpart[item][q] (F partial, double).  Weighted: tile_el[t] is the element
// of tile t.
// m = 16 (8 waves, 128 bins per block) or 12 (12 waves, 144 bins).
hipError_t launch_sym_grad(hipStream_t st, int m, const AtomRec* pos, const int32_t* perm,
                           const Range* tiles, const int* tile_el, const float* ftab, int nq,
                           double dq, const SymItem* items, int n_items, bool weighted,
                           float* slots, double* fpart);
// out[perm[p]][c][q] = gscale[q] * sum_{e in tile} slots[entries[e]][c][q][row]
// (FP64, list order), float or double output.
hipError_t launch_reduce_slots(hipStream_t st, const float* slots, const int* entries,
                               const Range* tiles, int n_tiles, const int32_t* perm, int nq,
                               const double* gscale, void* out, bool out_f64);
// Contracted forces over unordered pairs: items are tile pairs (I, J), J <= I,
// of at most 64 positions of one chain in perm (Item.tile = I's and Item.pad =
// J's chain-global tile index; weighted: Item.split = e_I | e_J << 16, each
// tile of one element); n_groups Q groups of `span` bins (a multiple of 32)
// each.  part[(item * n_groups + g)][side][64][3], side 0 = I, 1 = J.
constexpr size_t kPairForceLds = 64 * 1024;
hipError_t launch_pair_force(hipStream_t st, const AtomRec* pos, const int32_t* perm,
                             const float* ftab, int nq, double dq, const Item* items, int n_items,
                             int n_groups, int span, bool weighted, const float* u,
                             const float* uq, float* part);
// force[perm[p]][c] = conv * sum_{e in entries[tile]} part[e][row][c] (FP64)
// for the tile's positions p = i_base + row, where an entry is work * 2 +
// side; tiles[t] = {entry begin, end, i_base, i_end}.
hipError_t launch_reduce_pair_force(hipStream_t st, const float* part, const int* entries,
                                    const Range* tiles, int n_tiles, const int32_t* perm,
                                    double conv, double* out);
// uq[b][q] = u[b][q] * float32(q dq)
hipError_t launch_uq(hipStream_t st, const float* u, int n_chain, int nq, double dq, float* uq);
// Unordered-pair F only: fpart[item][q] (double), m Q bins per wave (64, 72,
// 80, 88, 96 or 112; compensated sums: 64 only).  Item indices are positions
// in perm; weighted: item.split is the element of the item's i tile and
// item.pad that of its j range.  forms (nullable): forms[0], forms[1] gain
// the (wave, j atom) steps run in the three-term and in the Reinsch form.
hipError_t launch_fq(hipStream_t st, int m, bool compensated, const AtomRec* pos,
                     const int32_t* perm, const float* ftab, int nq, double dq, const Item* items,
                     int n_items, bool weighted, double* fpart, unsigned long long* forms);
// out[r][q] = scale[r][q] * sum_{p in ranges[r]} part[p][q], r < n_rows (FP64)
hipError_t launch_reduce_ranges(hipStream_t st, const double* part, const Range* ranges,
                                int n_rows, int nq, const double* scale, double* out_f64,
                                float* out_f32);
// grad[i][c][q] = gscale[q] * sum_s gpart[s][i][c][q]
hipError_t launch_reduce_grad(hipStream_t st, const float* gpart, int n_split, int64_t n,
                              int nq, const double* gscale, void* out, bool out_f64);
// per-chain normalisation: na, 2w/na, w/na, -w/na  ([chain][q])
hipError_t launch_norm(hipStream_t st, const float* ftab, const int64_t* counts, int n_elem,
                       int nq, const int64_t* n_atoms, int n_chain, bool weighted, double* na,
                       double* fscale_unordered, double* fscale_ordered, double* gscale);
hipError_t launch_pair_geometry(hipStream_t st, const AtomRec* pos, int64_t k_offset,
                                int64_t k_count, float* d, float* r);

// ---- transform / potential (transform.hip) ----
// C[m][ldc] = A[m][K] . B[n][K]^T in FP32 MFMA; A float or double.
hipError_t launch_gemm_nt(hipStream_t st, const void* a, bool a_f64, int64_t m, int k,
                          const float* b, int n, double* c_f64, float* c_f32);
// Per chain b (one block each): potential value/scale and the adjoint
// w[b][t] (FP64).  out5[b] = {value, scale, a_opt, res.res, g.res}.
hipError_t launch_potential(hipStream_t st, int potential, const double* gobs,
                            const double* gcalc, int n, int n_chain, double* out5, double* w);
// y[row] = sum_t X[row][t] * w[t]  (FP64), one wave per row
hipError_t launch_gemv_rows(hipStream_t st, const double* x, int64_t rows, int n,
                            const double* w, double* y);
// u[b][q] = gscale[b][q] * sum_t T[t][q] w[b][t]   (FP64 accumulation)
hipError_t launch_adjoint(hipStream_t st, const float* t, int n_r, int nq, const double* w,
                          const double* gscale, int n_chain, float* u);
// u[q] = gscale[q] * w[q - q_first] (0 below q_first)
hipError_t launch_adjoint_fq(hipStream_t st, int nq, int q_first, const double* w,
                             const double* gscale, float* u);
hipError_t launch_f64_to_f32(hipStream_t st, const double* x, int64_t n, float* y);
hipError_t launch_f32_to_f64(hipStream_t st, const float* x, int64_t n, double* y);

// ---- sampler (hmc.hip) ----
struct HmcChain {
  int atom_begin, atom_end;
};
// Start of a proposal: save positions, draw momenta p ~ N(0, m/beta) from the
// counter-based generator (seed, proposal counter), kinetic energy K0[b],
// first half kick p += eps/2 F.
hipError_t launch_hmc_begin(hipStream_t st, AtomRec* pos, double* pos_save, double* mom,
                            const double* force, const double* mass, const HmcChain* chains,
                            int n_chain, double beta, double eps, uint64_t seed,
                            const int64_t* counter, double* kin0);
// x += eps p / m
hipError_t launch_hmc_drift(hipStream_t st, AtomRec* pos, const double* mom, const double* mass,
                            int64_t n, double eps);
// p += h F
hipError_t launch_hmc_kick(hipStream_t st, double* mom, const double* force, int64_t n, double h);
// End of a proposal: K1, Metropolis with exp(beta (H0 - H1)), restore or
// accept (positions, force, energy), record energy/acceptance, bump counter.
// force_cur null: no force copy.  viol (nullable): chains flagged there are
// rejected, counted in viol_count and their flag cleared.
hipError_t launch_hmc_end(hipStream_t st, AtomRec* pos, const double* pos_save, const double* mom,
                          const double* mass, double* force_cur, const double* force_new,
                          double* e_cur, const double* out5_new, double conv,
                          const HmcChain* chains, int n_chain, double beta, uint64_t seed,
                          int64_t* counter, const double* kin0, double* e_hist,
                          int32_t* acc_hist, int64_t hist_cap, int* viol, int* viol_count);
// e_cur[b] = conv * out5[b].value
hipError_t launch_energy_from_out5(hipStream_t st, const double* out5, int n_chain, double conv,
                                   double* e);

// ---- learned integrator (surrogate.hip) ----
// Network parameter block: bf16 W1[16][64], bf16 W2[64][64], f32 b1[64],
// b2[64], descriptor mean[16], inv_scale[16].
constexpr int kSurNetBytes = 16 * 64 * 2 + 64 * 64 * 2 + 64 * 4 * 2 + 16 * 4 * 2;
constexpr int kSurHidden = 64;
constexpr int kSurRbf = 16;
constexpr int kSurGram = 64 * 64 + 64 + 1;  // Phi Phi^T, Phi y, y.y per chain
// tiles: Item{i_base, i_end, jbeg, jend = the chain's atoms, chain}.
hipError_t launch_nbr_build(hipStream_t st, const AtomRec* pos, int n, const Item* tiles,
                            int n_tiles, float rl, int max_nbr, int* nbr, int* cnt, int* viol);
// rbf = {rc, mu0, dmu, eta}
hipError_t launch_descriptors(hipStream_t st, const AtomRec* pos, int n, const int* nbr,
                              const int* cnt, const float* rbf, float* d);
hipError_t launch_mlp(hipStream_t st, const float* d, int n, const Item* tiles, int n_tiles,
                      const void* net, const unsigned short* w3, int w3_stride, float* c,
                      float* e_atom);
hipError_t launch_sur_force(hipStream_t st, const AtomRec* pos, int n, const int* nbr,
                            const int* cnt, const float* rbf, const float* c, double* force,
                            float* phi);
// chunks: Item{jbeg, jend = flat column range, split = partial slot, chain};
// ranges[b] = {first slot, end slot}.  out[b] (FP64, kSurGram values).
hipError_t launch_gram(hipStream_t st, const float* phi, int64_t n3, const double* y,
                       const Item* chunks, int n_chunks, float* part, const Range* ranges,
                       int n_chain, double* out);
// out[b][k] = sum_{a in chain b} e_atom[k * n + a] (FP64), off: chain offsets (n_chain + 1).
hipError_t launch_chain_feature_energy(hipStream_t st, const float* e_atom, int64_t n,
                                       const int64_t* off, int n_chain, double* out);
hipError_t launch_disp_check(hipStream_t st, const AtomRec* pos, const double* pos_save,
                             const int* atom_chain, int64_t n, double lim, int* viol);

// Records the calling thread's message for pyiid_last_error and returns code.
int set_error(int code, const char* fmt, ...);

}  // namespace pyiid
