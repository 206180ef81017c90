/*
 * pyiid_hip.h - C ABI of the MI355X (gfx950) backend for pyIID's scattering
 * path: Debye F(Q), its position gradient, the F(Q) -> G(r) transform and the
 * Rw / chi^2 energy and forces.
 *
 * Conventions
 *  - Every call returns 0 on success and a negative PYIID_E* code on failure;
 *    pyiid_last_error() then returns a message (thread-local).
 *  - No exceptions cross the boundary.  Plain pointers and sizes only.
 *  - Pointer arguments are host pointers unless PYIID_DEVICE is set in
 *    `flags`, in which case they are device pointers on the context's device
 *    (e.g. torch.Tensor.data_ptr()).  Host results are complete when the call
 *    returns; device results are ordered on the context stream.
 *  - A context owns one HIP stream (or borrows one, pyiid_set_stream), the
 *    device-resident structure and tables, and its workspaces.  Calls on one
 *    context must not overlap; distinct contexts are independent.
 *
 * Reference citations are path:line in the pyIID repository.
 */
#ifndef PYIID_HIP_H
#define PYIID_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PYIID_ABI_VERSION 1

/* flags */
#define PYIID_DEVICE 1 /* pointers are device pointers */
#define PYIID_F64 2    /* gradient arrays are double (else float) */

/* status codes */
#define PYIID_OK 0
#define PYIID_E_INVALID -1  /* bad argument / shape */
#define PYIID_E_STATE -2    /* required state not set (positions, tables, transform) */
#define PYIID_E_HIP -3      /* HIP runtime error */
#define PYIID_E_NOMEM -4    /* device allocation failed */
#define PYIID_E_NODEVICE -5 /* no HIP device */

/* scatter-table slots: the two per-atom tables of _wrap_atoms
 * (pyiid/experiments/elasticscatter/__init__.py:127-148) */
#define PYIID_SLOT_FQ 0  /* 'F(Q) scatter' */
#define PYIID_SLOT_PDF 1 /* 'PDF scatter' */

/* potentials of Calc1D (pyiid/calc/calc_1d.py:28-34) */
#define PYIID_POT_RW 0
#define PYIID_POT_CHI_SQ 1

typedef struct pyiid_ctx pyiid_ctx;

/* ---- library / context ------------------------------------------------ */
int pyiid_abi_version(void);
const char *pyiid_last_error(void);
/* Replaces the device probe check_gpu() / get_gpus_mem()
 * (elasticscatter/__init__.py:29-38, experiments/__init__.py:8-19). */
int pyiid_device_count(int *count);
int pyiid_create(int device, pyiid_ctx **ctx);
int pyiid_destroy(pyiid_ctx *ctx);
/* Use `stream` (a hipStream_t; NULL restores the context's own stream). */
int pyiid_set_stream(pyiid_ctx *ctx, void *stream);
int pyiid_synchronize(pyiid_ctx *ctx);
/* Kernel timing with HIP events on the launch stream: when enabled, every
 * launch of the Debye kernels is bracketed by events and its duration summed.
 * kind: 0 = F+grad kernel (ordered or symmetric, pyiid_grad_shape tells
 *       which), 1 = F-only kernel, 2 = transform GEMM,
 *       3 = pair-geometry kernel, 4 = contracted-force (unordered pair) kernel. */
int pyiid_enable_timing(pyiid_ctx *ctx, int enable);
int pyiid_kernel_time(pyiid_ctx *ctx, int kind, double *total_ms, int64_t *launches);
/* While timing is enabled, the F-only kernel counts its (wave, j atom) steps
 * by recurrence form: the three-term form, or the Reinsch form that a wave
 * takes when any of its pairs has a per-bin phase step within 0.2 of a
 * multiple of pi.  Counts since timing was last enabled.  No counterpart in
 * the reference. */
int pyiid_fq_form_counts(pyiid_ctx *ctx, int64_t *three_term, int64_t *reinsch);
/* Q bins per wave of the F-only kernel: 64, 72, 80, 88, 96 or 112 fixes the
 * launch shape; 0 (the default) lets the autotuner time the six once per
 * problem shape and keep the fastest.  The shapes differ only in how the bins are
 * grouped, so F agrees between them to FP32 rounding.  No counterpart in the
 * reference (its kernels have no launch-shape choice). */
int pyiid_set_fq_bins_per_wave(pyiid_ctx *ctx, int m);
/* Q bins per wave of the ordered-pair F + gradient kernel: 16 or 32 runs
 * pyiid_grad_fq on that kernel with that launch shape (and the default j
 * split); 0 (the default) lets the autotuner choose between the symmetric
 * (unordered-pair) kernel and the ordered one, and time the ordered one's two
 * shapes with three split sizes each, once per problem shape.  Kernels and
 * shapes differ only in how the sums are grouped, so the results agree to
 * FP32 rounding.  No counterpart in the reference (its kernels have no
 * launch-shape choice). */
int pyiid_set_grad_bins_per_wave(pyiid_ctx *ctx, int m);
/* Tiles of 64 atoms per I group and per J range of a work item of the
 * symmetric F + gradient kernel: 4, 8, 16 or 32 runs pyiid_grad_fq on that
 * kernel with that grouping (unless compensated sums are on or the ordered
 * kernel's bins per wave are fixed); 0 (the default) leaves the choice to the
 * autotuner.  Larger groups write fewer partial-sum slots (about 3 nq 64 x 4
 * bytes each, tiles^2 / k of them) and give fewer, longer blocks.  No
 * counterpart in the reference. */
int pyiid_set_grad_tiles_per_item(pyiid_ctx *ctx, int k);
/* Q bins per wave of the symmetric F + gradient kernel: 16 (8 waves per
 * block) or 12 (12 waves per block) fixes the launch shape; 0 (the default)
 * lets the autotuner time both.  The results agree to FP32 rounding.  No
 * counterpart in the reference. */
int pyiid_set_grad_sym_bins_per_wave(pyiid_ctx *ctx, int m);
/* Kernel of pyiid_grad_fq: PYIID_GRAD_AUTO (the default) lets the autotuner
 * choose, PYIID_GRAD_ORDERED and PYIID_GRAD_SYMMETRIC restrict it to one
 * kernel (whose launch shapes are still autotuned unless fixed by the setters
 * above).  Compensated sums always run the ordered kernel.  No counterpart in
 * the reference. */
#define PYIID_GRAD_AUTO 0
#define PYIID_GRAD_ORDERED 1
#define PYIID_GRAD_SYMMETRIC 2
int pyiid_set_grad_kernel(pyiid_ctx *ctx, int kernel);
/* Launch shape of the last pyiid_grad_fq call: *symmetric = 1 for the
 * symmetric (unordered-pair) kernel, with its Q bins per wave and tiles per
 * item in *group; 0 for the ordered kernel, with its Q bins per wave and its
 * number of j ranges per tile in *group.  No counterpart in the reference. */
int pyiid_grad_shape(pyiid_ctx *ctx, int *symmetric, int *bins_per_wave, int *group);
/* Autotuning of the Debye kernels' launch shapes (F + gradient kernel and its
 * grouping, F-only bins per wave, contracted-force Q groups), on by
 * default: the first call for a problem shape times the candidates and keeps
 * the fastest.  Off (enable = 0): every call uses fixed default shapes (the
 * ordered F + gradient kernel unless pyiid_set_grad_tiles_per_item is set), so
 * results are the same bit for bit from run to run and context to context;
 * with it on they agree to FP32 rounding.  No counterpart in the reference. */
int pyiid_set_autotune(pyiid_ctx *ctx, int enable);
/* Summation of the Debye kernels' FP32 per-bin sums.
 *  - PYIID_PRECISION_DEFAULT: plain FP32 sums over runs of at most 2,048
 *    j atoms, then FP64 across runs and tiles.
 *  - PYIID_PRECISION_COMPENSATED: compensated FP32 sums.  The terms of each
 *    256-atom j tile gather in a second register, which an error-free TwoSum
 *    folds into the running sum.  The F + gradient kernel then runs 16 bins
 *    per wave in 8-wave blocks.
 * The reference sums each pair chunk in FP64 (atomics/cpu_atomics.py:78);
 * this is the FP32 device counterpart of that choice. */
#define PYIID_PRECISION_DEFAULT 0
#define PYIID_PRECISION_COMPENSATED 1
int pyiid_set_precision(pyiid_ctx *ctx, int mode);

/* ---- structure and tables ---------------------------------------------- */
/* Atom positions, float32 N x 3 (the reference casts positions to float32:
 * flat_multi_cpu_wrap.py:9-19, gpu_wrap.py:12-24). */
int pyiid_set_positions(pyiid_ctx *ctx, const float *pos, int64_t n, int flags);
/* Element index per atom in [0, n_elem) (int32, N).  Rows of the scatter
 * tables below are indexed by it.  Replaces the per-atom N x Q float arrays
 * 'F(Q) scatter' / 'PDF scatter' that _wrap_atoms stores on the atoms
 * (elasticscatter/__init__.py:127-148) by an E x Q table plus this index. */
int pyiid_set_elements(pyiid_ctx *ctx, const int32_t *elem, int64_t n, int flags);
/* A batch of n_chain independent structures stored back to back: chain b is
 * atoms [offsets[b], offsets[b+1]) of pos (float32 n_total x 3); offsets is a
 * host array of n_chain + 1 entries from 0 to n_total.  The reference handles
 * one structure per call (wrap_fq / wrap_fq_grad); batches serve the sampler
 * (one F(Q), G(r), energy and force evaluation for every chain per launch).
 * Elements reset to 0 when the layout changes (set them after this call). */
int pyiid_set_batch(pyiid_ctx *ctx, const float *pos, int64_t n_total, const int64_t *offsets,
                    int32_t n_chain, int flags);
/* Current positions (float32 n_total x 3), e.g. after sampling. */
int pyiid_get_positions(pyiid_ctx *ctx, float *pos, int flags);
/* Per-element form-factor table f[e][q], float32 n_elem x n_q, for `slot`. */
int pyiid_set_scatter(pyiid_ctx *ctx, int slot, const float *table, int32_t n_elem,
                      int32_t n_q, int flags);

/* ---- Debye sums -------------------------------------------------------- */
/* F(Q) on the full grid from Q=0: 2/na * sum_{i>j} f_i f_j sin(Q r)/r with
 * Q = m * float32(qbin), nan_to_num.  Writes n_q floats per chain
 * (n_chain x n_q).
 * Replaces wrap_fq(atoms, qbin, sum_type) of the CPU and GPU processors
 * (cpu_wrappers/flat_multi_cpu_wrap.py:22-60, gpu_wrappers/gpu_wrap.py:119-156). */
int pyiid_fq(pyiid_ctx *ctx, int slot, double qbin, float *fq, int flags);
/* -1/2 dF/dq as the reference returns it: (N, 3, n_q), Q fastest; float32,
 * or double with PYIID_F64.  `fq` (nullable) receives F(Q) from the same
 * fused kernel.  Replaces wrap_fq_grad(atoms, qbin, sum_type)
 * (flat_multi_cpu_wrap.py:63-102, gpu_wrap.py:159-194). */
int pyiid_grad_fq(pyiid_ctx *ctx, int slot, double qbin, void *grad, float *fq, int flags);
/* Flat pair geometry for pairs k in [k_offset, k_offset + k_count) with
 * k = j + i(i-1)/2, i > j: d[k] = q_i - q_j (k x 3 float) and r[k] = |d[k]|.
 * Replaces get_d_array + get_r_array (kernels/cpu_flat.py:14-52,
 * kernels/gpu_flat.py:11-49).  d or r may be NULL. */
int pyiid_pair_geometry(pyiid_ctx *ctx, int64_t k_offset, int64_t k_count, float *d, float *r,
                        int flags);
/* na(Q) = N * mean_{i>j} f_i f_j for `slot` (double, n_chain x n_q).  Replaces
 * get_normalization_array + the mean in the wrappers
 * (cpu_flat.py:55-74, flat_multi_cpu_wrap.py:52-54). */
int pyiid_normalization(pyiid_ctx *ctx, int slot, double *na, int flags);

/* ---- F(Q) -> G(r) ------------------------------------------------------ */
/* Build the transform of get_pdf_at_qmin (kernels/master_kernel.py:39-104)
 * for F of length n_q on the grid rgrid (n_r doubles, host memory). */
int pyiid_set_transform(pyiid_ctx *ctx, double rstep, double qstep, const double *rgrid,
                        int32_t n_r, double qmin, int32_t n_q);
/* G(r) from F(Q) (n_q floats) -> n_r doubles.  Replaces get_pdf_at_qmin. */
int pyiid_pdf(pyiid_ctx *ctx, const float *fq, double *gr, int flags);
/* grad G = grad F . T^T for `rows` rows of length n_q (float, or double with
 * PYIID_F64) -> rows x n_r doubles.  Replaces grad_pdf
 * (master_kernel.py:272-290, gpu_wrap.py:197-284); rows = 3N. */
int pyiid_grad_pdf(pyiid_ctx *ctx, const void *grad_fq, int64_t rows, double *grad_gr,
                   int flags);

/* ---- Rw / chi^2 -------------------------------------------------------- */
/* Rw (or chi^2) and scale of gcalc against gobs (n doubles each).
 * Replaces get_rw / get_chi_sq (master_kernel.py:206-266). */
int pyiid_potential(pyiid_ctx *ctx, int potential, const double *gobs, const double *gcalc,
                    int32_t n, double *value, double *scale, int flags);
/* Gradient of the potential for `rows` rows of grad_gcalc (rows x n doubles)
 * -> rows doubles.  Replaces get_grad_rw / get_grad_chi_sq
 * (master_kernel.py:293-375) as called by wrap_grad_rw / wrap_grad_chi_sq
 * (calc/__init__.py:56-105). */
int pyiid_potential_grad(pyiid_ctx *ctx, int potential, const double *grad_gcalc, int64_t rows,
                         const double *gobs, const double *gcalc, int32_t n, double *grad,
                         int flags);
/* Calc1D energy and forces for a PDF target (calc_1d.py:78-95 with
 * exp_function=get_pdf, exp_grad_function=get_grad_pdf), for every chain:
 * energy[b] = conv * P(G_b), forces = conv * dP/dG . gradG in the
 * reference's convention (gradG built from the -1/2 dF/dq array).  G is
 * computed from the `slot` table at `qbin`; gobs holds n_chain x n_r points.
 * The gradient array is never materialised: the adjoint u = T^T w is
 * contracted inside the Debye gradient kernel.  energy: n_chain doubles;
 * forces: n_total x 3 doubles (nullable). */
int pyiid_pdf_energy_forces(pyiid_ctx *ctx, int slot, double qbin, const double *gobs,
                            int potential, double conv, double *energy, double *forces,
                            int flags);
/* Same for an F(Q) target (exp_function=get_fq, exp_grad_function=get_grad_fq):
 * gobs has n_q - q_first points, the F grid sliced from bin q_first
 * (elasticscatter/__init__.py:331-332,495-496). */
int pyiid_fq_energy_forces(pyiid_ctx *ctx, int slot, double qbin, int32_t q_first,
                           const double *gobs, int potential, double conv, double *energy,
                           double *forces, int flags);

/* ---- device-resident Hamiltonian Monte Carlo ---------------------------- */
/* Prepare the sampler for every chain of the current structure against a
 * PDF target (gobs: n_chain x n_r of the transform), potential and conv as
 * in Calc1D (calc_1d.py:15-42), per-atom masses (n_total doubles).  Computes
 * the initial energy and forces.  The sampler integrates exact forces
 * -dE/dq, E = conv * P; the reference's Calc1D forces are half of that. */
int pyiid_hmc_setup(pyiid_ctx *ctx, int slot, double qbin, const double *gobs, int potential,
                    double conv, const double *masses, int flags);
/* n_proposals HMC proposals per chain, each n_leapfrog velocity-Verlet steps
 * of size `step` (pyiid/sim/__init__.py:10-38: p += eps/2 F, x += eps p/m,
 * p += eps/2 F), momenta p ~ N(0, m/beta), accepted when
 * u < exp(beta (H0 - H1)) (pyiid/sim/hmc.py:66-92, beta = its `temp`).
 * Everything stays on the device; with use_graph the proposal is captured
 * once as a HIP graph and replayed.  energies / accepted (nullable):
 * n_proposals x n_chain, the energy after each decision and the decision.
 * times_ms (nullable, host): completion time of each proposal in ms since
 * the call started, from HIP events on the context stream. */
int pyiid_hmc_run(pyiid_ctx *ctx, int n_proposals, int n_leapfrog, double step, double beta,
                  uint64_t seed, int use_graph, double *energies, int32_t *accepted,
                  double *times_ms, int flags);

/* ---- learned HMC integrator --------------------------------------------- */
/* No counterpart in the reference: pyiid/sim/hmc.py:33-63 integrates with the
 * exact forces only.  Here a per-atom network drives the leapfrog and the
 * exact Rw (pyiid_hmc_setup's potential) still decides every Metropolis step
 * (hmc.py:66-92), so the sampled distribution is unchanged.
 *
 * Descriptors: D[i][m] = sum_{j != i, r_ij < rc} exp(-eta (r_ij - mu_m)^2) *
 * (cos(pi r_ij / rc) + 1) / 2, m < 16, mu_m = mu0 + m * dmu, over the atoms of
 * i's chain.  Network per atom: x = (D - d_mean) * d_inv_scale, 64 SiLU units,
 * 64 SiLU units, one output; weights in BF16, FP32 accumulation.  Surrogate
 * energy of a chain: sum over its atoms.  params = {rc, skin, mu0, dmu, eta}
 * (A, 1/A^2); neighbour lists hold pairs within rc + skin of a proposal's
 * start, and a proposal that moves an atom farther than skin / 2 is rejected.
 * w1: 16 x 64, b1: 64, w2: 64 x 64, b2: 64, d_mean / d_inv_scale: 16 (host
 * float32, rounded to BF16 here).  max_nbr: neighbour-list capacity per atom.
 * Resets the output-layer fit. */
int pyiid_sur_setup(pyiid_ctx *ctx, const float *params, const float *w1, const float *b1,
                    const float *w2, const float *b2, const float *d_mean,
                    const float *d_inv_scale, int32_t max_nbr);
/* Output weights w3 (n_chain x 64 floats; rounded to BF16), one row per chain. */
int pyiid_sur_set_output(pyiid_ctx *ctx, const float *w3, int flags);
/* Descriptors of the current positions (n_total x 16 floats). */
int pyiid_sur_descriptors(pyiid_ctx *ctx, float *d, int flags);
/* Surrogate forces -dE_s/dq (n_total x 3 doubles, nullable) and per-atom
 * network outputs (n_total floats, nullable) at the current positions. */
int pyiid_sur_forces(pyiid_ctx *ctx, double *forces, float *atom_energy, int flags);
/* Feature forces: row k (k < 64) holds the surrogate forces with w3 = e_k
 * for every chain, 64 x (3 n_total) floats. */
int pyiid_sur_features(pyiid_ctx *ctx, float *phi, int flags);
/* Feature energies: e[b][k] = sum over chain b's atoms of the network
 * output with w3 = e_k (the chain energies of the 64 output units),
 * n_chain x 64 doubles. */
int pyiid_sur_feature_energies(pyiid_ctx *ctx, double *e, int flags);
/* The output weights in use (n_chain x 64 floats, the BF16 values). */
int pyiid_sur_get_output(pyiid_ctx *ctx, float *w3, int flags);
/* Weight lambda (A^-2, >= 0; default 0, forces only) of the energy rows of
 * the fit (see pyiid_sur_train). */
int pyiid_sur_set_energy_weight(pyiid_ctx *ctx, double weight);
/* One training step at the current positions: exact energies E_b and forces
 * F_b (Debye path, as in pyiid_hmc_setup), feature forces Phi_b and feature
 * energies e_b.  Each chain's running normal equations are scaled by
 * `forget`, then take the force rows min |Phi_b w - F_b|^2 and, when an
 * earlier training step ran under the same targets and layout, the energy
 * row lambda |(e_b - e_b') w - (E_b - E_b')|^2 against that step's e_b', E_b'
 * (the exact Rw change between the two states).  Ridge solve
 * (ridge * trace / 64) for w3.  r2 (nullable, n_chain doubles):
 * 1 - |F - Phi w|^2 / |F|^2 on this sample's forces. */
int pyiid_sur_train(pyiid_ctx *ctx, double forget, double ridge, double *r2, int flags);
/* n_proposals proposals driven by the surrogate: forces from the network,
 * velocity Verlet and momenta as in pyiid_hmc_run, acceptance with the exact
 * energy of the end point.  A training step runs before proposal p whenever
 * p % train_every == 0 (train_every = 0: only if no fit exists).  The
 * proposals until the next step use exact forces (pyiid_hmc_run's proposal)
 * when the mean r2 of the training step is below min_r2, or when the
 * learned-force proposals since the previous step were accepted less often
 * than min_acceptance (0 disables the test; after an exact block the learned
 * forces are tried again).  With train_every > 3 the acceptance is also
 * checked after a block's first three learned-force proposals, and a block
 * that falls short there finishes with exact forces.  energies / accepted /
 * exact_mode (per proposal) / times_ms as in pyiid_hmc_run; violations
 * (n_chain int32, nullable): proposals rejected for leaving the neighbour
 * lists' range, counted since the layout was set. */
int pyiid_sur_hmc_run(pyiid_ctx *ctx, int n_proposals, int n_leapfrog, double step, double beta,
                      uint64_t seed, int use_graph, int train_every, double forget, double ridge,
                      double min_r2, double min_acceptance, double *energies, int32_t *accepted,
                      int32_t *exact_mode, double *times_ms, int32_t *violations, int flags);

/* ---- F(Q) without device allocation ------------------------------------ */
#define PYIID_NOALLOC_MAX_ATOMS 65536
#define PYIID_NOALLOC_MAX_ELEMENTS 32
#define PYIID_NOALLOC_MAX_BINS 4096
#define PYIID_NOALLOC_SCRATCH_MIB 37
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory, e.g. code under HIP graph capture, tight serving loops, or
 * harnesses that forbid hipMalloc.
 *
 * Computes F(Q) of one structure as pyiid_fq does.  There is no context and
 * no device allocation.
 *  - Arguments.  pos (float32 n x 3), elem (int32 n scatter-table rows;
 *    NULL means all 0), table (float32 n_elem x n_q) and fq (float32 n_q)
 *    are device pointers on `device`.
 *  - Stream.  The work is ordered on `stream` (a hipStream_t; NULL is the
 *    null stream), and the call returns without waiting for it.
 *  - Scratch.  Intermediates live in a static device array of at most
 *    PYIID_NOALLOC_SCRATCH_MIB MiB on each GPU, held until the process ends.
 *    Every call on a GPU shares that array, so calls that overlap on one GPU
 *    (on other streams or threads) race.
 *  - Limits.  Beyond the PYIID_NOALLOC_MAX_* limits the call returns
 *    PYIID_E_INVALID.  An element index outside [0, n_elem) makes every bin
 *    of fq NaN.
 *  - First call.  The first call on a device looks the array up; make it
 *    before capturing a graph.
 *  - Restores the caller's current device. */
int pyiid_fq_noalloc(int device, void *stream, const float *pos, const int32_t *elem, int64_t n,
                     const float *table, int32_t n_elem, int32_t n_q, double qbin, float *fq);

#ifdef __cplusplus
}
#endif
#endif /* PYIID_HIP_H */
