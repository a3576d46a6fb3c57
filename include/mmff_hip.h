/*
 * C ABI of libmmff_hip.so -- batched MMFF94 conformer minimization on gfx950 (MI355X).
 *
 * Every buffer is a caller-owned DEVICE pointer (torch-ROCm tensors' data_ptr()). Hot
 * calls allocate nothing: workspace sizes are queried and the caller passes the
 * buffers. All calls are stream-ordered and asynchronous; they return 0 on success or a
 * negative mmffhip_status, with a message from mmffhip_last_error() (thread-local).
 * No C++ exception crosses this boundary.
 *
 * Replaced reference interfaces (mlxmolkit, file:line):
 *   mmffhip_energy_grad     <- mmff_energy_grad_metal / mmff_energy_grad_metal_mega
 *                              (mlxmolkit/mmff_metal_kernel.py:713-775) and the TG term
 *                              library (mlxmolkit/mmff_bfgs_header.metal:14-402)
 *   mmffhip_minimize_nk     <- the mmff_bfgs_tg / mmff_lbfgs_tg kernel launches of
 *                              mmff_minimize_nk (mlxmolkit/mmff_minimize.py:267-346)
 *   mmffhip_bb_fused        <- mmff_optimize_metal_fused (mlxmolkit/mmff_metal_kernel.py:1104-1150)
 *   mmffhip_multimol_*      <- the per-iteration mega-kernel launches + MLX L-BFGS of
 *                              mmff_optimize_metal_multi_mol (mlxmolkit/mmff_metal_optimizer.py:307-495)
 *                              and of mmff_optimize_metal_batch (:77-202)
 */
#ifndef MMFF_HIP_H
#define MMFF_HIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MMFFHIP_ABI_VERSION 2

enum mmffhip_status {
  MMFFHIP_OK = 0,
  MMFFHIP_ERR_ARG = -1,     /* invalid argument / inconsistent batch */
  MMFFHIP_ERR_HIP = -2,     /* a HIP runtime call failed */
  MMFFHIP_ERR_SIZE = -3,    /* molecule / workspace too large for the chosen plan */
  MMFFHIP_ERR_DEVICE = -4   /* no usable gfx950 device */
};

enum mmffhip_family { MMFFHIP_FAMILY_TG = 0, MMFFHIP_FAMILY_MEGA = 1 };
enum mmffhip_algo { MMFFHIP_BFGS = 0, MMFFHIP_LBFGS = 1 };
/* where k_minimize_nk keeps the dense inverse Hessian (L-BFGS: its history): AUTO = GLOBAL_SYM
 * for dense BFGS, and for L-BFGS the history in LDS when it fits, else in HBM; LDS = packed
 * upper triangle in LDS; GLOBAL = full matrix in HBM; GLOBAL_PACKED
 * = packed upper triangle in HBM (dense BFGS; L-BFGS treats it as GLOBAL); GLOBAL_SYM = one
 * 4x4 block per pair of row quads in HBM, half the bytes of GLOBAL (dense BFGS; L-BFGS treats
 * it as GLOBAL) */
enum mmffhip_hess_place {
  MMFFHIP_HESS_AUTO = 0,
  MMFFHIP_HESS_LDS = 1,
  MMFFHIP_HESS_GLOBAL = 2,
  MMFFHIP_HESS_GLOBAL_PACKED = 3,
  MMFFHIP_HESS_GLOBAL_SYM = 4
};

/* Device-side ragged batch. Per-molecule term tables are shared by that molecule's
 * conformers; conformer c reads molecule conf_mol[c] and its 3N coordinates start at
 * float offset conf_off[c]. mol_desc holds 16 int32 per molecule:
 *   [0] n_atoms  [1] bond_off [2] nb  [3] ang_off [4] na  [5] sb_off [6] nsb
 *   [7] oop_off  [8] noop     [9] tor_off [10] ntor [11] pair_off [12] npair
 * Term tables are SoA with atom indices packed as 16-bit pairs:
 *   bond_ij  u32  i | j<<16            bond_p  f32x2 {kb, r0}
 *   ang_idx  u32x2 {i|j<<16, k|lin<<16} ang_p  f32x2 {ka, theta0_deg}
 *   sb_idx   u32x2 {i|j<<16, k}        sb_p    f32x4 {kba_ijk, kba_kji, r0_ij, r0_kj}, sb_t0 f32 (deg)
 *   oop_idx  u32x2 {i|j<<16, k|l<<16}  oop_k   f32
 *   tor_idx  u32x2 {i|j<<16, k|l<<16}  tor_p   f32x4 {V1, V2, V3, 0}
 *   pair_ij  u32  i | j<<16            pair_p  f32x4 {R*, eps, q_i q_j * (0.75 if 1-4 else 1),
 *                                                    1 / max(R*, 1e-8)}
 * (vdW and electrostatic pairs of the reference are merged into one pair list.) */
typedef struct {
  const int32_t *mol_desc;
  const uint32_t *bond_ij;
  const void *bond_p;
  const void *ang_idx;
  const void *ang_p;
  const void *sb_idx;
  const void *sb_p;
  const float *sb_t0;
  const void *oop_idx;
  const float *oop_k;
  const void *tor_idx;
  const void *tor_p;
  const uint32_t *pair_ij;
  const void *pair_p;
  const int32_t *conf_mol;
  const int32_t *conf_off;
  int32_t n_mols;
  int32_t n_confs;
  int32_t max_atoms;  /* largest molecule in the batch */
} mmffhip_batch;

typedef struct {
  int32_t algo;        /* mmffhip_algo */
  int32_t max_iters;
  float grad_tol;
  int32_t diel;        /* 1 constant, 2 distance-dependent dielectric */
  int32_t lbfgs_m;     /* history length for MMFFHIP_LBFGS (reference nk kernel uses 8) */
  int32_t hess_place;  /* mmffhip_hess_place */
  int32_t block;       /* threads per workgroup, a multiple of 64 up to 1024 with an instantiated
                          kernel for (algo, placement, staging); 0 = plan default */
  int32_t ls_grad;     /* 1: evaluate E+G at every line-search trial (saves the separate
                          gradient pass when the first trial is accepted) */
  const float *precond;  /* optional per-atom 3x3 initial inverse-Hessian blocks, 9 floats per
                            atom at (conf_off/3 + atom)*9, in units of (0.1 * Hessian block)^-1
                            (the kernel multiplies them by 0.1 / initial gradient scale and
                            starts from their symmetric part (B + B^T) / 2); BFGS restarts from
                            the identity if a line search fails. NULL = identity (reference) */
  int32_t table_lds_bytes; /* >0: stage each molecule's term tables in LDS, reserving this many
                              bytes: all tables beside an LDS Hessian; beside a global Hessian
                              only the bonded tables (pairs stay in global memory), and only
                              when that keeps the workgroups per CU. Ignored when they do not
                              fit. A conformer whose staged tables need more gets status 2 and
                              E = NaN */
  const int32_t *xcd_offsets; /* optional host array of 9 rising offsets from 0 to n_confs that
                                 split `order` into 8 per-XCD work queues: a workgroup on XCD x
                                 (workgroups go to the 8 XCDs round-robin) takes conformers from
                                 order[xcd_offsets[x] .. xcd_offsets[x+1]) first, then from the
                                 other queues. NULL = one queue over `order` */
} mmffhip_nk_opts;

int mmffhip_abi_version(void);
const char *mmffhip_last_error(void);
/* Number of compute units and XCDs of the current device (for planners). */
int mmffhip_device_info(int32_t *n_cu, int32_t *n_xcd, int32_t *lds_bytes);

/* Single-point energies and gradients. e_out: n_confs doubles. e_terms (optional):
 * 7 doubles per conformer {bond, angle, strbend, oop, torsion, vdw, ele}. g_out (optional):
 * float gradient in the coordinate layout. term_mask selects terms (bit k = term k). */
int mmffhip_energy_grad(const mmffhip_batch *b, int family, int diel, const float *x, double *e_out,
                        double *e_terms, float *g_out, uint32_t term_mask, void *stream);

/* Workspace bytes mmffhip_minimize_nk needs for this batch and options. */
int mmffhip_minimize_nk_workspace(const mmffhip_batch *b, const mmffhip_nk_opts *o, size_t *bytes);

/* The launch mmffhip_minimize_nk would make for this batch and options: the persistent grid
 * (resident workgroups, at most one per conformer), threads per workgroup, LDS bytes per
 * workgroup and the Hessian placement chosen (MMFFHIP_HESS_LDS / _GLOBAL / _GLOBAL_PACKED /
 * _GLOBAL_SYM).
 * block, lds_bytes and hess_place may be NULL. No counterpart in the reference, which
 * launches one threadgroup per conformer (mlxmolkit/mmff_minimize.py:310-311,340-341): callers
 * use it to size their size buckets against the device's residency. */
int mmffhip_minimize_nk_plan(const mmffhip_batch *b, const mmffhip_nk_opts *o, int32_t *grid, int32_t *block,
                             int32_t *lds_bytes, int32_t *hess_place);

/* In-kernel BFGS / L-BFGS minimization of every conformer (reference run_mmff semantics).
 * x_in / x_out may alias. order: conformer processing order (length n_confs, e.g. longest
 * first), NULL = natural order. counters (optional): 3 int32 per conformer
 * {iterations, gradient evaluations, energy evaluations}. workspace: 256-byte aligned device
 * memory of at least the queried size (hipMalloc and PyTorch allocations are). */
int mmffhip_minimize_nk(const mmffhip_batch *b, const mmffhip_nk_opts *o, const float *x_in, float *x_out,
                        double *e_out, int32_t *status_out, int32_t *counters, const int32_t *order,
                        void *workspace, size_t workspace_bytes, void *stream);

/* mmffhip_minimize_nk without a workspace argument: the work queue and the global Hessian /
 * L-BFGS history slots live in a 64 MiB device array inside the library (one per GPU, held
 * until the process ends), so the call allocates nothing and needs no caller buffer beyond
 * inputs and outputs. No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). The persistent grid is capped to the slots that fit; a plan that needs
 * more than 64 MiB for one slot per compute unit (on 256 CUs: a full global Hessian above
 * n = 3N = 256, symmetric blocks above n = 356) fails with MMFFHIP_ERR_SIZE. Results are identical to mmffhip_minimize_nk's.
 * Every call on one GPU shares that GPU's array, which keeps the last call's contents: calls
 * that overlap on one GPU (concurrent streams or threads) race and must be serialised by the
 * caller. */
int mmffhip_minimize_nk_noalloc(const mmffhip_batch *b, const mmffhip_nk_opts *o, const float *x_in, float *x_out,
                                double *e_out, int32_t *status_out, int32_t *counters, const int32_t *order,
                                void *stream);

/* Barzilai-Borwein steepest descent of mmff_optimize_metal_fused (MEGA family).
 * e_out: energy of the last pre-step geometry, as the reference returns. */
int mmffhip_bb_fused(const mmffhip_batch *b, const float *x_in, float *x_out, double *e_out, int32_t max_iters,
                     float max_step, float grad_lr, void *stream);

/* Lock-step host-L-BFGS of mmff_optimize_metal_multi_mol, run on the device in chunks of
 * eval_freq iterations. State lives in a caller workspace; counters in flags decide the
 * global early exit without host synchronisation, so the whole sequence is capturable. */
int mmffhip_multimol_workspace(const mmffhip_batch *b, int32_t lbfgs_m, size_t *bytes);
int mmffhip_multimol_run(const mmffhip_batch *b, const float *x_in, float *x_best, double *e_final,
                         float *grad_norm, int32_t *per_conf_iters, int32_t *converged, int32_t *iters_used,
                         int32_t max_iters, int32_t eval_freq, float max_step_per_atom, int32_t lbfgs_m,
                         float grad_scale, float grad_tol, void *workspace, size_t workspace_bytes,
                         void *stream);

/* The same engine with flags. MMFFHIP_MM_BATCH gives mmff_optimize_metal_batch
 * (mlxmolkit/mmff_metal_optimizer.py:77-202): max_iters iterations with the best-x guard at
 * every sync point and no convergence exit, no polish pass, and grad_norm = ||g_sc||_2 of the
 * last iterate (:186); per_conf_iters = max_iters, converged = 0, iters_used = max_iters. */
#define MMFFHIP_MM_BATCH 1
int mmffhip_multimol_run_ex(const mmffhip_batch *b, const float *x_in, float *x_best, double *e_final,
                            float *grad_norm, int32_t *per_conf_iters, int32_t *converged, int32_t *iters_used,
                            int32_t max_iters, int32_t eval_freq, float max_step_per_atom, int32_t lbfgs_m,
                            float grad_scale, float grad_tol, int32_t flags, void *workspace,
                            size_t workspace_bytes, void *stream);

/* Start geometries on the device (restraint embedding, FP64): stands in for the reference
 * pipeline's device DG -> 4D->3D -> ETK stages (mlxmolkit/conformer_pipeline_v2.py:176-290),
 * so the minimizer's inputs are produced where it reads them. mol_desc: 4 int32 per molecule
 * {n_atoms, restraint offset, restraint count, 0}; restr_ij: i | j<<16; restr_d: target
 * distance d0 > 0 of an equality spring 100 (d - d0)^2, or -d_min for a lower-bound wall
 * 10 (d_min - d)^2 when d < d_min. Conformer c embeds molecule conf_mol[c] from
 * seeds[c] (splitmix64 stream) and writes its 3N floats at x_out + conf_off[c], with
 * N(0, noise) added to every coordinate. */
int mmffhip_embed(const int32_t *mol_desc, const uint32_t *restr_ij, const float *restr_d, const int32_t *conf_mol,
                  const int32_t *conf_off, const uint64_t *seeds, int32_t n_confs, int32_t max_atoms, float noise,
                  float *x_out, void *stream);

/* Learned initial inverse Hessian: a per-atom MLP (104 features -> 64 -> 64 -> 6, ReLU) whose 6
 * outputs are the Cholesky factor of a 3x3 SPD block (diagonal through softplus + 1e-3). No
 * counterpart in the reference; it produces the optional mmffhip_nk_opts.precond blocks.
 * params: n_params floats, row-major W1 [64][104], b1, W2 [64][64], b2, W3 [6][64], b3.
 * feat: n_atoms x 104 floats. blocks: n_atoms x 9 floats (row-major 3x3). */
int mmffhip_mlp_dims(int32_t *n_in, int32_t *n_hidden, int32_t *n_out, int32_t *n_params);
/* Inference on MFMA in BF16 with FP32 accumulation. */
int mmffhip_mlp_infer(const float *params, const float *feat, int32_t n_atoms, float *blocks, void *stream);
/* Training: FP32 forward and backward of sum_a weight_a * ||block_a - target_a||_F^2.
 * loss (optional): per-atom unweighted loss; grad (optional): n_params floats, the gradient
 * summed over atoms in a fixed order (deterministic); weight may be NULL (all ones). */
int mmffhip_mlp_workspace(int32_t n_atoms, size_t *bytes);
int mmffhip_mlp_forward_backward(const float *params, const float *feat, int32_t n_atoms, const float *target,
                                 const float *weight, float *blocks, float *loss, float *grad, void *workspace,
                                 size_t workspace_bytes, void *stream);

#ifdef __cplusplus
}
#endif
#endif
