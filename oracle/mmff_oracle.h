/*
 * FP64 CPU oracle for batched MMFF94 minimization.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker the HIP product path
 * is compared against. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it. The product path never links,
 * loads or falls back to it.
 *
 * Everything here is double precision. Parameters arrive as float32 exactly
 * as the reference's MMFFParams stores them (mlxmolkit/mmff_params.py:30-81)
 * and are promoted to double on use.
 */
#ifndef MMFF_ORACLE_H
#define MMFF_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One molecule's term tables, field-for-field the reference MMFFParams
 * (mlxmolkit/mmff_params.py:34-80). Angles are in degrees. */
typedef struct {
    int32_t n_atoms;
    int32_t nb;   const int32_t *b_i, *b_j;         const float *b_kb, *b_r0;
    int32_t na;   const int32_t *a_i, *a_j, *a_k;   const float *a_ka, *a_t0; const int32_t *a_lin;
    int32_t nsb;  const int32_t *s_i, *s_j, *s_k;   const float *s_kij, *s_kkj, *s_r0ij, *s_r0kj, *s_t0;
    int32_t noop; const int32_t *o_i, *o_j, *o_k, *o_l; const float *o_koop;
    int32_t nt;   const int32_t *t_i, *t_j, *t_k, *t_l; const float *t_v1, *t_v2, *t_v3;
    int32_t nv;   const int32_t *v_i, *v_j;         const float *v_r, *v_eps;
    int32_t ne;   const int32_t *e_i, *e_j;         const float *e_qq; const int32_t *e_14;
} oracle_mol;

/* Formula families (the two families of the reference, see term docs in .c) */
enum { ORACLE_FAMILY_TG = 0,   /* mlxmolkit/mmff_bfgs_header.metal (run_mmff / nk path) */
       ORACLE_FAMILY_MEGA = 1  /* mlxmolkit/mmff_metal_kernel.py:274-505 (multi_mol / fused) */ };

/* term mask bits */
enum { T_BOND = 1, T_ANGLE = 2, T_STRBEND = 4, T_OOP = 8, T_TORSION = 16, T_VDW = 32, T_ELE = 64, T_ALL = 127 };

/* Energy and (optionally) gradient of one conformer. x and g are (3N).
 * e_terms (optional, length 7) receives the per-term-type energies.
 * diel: 1 constant, 2 distance-dependent (TG family only; MEGA is always 1). */
/* Analysis switch (default off): emulate the HIP kernels' FP32 coordinates and FP32 term
 * results inside the FP64 oracle (see mmff_oracle.c). */
void oracle_set_fp32_emulation(int on);

double oracle_energy_grad(const oracle_mol *m, int family, int diel, const double *x,
                          double *g, uint32_t term_mask, double *e_terms);

/* Per-term-type sums of |term energy| (scale of FP32 rounding in each total). */
void oracle_energy_terms_abs(const oracle_mol *m, int family, int diel, const double *x, uint32_t term_mask,
                             double *abs7);

/* In-kernel dense BFGS of mlxmolkit/mmff_bfgs_source_tg.metal:1-372.
 * x (3N) in/out. Returns status (0 converged / 1 hit max_iters). */
int oracle_bfgs_nk(const oracle_mol *m, double *x, int max_iters, double grad_tol, int diel,
                   double *e_out, int32_t *counters /* [iters, grads, evals] or NULL */);

/* The same BFGS started from per-atom 3x3 inverse-Hessian blocks h0 (9 doubles per atom) scaled by
 * 0.1 / initial gradient scale, restarting from the identity if a line search fails: the learned
 * preconditioner of the HIP minimizer (no counterpart in the reference). h0 == NULL is
 * oracle_bfgs_nk. */
int oracle_bfgs_nk_h0(const oracle_mol *m, double *x, int max_iters, double grad_tol, int diel, const double *h0,
                      double *e_out, int32_t *counters);

/* In-kernel L-BFGS of mlxmolkit/mmff_lbfgs_source_tg.metal:1-298 with history m_hist. */
int oracle_lbfgs_nk(const oracle_mol *m, double *x, int max_iters, double grad_tol, int diel,
                    int m_hist, double *e_out, int32_t *counters);

/* Barzilai-Borwein steepest descent of mlxmolkit/mmff_metal_kernel.py:863-1073. */
void oracle_bb_fused(const oracle_mol *m, double *x, int max_iters, double max_step,
                     double grad_lr, double *e_out);

/* Threaded batch drivers. confs are ragged: conformer c uses mols[conf_mol[c]] and
 * its coordinates start at x[conf_off[c]]. */
void oracle_bfgs_nk_batch(const oracle_mol *mols, const int32_t *conf_mol, const int64_t *conf_off,
                          int n_confs, double *x, int max_iters, double grad_tol, int diel,
                          int use_lbfgs, int m_hist, double *e_out, int32_t *status_out,
                          int32_t *counters_out /* 3 per conf or NULL */, int n_threads);

/* oracle_bfgs_nk_batch with learned blocks h0 at (conf_off[c] / 3 + atom) * 9 (dense BFGS only). */
void oracle_bfgs_nk_batch_h0(const oracle_mol *mols, const int32_t *conf_mol, const int64_t *conf_off, int n_confs,
                             double *x, int max_iters, double grad_tol, int diel, int use_lbfgs, int m_hist,
                             const double *h0, double *e_out, int32_t *status_out, int32_t *counters_out,
                             int n_threads);

void oracle_energy_grad_batch(const oracle_mol *mols, const int32_t *conf_mol, const int64_t *conf_off,
                              int n_confs, int family, int diel, const double *x, double *e_out,
                              double *g_out /* same layout as x, or NULL */, uint32_t term_mask,
                              double *e_terms_out /* 7 per conf or NULL */, int n_threads);

void oracle_bb_fused_batch(const oracle_mol *mols, const int32_t *conf_mol, const int64_t *conf_off,
                           int n_confs, double *x, int max_iters, double max_step, double grad_lr,
                           double *e_out, int n_threads);

/* Host-driven lock-step L-BFGS of mmff_optimize_metal_multi_mol
 * (mlxmolkit/mmff_metal_optimizer.py:217-535), MEGA family.
 * Outputs: x -> best positions, e_out final energy at best, gnorm_out per-conformer
 * max per-atom scaled gradient norm of the last iterate, per_conf_iters, conv.
 * Returns iters_used of the main loop. */
int oracle_multi_mol(const oracle_mol *mols, const int32_t *conf_mol, const int64_t *conf_off,
                     int n_confs, double *x, int max_iters, int eval_freq, double max_step_per_atom,
                     int lbfgs_m, double grad_scale, double grad_tol, double *e_out,
                     double *gnorm_out, int32_t *per_conf_iters, int32_t *conv, int n_threads);

/* The same engine with flags; ORACLE_MM_BATCH gives mmff_optimize_metal_batch
 * (mlxmolkit/mmff_metal_optimizer.py:77-202): no convergence exit, no polish pass, gnorm_out =
 * ||g_sc||_2 of the last iterate. */
#define ORACLE_MM_BATCH 1
int oracle_multi_mol_ex(const oracle_mol *mols, const int32_t *conf_mol, const int64_t *conf_off,
                        int n_confs, double *x, int max_iters, int eval_freq, double max_step_per_atom,
                        int lbfgs_m, double grad_scale, double grad_tol, double *e_out,
                        double *gnorm_out, int32_t *per_conf_iters, int32_t *conv, int n_threads, int flags);

#ifdef __cplusplus
}
#endif
#endif
