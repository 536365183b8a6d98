/*
 * Batched CP_AL: solve many independent LPs in one device-resident run.
 *
 * No counterpart in the reference API (the reference solves one LP per
 * process call, LP/CP_AL/include/cp_al.h:28-47).  Every instance follows the
 * single-instance algorithm of solve_lp_problem exactly (same rescaling, step
 * sizes, blocks of termination_evaluation_frequency iterations, restart and
 * termination rules), with its own state kept on the device, so a caller that
 * looped over solve_lp_problem gets the same per-instance results.  Two
 * differences: the power-iteration start vector of every instance is drawn
 * from a freshly seeded std::mt19937(1) (as after pdal_cp_al_reset_rng), and
 * time_sec_limit bounds the whole batch (instances still running when it
 * expires end with TIME_LIMIT).  The presolve and the environment-driven
 * controllers of the reference are not applied, as in solve_lp_problem here.
 *
 * Lifetime: create uploads, rescales and estimates every instance's step size
 * once; solve may be called repeatedly (each call restarts every instance
 * from its start point); the problems may be freed after create.
 */
#pragma once

#include "cp_al_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pdal_cp_al_batch pdal_cp_al_batch_t;

/* Set up a batch of `count` problems (primal_start / dual_start of each
 * problem, when present, are its start point).  params NULL = defaults.
 * Returns NULL on error (invalid parameters, > 2^31-1 total rows, columns or
 * nonzeros, device failure). */
pdal_cp_al_batch_t *pdal_cp_al_batch_create(lp_problem_t *const *problems, int count,
                                            const cp_al_parameters_t *params);

/* Replace the start points of all instances: x is the concatenation of the
 * instances' primal starts (sum of num_variables entries), y of their dual
 * starts; either may be NULL to keep the current one.  Returns 0 on success. */
int pdal_cp_al_batch_set_starts(pdal_cp_al_batch_t *batch, const double *x, const double *y);

/* Device addresses (current device) of the batch's unscaled start buffers:
 * *x holds sum n_i doubles and *y sum m_i, concatenated in instance order as
 * set_starts takes them.  A producer on the device (pdal_pod_expand_device)
 * writes the starts there, then pdal_cp_al_batch_set_starts_from_buffers
 * takes them for the next solves as set_starts would; no host copy.  The
 * buffers live as long as the batch.  Returns 0 on success. */
int pdal_cp_al_batch_start_buffers(pdal_cp_al_batch_t *batch, double **x, double **y);
int pdal_cp_al_batch_set_starts_from_buffers(pdal_cp_al_batch_t *batch);

/* Solve every instance; returns the number that reached OPTIMAL, or -1. */
int pdal_cp_al_batch_solve(pdal_cp_al_batch_t *batch);

/* Result of instance i of the last solve (free with cp_al_result_free);
 * NULL before the first solve or for an invalid index. */
cp_al_result_t *pdal_cp_al_batch_result(pdal_cp_al_batch_t *batch, int i);

/* out[0..17] = {set-up seconds, last solve seconds (wall clock, device
 * resident), hipGraph replays, instances, OPTIMAL instances, sum of
 * iteration counts, lanes per row of A, lanes per row of At, block-table
 * compactions (stopped instances are dropped from the launch tables and the
 * graph re-captured whenever the active count halves), seconds of the last
 * solve until half of the instances had stopped, seconds until 99 % had,
 * instances whose At.y / A.x_bar runs LDS-staged (one workgroup per instance,
 * gathered segment in LDS) while every instance runs: At.y staged, A.x_bar
 * staged whole, A.x_bar staged in two column halves, 1 when the batch runs
 * the shared-A engine (every instance has the same constraint matrix: the
 * instances' vectors side by side, one pass over A serves them all; up to 512
 * instances; its compactions pack the running instances into narrower
 * layouts), 1 when the last shared-A solve handed its last few running
 * instances (8 or fewer) to the block-diagonal engine, 1 when the shared-A
 * engine holds the variable bounds once for all instances (their scaled
 * bounds are equal bit for bit), 1 when its dual pass takes every row as an
 * equality (l = u bit for bit in every instance: the projection onto [l, u]
 * is l)}.  `out` holds 18 doubles. */
void pdal_cp_al_batch_stats(const pdal_cp_al_batch_t *batch, double *out);

/* Launch `iters` plain iterations over all instances kernel by kernel
 * (instances reset to their starts first) and time the two fused SpMV kernels
 * with hipEvents on the batch's stream, the first iteration excluded.
 * out[0..7] = {avg ms primal kernel (At.y over all instances), avg ms dual
 * kernel (A.x_bar), lanes per row of At, lanes per row of A, total nonzeros,
 * total columns, total rows, workgroups of the primal kernel}.  Returns 0 on
 * success.  A following solve starts from the starts again. */
int pdal_cp_al_batch_time_kernels(pdal_cp_al_batch_t *batch, int iters, double *out);

/* Primal weight of every instance (count doubles): after a solve, the value
 * the PID controller ended with; before the first solve, the start value. */
int pdal_cp_al_batch_primal_weights(const pdal_cp_al_batch_t *batch, double *out);

/* Start value of the primal weight per instance (entries that are not finite
 * and positive keep the default); the step schedule follows it.  A warm start
 * near the solution converges in the first blocks only with a primal weight
 * near the one a cold solve ends with. */
int pdal_cp_al_batch_set_primal_weights(pdal_cp_al_batch_t *batch, const double *pw);

/* New objective and bound vectors for the batch's matrices, each the
 * concatenation over the instances in instance order (c, var_lb, var_ub: sum
 * of num_variables entries; con_lb, con_ub: sum of num_constraints); NULL
 * keeps the current vector.  The matrices keep their scaling and step-size
 * estimate, which depend on the matrices only: the vectors are scaled as a
 * fresh pdal_cp_al_batch_create would scale them (the same scaling passes in
 * the same order, then the bound and objective scales), so the next solve
 * gives the bits of a batch created with these vectors, at the cost of a few
 * passes over the vectors instead of a full set-up.  The starts become zero
 * and the primal weights their defaults (set_starts and set_primal_weights
 * apply afterwards).  set_vectors reads host arrays, set_vectors_device device
 * arrays of the current device.  Returns 0 on success.  No counterpart in the
 * reference (one LP per call); it serves parametric families that share A
 * (b(theta), c(theta)) without paying the set-up per batch. */
int pdal_cp_al_batch_set_vectors(pdal_cp_al_batch_t *batch, const double *c, const double *con_lb,
                                 const double *con_ub, const double *var_lb, const double *var_ub);
int pdal_cp_al_batch_set_vectors_device(pdal_cp_al_batch_t *batch, const double *c, const double *con_lb,
                                        const double *con_ub, const double *var_lb, const double *var_ub);

void pdal_cp_al_batch_free(pdal_cp_al_batch_t *batch);

#ifdef __cplusplus
}
#endif
