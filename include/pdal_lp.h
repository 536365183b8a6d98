/*
 * Additional C entry points of libcp_al.so that have no counterpart in the
 * reference API.  They exist for tests and measurement only; a drop-in caller
 * of cp_al.h never needs them.
 */
#pragma once

#include "cp_al_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Reseed the process-global power-iteration generator to std::mt19937(1).
 * The reference keeps that generator for the life of the process
 * (LP/CP_AL/src/utils.cu:27-28), so the k-th solve of a process starts its
 * power iteration from the k-th block of draws; tests call this before each
 * solve to make runs independent of call order. */
void pdal_cp_al_reset_rng(void);

/* Set up the device state for `prob` (upload, transpose, rescale, power
 * iteration) and launch `iters` plain iterations kernel by kernel, timing each
 * fused SpMV kernel with hipEvents on the solver's stream.  out[0..6] =
 * {avg ms primal (At.y + primal update), avg ms dual (A.x_bar + dual update),
 * avg ms predictor (0 if beta == 0), avg ms per iteration, ms of one
 * evaluation pass, lanes per row for At, lanes per row for A}; a negative
 * lane count -S means the pass runs over the column-sliced copy of the matrix
 * with S column slices (one workgroup per row range; taken when the gathered
 * vector is far larger than an XCD's L2 and that pass is faster).
 * Returns 0 on success, negative on error. */
int pdal_cp_al_time_kernels(const lp_problem_t *prob, const cp_al_parameters_t *params, int iters, double *out);

/* Device-resident solve session.  create: upload, transpose, rescale and power
 * iteration for `prob` (which the caller keeps alive until free).  solve: run
 * the main loop from the problem's start point and return a result exactly as
 * solve_lp_problem would (free it with cp_al_result_free); may be called any
 * number of times.  time_kernels: as pdal_cp_al_time_kernels on the session's
 * current iterate.  setup_seconds: wall time of create.  NULL / negative on
 * error. */
void *pdal_cp_al_session_create(lp_problem_t *prob, const cp_al_parameters_t *params);
cp_al_result_t *pdal_cp_al_session_solve(void *session);
int pdal_cp_al_session_time_kernels(void *session, int iters, double *out);
double pdal_cp_al_session_setup_seconds(void *session);
void pdal_cp_al_session_free(void *session);

/* reference: LP/CP_AL/src/utils.cu:230-253 (exported there through utils.h) */
const char *termination_reason_to_string(termination_reason_t reason);

/* Counter calibration: copy `bytes` bytes (rounded down to whole doubles)
 * between two fresh device buffers with a plain grid-stride copy kernel,
 * `reps` times; returns the average kernel time in ms (hipEvents), or a
 * negative value on error.  The FETCH_SIZE / WRITE_SIZE counters of these
 * launches read `bytes` each when the counters are exact, which is how the
 * HBM traffic of the solver kernels is scaled. */
double pdal_copy_calibration(long long bytes, int reps);

/* Gather calibration: `count` FP64 loads at uniformly random indices of a
 * `len`-entry device vector (4-byte indices streamed, one running sum per
 * thread), `reps` times, with 1, 4 and 8 independent loads in flight per
 * thread over three grid sizes; returns the fastest average kernel time in ms
 * (hipEvents), or a negative value on error.  It is the floor of an SpMV pass whose column
 * indices are uniform: that pass makes the same loads and streams the values
 * and the result besides. */
double pdal_gather_calibration(long long count, long long len, int reps);

#ifdef __cplusplus
}
#endif
