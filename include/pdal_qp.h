/*
 * Additional C entry points of libfa_cp.so that have no counterpart in the
 * reference API.  They exist for tests and measurement only; a drop-in caller
 * of fa_cp.h never needs them.
 */
#pragma once

#include "fa_cp_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Reseed the power-iteration generators: the process-global std::mt19937(1)
 * with its normal-distribution cache (start vector of the A power iteration,
 * QP/FA_CP/src/utils.cu:28-29) and glibc's rand() via srand(1) (start vectors
 * of the Q eigenvalue estimates, utils.cu:184, 254).  The reference never
 * reseeds, so the k-th solve of a process continues both streams; tests call
 * this before each solve to make runs independent of call order. */
void pdal_fa_cp_reset_rng(void);

/* Name of a termination reason ("OPTIMAL", "ITERATION_LIMIT", ...). */
const char *pdal_fa_cp_termination_reason_to_string(termination_reason_t reason);

#ifdef __cplusplus
}
#endif
