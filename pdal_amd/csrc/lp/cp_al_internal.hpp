// Internal declarations shared by the translation units of libcp_al.so.
#pragma once

#include "cp_al_types.h"

extern "C" const char *termination_reason_to_string(termination_reason_t reason);

namespace pdal {
namespace lp {

// Environment helpers with the reference's parsing rules (solver.cu:377-413).
double env_double(const char *name, double def);
int env_int(const char *name, int def);
bool env_set(const char *name);

// Warn once per solve about reference controllers this build does not run.
void warn_unsupported_controllers(bool batch);  // batch: the batched solver (fewer controllers)

// Plain-path solve on the current HIP device (throws pdal::HipError).
cp_al_result_t *optimize_gpu(const cp_al_parameters_t *params, lp_problem_t *problem);

// Reseed the process-global power-iteration generator to mt19937(1).
void reset_rng();

// Set up the device state and time `iters` iterations kernel by kernel.
int time_kernels_gpu(const cp_al_parameters_t *params, const lp_problem_t *problem, int iters, double *out);

// Device-resident session: set up once, solve many times (pdal_lp.h).
void *session_create(const cp_al_parameters_t *params, lp_problem_t *problem);
cp_al_result_t *session_solve(void *session);
int session_time_kernels(void *session, int iters, double *out);
double session_setup_seconds(void *session);
void session_free(void *session);

// Plain copy kernel timing for counter calibration (pdal_lp.h).
double copy_calibration(long long bytes, int reps);
double gather_calibration(long long count, long long len, int reps);

// Batched solver (pdal_batch.h).
void *batch_create(lp_problem_t *const *problems, int count, const cp_al_parameters_t *params);
int batch_solve(void *batch);
void batch_set_starts(void *batch, const double *x, const double *y);
void batch_start_buffers(void *batch, double **x, double **y);
void batch_set_starts_from_buffers(void *batch);
cp_al_result_t *batch_result(void *batch, int i);
void batch_stats(void *batch, double *out);
void batch_free(void *batch);
void batch_time_kernels(void *batch, int iters, double *out);
void batch_primal_weights(void *batch, double *out);
void batch_set_primal_weights(void *batch, const double *pw);
void batch_set_vectors(void *batch, const double *c, const double *l, const double *u, const double *lb,
                       const double *ub, bool device);

}  // namespace lp
}  // namespace pdal
