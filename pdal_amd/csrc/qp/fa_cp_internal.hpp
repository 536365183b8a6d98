// Internal interface between the FA_CP C API (fa_cp_api.cpp, host-only C++)
// and the HIP solver (fa_cp_solver.hip).
#pragma once

#include "qp_variant.hpp"

namespace pdal {
namespace qp {

// Solve on the current HIP device; throws pdal::HipError / std::exception on failure.
result_t *optimize_gpu(const params_t *params, const qp_problem_t *prob);

// Re-seed the power-iteration generators (std::mt19937(1) + normal cache, srand(1)).
void reset_rng();

const char *termination_reason_name(termination_reason_t r);

}  // namespace qp
}  // namespace pdal
