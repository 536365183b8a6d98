// Which QP solver a translation unit builds: FA_CP (libfa_cp.so, the default)
// or, with -DPDAL_LIN_CP_AL, the linearized CP-AL solver RHR-Lin-CP-AL
// (librhr_lin_cp_al.so).  The two share problem construction, scaling, power
// iterations, residuals, infeasibility detection and termination; they differ
// in their parameter and result types and in the iteration itself.
#pragma once

#ifdef PDAL_LIN_CP_AL
#include "lin_cp_al.h"
#else
#include "fa_cp.h"
#endif

namespace pdal {
namespace qp {

#ifdef PDAL_LIN_CP_AL
using params_t = cp_al_parameters_t;
using result_t = cp_al_result_t;
// the shared code names the Q types by their FA_CP names (same values)
constexpr quad_obj_type_t FA_CP_SPARSE_Q = CP_AL_SPARSE_Q;
constexpr quad_obj_type_t FA_CP_DIAG_Q = CP_AL_DIAG_Q;
constexpr quad_obj_type_t FA_CP_LOW_RANK_PLUS_SPARSE_Q = CP_AL_LOW_RANK_PLUS_SPARSE_Q;
constexpr quad_obj_type_t FA_CP_LOW_RANK_Q = CP_AL_LOW_RANK_Q;
constexpr quad_obj_type_t FA_CP_NON_Q = CP_AL_NON_Q;
constexpr bool kLinearized = true;
#else
using params_t = fa_cp_parameters_t;
using result_t = fa_cp_result_t;
constexpr bool kLinearized = false;
#endif

}  // namespace qp
}  // namespace pdal
