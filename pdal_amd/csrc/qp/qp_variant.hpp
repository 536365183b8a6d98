// Which QP solver a translation unit builds: FA_CP (libfa_cp.so, the default);
// with -DPDAL_LIN_CP_AL the linearized CP-AL solver RHR-Lin-CP-AL
// (librhr_lin_cp_al.so); with -DPDAL_PLANTED_FA_CP the planted-KKT FA_CP
// RHR-FA-CP (librhr_fa_cp.so); with -DPDAL_PLANTED_CP_AL the planted-KKT CP-AL
// RHR-CP-AL (librhr_cp_al.so).  They share problem construction, scaling, power
// iterations, residuals, infeasibility detection and termination; they differ
// in their parameter and result types and in the iteration itself.
// PDAL_CP_AL_TYPES marks the builds with the cp_al_* parameter and result
// types (the linearized and the planted-KKT CP-AL solvers).
#pragma once

#if defined(PDAL_LIN_CP_AL)
#define PDAL_CP_AL_TYPES 1
#include "lin_cp_al.h"
#elif defined(PDAL_PLANTED_CP_AL)
#define PDAL_CP_AL_TYPES 1
#include "planted_cp_al.h"
#elif defined(PDAL_PLANTED_FA_CP)
#include "planted_fa_cp.h"
#else
#include "fa_cp.h"
#endif

namespace pdal {
namespace qp {

#ifdef PDAL_CP_AL_TYPES
using params_t = cp_al_parameters_t;
using result_t = cp_al_result_t;
// the shared code names the Q types by their FA_CP names (same values)
constexpr quad_obj_type_t FA_CP_SPARSE_Q = CP_AL_SPARSE_Q;
constexpr quad_obj_type_t FA_CP_DIAG_Q = CP_AL_DIAG_Q;
constexpr quad_obj_type_t FA_CP_LOW_RANK_PLUS_SPARSE_Q = CP_AL_LOW_RANK_PLUS_SPARSE_Q;
constexpr quad_obj_type_t FA_CP_LOW_RANK_Q = CP_AL_LOW_RANK_Q;
constexpr quad_obj_type_t FA_CP_NON_Q = CP_AL_NON_Q;
constexpr bool kLinearized = true;  // sigma / lambda parametrized steps
#else
using params_t = fa_cp_parameters_t;
using result_t = fa_cp_result_t;
constexpr bool kLinearized = false;
#endif

#ifdef PDAL_PLANTED_CP_AL
constexpr bool kPlantedCpAl = true;
#else
constexpr bool kPlantedCpAl = false;
#endif

// Planted-KKT FA_CP: the AL kernels treat every row as an equality with the
// lower bound as right-hand side (QP/PLANTED_FA_CP/src/fa_cp_core_op.cu:147-183).
#ifdef PDAL_PLANTED_FA_CP
constexpr bool kEqualityAL = true;
#else
constexpr bool kEqualityAL = false;
#endif

}  // namespace qp
}  // namespace pdal
