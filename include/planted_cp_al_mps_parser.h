/*
 * MPS / QPS reader of librhr_cp_al.so.  Replaces
 * QP/PLANTED_CP_AL/include/mps_parser.h (implementation QP/PLANTED_CP_AL/src/mps_parser.c,
 * the FA_CP reader): the reader of libfa_cp.so, returning the planted-KKT
 * CP-AL solver's qp_problem_t (same layout); free it with qp_problem_free.  NULL when
 * the file cannot be read or parsed.
 */
#pragma once

#include "planted_cp_al_types.h"

#ifdef __cplusplus
extern "C" {
#endif

qp_problem_t *read_mps_file(const char *filename);

#ifdef __cplusplus
}
#endif
