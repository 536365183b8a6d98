/*
 * MPS / QPS reader of librhr_fa_cp.so (the planted-KKT FA_CP).  Replaces
 * QP/PLANTED_FA_CP/include/mps_parser.h (implementation
 * QP/PLANTED_FA_CP/src/mps_parser.c, the FA_CP reader): the reader of
 * libfa_cp.so; free the problem with qp_problem_free.  NULL when the file
 * cannot be read or parsed.
 */
#pragma once

#include "planted_fa_cp_types.h"

#ifdef __cplusplus
extern "C" {
#endif

qp_problem_t *read_mps_file(const char *filename);

#ifdef __cplusplus
}
#endif
