/*
 * MPS / QPS reader of libfa_cp.so.  Replaces QP/FA_CP/include/mps_parser.h:20
 * (implementation QP/FA_CP/src/mps_parser.c:371-562): reads free/fixed MPS
 * with QUADOBJ (lower triangle, mirrored) or QMATRIX (full) sections, plain or
 * gzip-compressed, into a freshly allocated qp_problem_t (free it with
 * qp_problem_free): A and Q in CSR with the entries of a row in file order, no
 * low-rank term.  OBJSENSE MAX negates c, Q and the constant.  NULL when the
 * file cannot be read or parsed.
 *
 * A program built against the reference includes this file in place of the
 * FA_CP mps_parser.h (same function name and signature).
 */
#pragma once

#include "fa_cp_types.h"

#ifdef __cplusplus
extern "C" {
#endif

qp_problem_t *read_mps_file(const char *filename);

#ifdef __cplusplus
}
#endif
