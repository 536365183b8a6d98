/*
 * MPS reader of libcp_al.so.  Replaces LP/CP_AL/include/mps_parser.h:22
 * (implementation LP/CP_AL/src/mps_parser.c:402): reads free/fixed MPS, plain
 * or gzip-compressed, into a freshly allocated CSR lp_problem_t.
 */
#pragma once

#include "cp_al_types.h"

#ifdef __cplusplus
extern "C" {
#endif

lp_problem_t *read_mps_file(const char *filename);

#ifdef __cplusplus
}
#endif
