// MPS / QPS reader of libfa_cp.so (QP/FA_CP/src/mps_parser.c:371-562): the
// shared reader's triplets and vectors as a freshly allocated qp_problem_t.
// A (m rows) and Q (n rows) are CSR with the entries of a row in file order;
// the low-rank term is empty (rank 0, a one-entry row pointer), as the
// reference's reader leaves it.
#include "fa_cp_mps_parser.h"

#include <cstdlib>

#include "../common/mps_reader.hpp"

#define PDAL_API extern "C" __attribute__((visibility("default")))

PDAL_API qp_problem_t *read_mps_file(const char *filename)
{
    pdal::mps::MpsData d;
    if (!pdal::mps::read_mps(filename, d, true)) return nullptr;
    qp_problem_t *p = (qp_problem_t *)calloc(1, sizeof(qp_problem_t));
    p->num_variables = d.n;
    p->num_constraints = d.m;
    p->num_rank_lowrank_obj = 0;
    p->objective_constant = d.obj_const;
    p->objective_vector = pdal::mps::malloc_copy(d.obj);
    p->variable_lower_bound = pdal::mps::malloc_copy(d.vlo);
    p->variable_upper_bound = pdal::mps::malloc_copy(d.vhi);
    p->constraint_lower_bound = pdal::mps::malloc_copy(d.lo);
    p->constraint_upper_bound = pdal::mps::malloc_copy(d.hi);
    p->constraint_matrix = (CsrComponent *)calloc(1, sizeof(CsrComponent));
    p->constraint_matrix_num_nonzeros = (int)d.av.size();
    pdal::mps::triplets_to_csr(d.m, d.ai, d.aj, d.av, &p->constraint_matrix->row_ptr,
                               &p->constraint_matrix->col_ind, &p->constraint_matrix->val);
    p->objective_sparse_matrix = (CsrComponent *)calloc(1, sizeof(CsrComponent));
    p->objective_sparse_matrix_num_nonzeros = (int)d.qv.size();
    pdal::mps::triplets_to_csr(d.n, d.qi, d.qj, d.qv, &p->objective_sparse_matrix->row_ptr,
                               &p->objective_sparse_matrix->col_ind, &p->objective_sparse_matrix->val);
    p->objective_lowrank_matrix = (CsrComponent *)calloc(1, sizeof(CsrComponent));
    p->objective_lowrank_matrix->row_ptr = (int *)calloc(1, sizeof(int));
    p->objective_lowrank_matrix_num_nonzeros = 0;
    return p;
}
