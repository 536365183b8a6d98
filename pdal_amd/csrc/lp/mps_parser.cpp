// MPS / MPS.gz reader of libcp_al.so (LP/CP_AL/src/mps_parser.c:402-900):
// the shared reader's triplets and vectors as a freshly allocated CSR
// lp_problem_t.  Quadratic sections are not known to the LP reader: their
// lines fall outside every section.
#include "mps_parser.h"

#include <cstdlib>

#include "../common/mps_reader.hpp"

#define PDAL_API extern "C" __attribute__((visibility("default")))

PDAL_API lp_problem_t *read_mps_file(const char *filename)
{
    pdal::mps::MpsData d;
    if (!pdal::mps::read_mps(filename, d, false)) return nullptr;
    lp_problem_t *pr = (lp_problem_t *)calloc(1, sizeof(lp_problem_t));
    pr->num_variables = d.n;
    pr->num_constraints = d.m;
    pr->constraint_matrix_num_nonzeros = (int)d.av.size();
    pr->objective_constant = d.obj_const;
    pr->objective_vector = pdal::mps::malloc_copy(d.obj);
    pr->variable_lower_bound = pdal::mps::malloc_copy(d.vlo);
    pr->variable_upper_bound = pdal::mps::malloc_copy(d.vhi);
    pr->constraint_lower_bound = pdal::mps::malloc_copy(d.lo);
    pr->constraint_upper_bound = pdal::mps::malloc_copy(d.hi);
    pdal::mps::triplets_to_csr(d.m, d.ai, d.aj, d.av, &pr->constraint_matrix_row_pointers,
                               &pr->constraint_matrix_col_indices, &pr->constraint_matrix_values);
    return pr;
}
