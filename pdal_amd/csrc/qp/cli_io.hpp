// Output files of the QP runners (fa_cp, rhr_lin_cp_al, rhr_fa_cp): the
// instance name, `<name>_summary.txt` and the `%.10g` solution files.
//
// Reference: QP/FA_CP/src/cli.c:80-99 (instance name = basename up to the
// first dot), :101-118 (solution files), :120-166 (summary keys); the variants'
// runners write the same files (QP/LIN_CP_AL/src/cli.c:42-150,
// QP/PLANTED_FA_CP/src/cli.c:78-186) without FA_CP's three sigma lines.
#pragma once

#include <cstdio>
#include <string>

namespace pdal {
namespace qpcli {

inline std::string instance_name(const char *path)
{
    std::string s(path);
    const size_t slash = s.find_last_of('/');
    if (slash != std::string::npos) s = s.substr(slash + 1);
    const size_t dot = s.find('.');
    if (dot != std::string::npos) s = s.substr(0, dot);
    return s;
}

inline std::string out_path(const char *dir, const std::string &name, const char *suffix)
{
    return std::string(dir) + "/" + name + suffix;
}

inline void save_solution(const double *v, int n, const char *dir, const std::string &name, const char *suffix)
{
    if (!v) return;
    const std::string path = out_path(dir, name, suffix);
    FILE *f = fopen(path.c_str(), "w");
    if (!f) {
        perror("Error opening solution file");
        return;
    }
    for (int i = 0; i < n; ++i) fprintf(f, "%.10g\n", v[i]);
    fclose(f);
}

// `extra(f)` writes the lines a runner adds after the objective gaps (FA_CP:
// the final AL sigma, the sigma rule and its update count).
template <class Result, class Extra>
void save_summary(const Result *r, const char *reason, const char *dir, const std::string &name, Extra extra)
{
    const std::string path = out_path(dir, name, "_summary.txt");
    FILE *f = fopen(path.c_str(), "w");
    if (!f) {
        perror("Error opening summary file");
        return;
    }
    fprintf(f, "Termination Reason: %s\n", reason);
    fprintf(f, "Runtime (sec): %e\n", r->cumulative_time_sec);
    fprintf(f, "Iterations Count: %d\n", r->total_count);
    if (r->total_inner_count > 0) fprintf(f, "Inner Iterations Count: %d\n", r->total_inner_count);
    fprintf(f, "Primal Objective Value: %e\n", r->primal_objective_value);
    fprintf(f, "Dual Objective Value: %e\n", r->dual_objective_value);
    fprintf(f, "Relative Primal Residual: %e\n", r->relative_primal_residual);
    fprintf(f, "Relative Dual Residual: %e\n", r->relative_dual_residual);
    fprintf(f, "Absolute Objective Gap: %e\n", r->objective_gap);
    fprintf(f, "Relative Objective Gap: %e\n", r->relative_objective_gap);
    extra(f);
    fprintf(f, "Rows: %d\n", r->num_constraints);
    fprintf(f, "Columns: %d\n", r->num_variables);
    fprintf(f, "Nonzeros: %d\n", r->num_nonzeros);
    if (r->feasibility_polishing_time > 0.0) {
        fprintf(f, "Feasibility Polishing Time (sec): %e\n", r->feasibility_polishing_time);
        fprintf(f, "Feasibility Polishing Iteration Count: %d\n", r->feasibility_iteration);
    }
    fclose(f);
}

}  // namespace qpcli
}  // namespace pdal
