// fa_cp: the QP runner command line, linked against libfa_cp.so.
//
// Reference: QP/FA_CP/src/cli.c:260-455 (getopt flags, defaults from
// set_default_parameters, read_mps_file -> optimize), :38-78 (sigma-rule
// names); the output files are written by cli_io.hpp, the trajectory CSV by
// the solver.
#include "fa_cp.h"
#include "fa_cp_mps_parser.h"
#include "pdal_qp.h"
#include "cli_io.hpp"

#include <getopt.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

namespace {

struct RuleName {
    const char *name;
    sigma_update_rule_t rule;
};

// Accepted spellings of --sigma_update (cli.c:38-78); the first spelling of a
// rule is the one the summary prints.
const RuleName kRules[] = {
    {"fixed", FA_CP_SIGMA_UPDATE_FIXED},
    {"none", FA_CP_SIGMA_UPDATE_FIXED},
    {"residual", FA_CP_SIGMA_UPDATE_RESIDUAL_BALANCE},
    {"kkt", FA_CP_SIGMA_UPDATE_KKT_BALANCE},
    {"displacement", FA_CP_SIGMA_UPDATE_DISPLACEMENT},
    {"progress", FA_CP_SIGMA_UPDATE_PROGRESS},
    {"greedy", FA_CP_SIGMA_UPDATE_GREEDY_SCORE},
    {"guarded", FA_CP_SIGMA_UPDATE_GUARDED},
    {"zero-aware", FA_CP_SIGMA_UPDATE_GUARDED},
    {"zero_aware", FA_CP_SIGMA_UPDATE_GUARDED},
    {"balanced_guarded", FA_CP_SIGMA_UPDATE_BALANCED_GUARDED},
    {"balanced-guarded", FA_CP_SIGMA_UPDATE_BALANCED_GUARDED},
    {"guarded_balanced", FA_CP_SIGMA_UPDATE_BALANCED_GUARDED},
    {"scaled_guarded", FA_CP_SIGMA_UPDATE_BALANCED_GUARDED},
};

bool parse_rule(const char *s, sigma_update_rule_t *out)
{
    for (const RuleName &r : kRules)
        if (strcmp(s, r.name) == 0) {
            *out = r.rule;
            return true;
        }
    return false;
}

const char *rule_name(sigma_update_rule_t rule)
{   // sigma_update_rule_to_string (QP/FA_CP/src/utils.cu:390-410)
    for (const RuleName &r : kRules)
        if (r.rule == rule) return r.name;
    return "unknown";
}

void usage(const char *prog)
{
    fprintf(stderr, "Usage: %s [OPTIONS] <mps_file> <output_dir>\n\n", prog);
    fprintf(stderr,
            "Arguments:\n"
            "  <mps_file>               Path to the input problem in MPS format (.mps .QPS or .mps.gz).\n"
            "  <output_dir>             Directory for <basename>_summary.txt, <basename>_primal_solution.txt,\n"
            "                           <basename>_dual_solution.txt.\n\n"
            "Options:\n"
            "  -h, --help                          Display this help message.\n"
            "  -v, --verbose <int>                 Verbosity level: 0=silent, 1=summary, 2=detailed (default: 1).\n"
            "      --time_limit <seconds>          Time limit in seconds (default: 3600.0).\n");
    fprintf(stderr, "      --iter_limit <iterations>       Iteration limit (default: %d).\n", INT32_MAX);
    fprintf(stderr,
            "      --eps_opt <tolerance>           Relative optimality tolerance (default: 1e-4).\n"
            "      --eps_feas <tolerance>          Relative feasibility tolerance (default: 1e-4).\n"
            "      --eps_infeas_detect <tolerance> Infeasibility detection tolerance (default: 1e-10).\n"
            "      --l_inf_ruiz_iter <int>         Iterations for L-inf Ruiz rescaling (default: 10).\n"
            "      --no_pock_chambolle             Disable Pock-Chambolle rescaling (default: enabled).\n"
            "      --pock_chambolle_alpha <float>  Value for Pock-Chambolle alpha (default: 1.0).\n"
            "      --no_bound_obj_rescaling        Disable bound objective rescaling (default: enabled).\n"
            "      --no_al_qp                      Use the non-augmented baseline branch.\n"
            "      --al_sigma <float>              AL penalty parameter of the primal subproblem (default: 0).\n"
            "      --sigma_update <rule>           fixed, residual, kkt, displacement, progress, greedy, guarded,\n"
            "                                      or balanced_guarded (default: guarded).\n"
            "      --sigma_update_gain <float>     Log-domain damping gain (default: 0.5).\n"
            "      --sigma_clip_low <float>        Lower per-restart multiplicative clip (default: 0.6).\n"
            "      --sigma_clip_high <float>       Upper per-restart multiplicative clip (default: 1.7).\n"
            "      --sigma_min <float>             Minimum positive sigma (default: 1e-8).\n"
            "      --sigma_max <float>             Maximum sigma (default: 10).\n"
            "      --sigma_rollback_tol <float>    Rollback factor (default: 1.25; nonpositive disables).\n"
            "      --sigma_zero_threshold <float>  Guarded sigma below this is zero (default: 1e-8).\n"
            "      --sigma_activation_threshold <float>     Primal bottleneck ratio (default: 2).\n"
            "      --sigma_normal_motion_threshold <float>  Minimum ||A dx||/||dx|| (default: 1e-10).\n"
            "      --sigma_activation_value <float>         Sigma on guarded activation (default: 0.1).\n"
            "      --eval_freq <int>               Termination evaluation frequency (default: 200).\n"
            "      --trajectory_csv <path>         Write iteration trajectory CSV for plotting.\n"
            "      --sv_max_iter <int>             Max iterations for singular value estimation (default: 5000).\n"
            "      --sv_tol <float>                Tolerance for singular value estimation (default: 1e-4).\n"
            "      --opt_norm <norm_type>          l2 or linf (default: linf).\n");
}

}  // namespace

int main(int argc, char *argv[])
{
    fa_cp_parameters_t params;
    set_default_parameters(&params);
    static const struct option opts[] = {
        {"help", no_argument, nullptr, 'h'},
        {"verbose", required_argument, nullptr, 'v'},
        {"time_limit", required_argument, nullptr, 1001},
        {"iter_limit", required_argument, nullptr, 1002},
        {"eps_opt", required_argument, nullptr, 1003},
        {"eps_feas", required_argument, nullptr, 1004},
        {"eps_infeas_detect", required_argument, nullptr, 1005},
        {"eps_feas_polish", required_argument, nullptr, 1006},
        {"feasibility_polishing", no_argument, nullptr, 'f'},
        {"l_inf_ruiz_iter", required_argument, nullptr, 1007},
        {"pock_chambolle_alpha", required_argument, nullptr, 1008},
        {"no_pock_chambolle", no_argument, nullptr, 1009},
        {"no_bound_obj_rescaling", no_argument, nullptr, 1010},
        {"sv_max_iter", required_argument, nullptr, 1011},
        {"sv_tol", required_argument, nullptr, 1012},
        {"eval_freq", required_argument, nullptr, 1013},
        {"opt_norm", required_argument, nullptr, 1014},
        {"no_al_qp", no_argument, nullptr, 1015},
        {"al_sigma", required_argument, nullptr, 1016},
        {"trajectory_csv", required_argument, nullptr, 1017},
        {"sigma_update", required_argument, nullptr, 1018},
        {"sigma_update_gain", required_argument, nullptr, 1019},
        {"sigma_clip_low", required_argument, nullptr, 1020},
        {"sigma_clip_high", required_argument, nullptr, 1021},
        {"sigma_min", required_argument, nullptr, 1022},
        {"sigma_max", required_argument, nullptr, 1023},
        {"sigma_rollback_tol", required_argument, nullptr, 1024},
        {"sigma_zero_threshold", required_argument, nullptr, 1025},
        {"sigma_activation_threshold", required_argument, nullptr, 1026},
        {"sigma_normal_motion_threshold", required_argument, nullptr, 1027},
        {"sigma_activation_value", required_argument, nullptr, 1028},
        {nullptr, 0, nullptr, 0}};
    int opt;
    while ((opt = getopt_long(argc, argv, "hv:fp", opts, nullptr)) != -1) {
        switch (opt) {
        case 'h': usage(argv[0]); return 0;
        case 'v': params.verbose = atoi(optarg); break;
        case 1001: params.termination_criteria.time_sec_limit = atof(optarg); break;
        case 1002: params.termination_criteria.iteration_limit = atoi(optarg); break;
        case 1003: params.termination_criteria.eps_optimal_relative = atof(optarg); break;
        case 1004: params.termination_criteria.eps_feasible_relative = atof(optarg); break;
        case 1005: params.termination_criteria.eps_infeasible = atof(optarg); break;
        case 1006: params.termination_criteria.eps_feas_polish_relative = atof(optarg); break;
        case 'f': params.feasibility_polishing = true; break;
        case 1007: params.l_inf_ruiz_iterations = atoi(optarg); break;
        case 1008: params.pock_chambolle_alpha = atof(optarg); break;
        case 1009: params.has_pock_chambolle_alpha = false; break;
        case 1010: params.bound_objective_rescaling = false; break;
        case 1011: params.sv_max_iter = atoi(optarg); break;
        case 1012: params.sv_tol = atof(optarg); break;
        case 1013: params.termination_evaluation_frequency = atoi(optarg); break;
        case 1014:
            if (strcmp(optarg, "l2") == 0) params.optimality_norm = NORM_TYPE_L2;
            else if (strcmp(optarg, "linf") == 0) params.optimality_norm = NORM_TYPE_L_INF;
            else {
                fprintf(stderr, "Error: opt_norm must be 'l2' or 'linf'\n");
                return 1;
            }
            break;
        case 1015: params.use_al_qp = false; break;
        case 1016: params.al_sigma = atof(optarg); break;
        case 1017:
            params.trajectory_csv = optarg;
            break;
        case 1018:
            if (!parse_rule(optarg, &params.sigma_update_rule)) {
                fprintf(stderr, "Error: sigma_update must be fixed, residual, kkt, displacement, progress, greedy, "
                                "guarded, or balanced_guarded\n");
                return 1;
            }
            break;
        case 1019: params.sigma_update_gain = atof(optarg); break;
        case 1020: params.sigma_clip_low = atof(optarg); break;
        case 1021: params.sigma_clip_high = atof(optarg); break;
        case 1022: params.sigma_min = atof(optarg); break;
        case 1023: params.sigma_max = atof(optarg); break;
        case 1024: params.sigma_rollback_tol = atof(optarg); break;
        case 1025: params.sigma_zero_threshold = atof(optarg); break;
        case 1026: params.sigma_activation_threshold = atof(optarg); break;
        case 1027: params.sigma_normal_motion_threshold = atof(optarg); break;
        case 1028: params.sigma_activation_value = atof(optarg); break;
        case '?': return 1;
        default: break;
        }
    }
    if (argc - optind != 2) {
        fprintf(stderr, "Error: You must specify an input file and an output directory.\n\n");
        usage(argv[0]);
        return 1;
    }
    const char *file = argv[optind];
    const char *outdir = argv[optind + 1];
    const std::string name = pdal::qpcli::instance_name(file);
    qp_problem_t *problem = read_mps_file(file);
    if (!problem) {
        fprintf(stderr, "Failed to read or parse the file.\n");
        return 1;
    }
    fa_cp_result_t *r = optimize(&params, problem);
    if (!r) {
        fprintf(stderr, "Solver failed.\n");
    } else {
        pdal::qpcli::save_summary(r, pdal_fa_cp_termination_reason_to_string(r->termination_reason), outdir, name,
                                  [r](FILE *f) {
                                      fprintf(f, "Final AL Sigma: %e\n", r->final_al_sigma);
                                      fprintf(f, "Sigma Update Rule: %s\n", rule_name(r->sigma_update_rule));
                                      fprintf(f, "Sigma Update Count: %d\n", r->sigma_update_count);
                                  });
        pdal::qpcli::save_solution(r->primal_solution, problem->num_variables, outdir, name, "_primal_solution.txt");
        pdal::qpcli::save_solution(r->dual_solution, problem->num_constraints, outdir, name, "_dual_solution.txt");
        fa_cp_result_free(r);
    }
    qp_problem_free(problem);
    return 0;
}
