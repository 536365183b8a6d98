// The runners of the three QP variants, built from this file with a variant flag:
//   -DPDAL_LIN_CP_AL     -> rhr_lin_cp_al, linked against librhr_lin_cp_al.so
//                           (QP/LIN_CP_AL/src/cli.c; CMakeLists.txt:143-158)
//   -DPDAL_PLANTED_FA_CP -> rhr_fa_cp, linked against librhr_fa_cp.so
//                           (QP/PLANTED_FA_CP/src/cli.c; CMakeLists.txt:143-158)
//   -DPDAL_PLANTED_CP_AL -> rhr_cp_al, linked against librhr_cp_al.so
//                           (QP/PLANTED_CP_AL/src/cli.c; CMakeLists.txt:143-158)
//
// All take FA_CP's common flags (getopt, defaults from set_default_parameters),
// accept only equality-constrained problems (validate_eqbox_problem: every row
// finite with |l - u| <= 1e-9 (1 + max(|l|, |u|))), and write FA_CP's output
// files without its three sigma lines.  Flags of their own: the linearized
// solver's --sigma_init (and its deprecated alias --al_lambda) and
// --reflection_coefficient; the planted CP-AL's --sigma_init,
// --reflection_coefficient and --trajectory_csv; the planted FA_CP's --al_sigma, the adaptive AL
// sigma flags, --reflection_coefficient, the restart and primal-weight PID
// flags and --trajectory_csv.  As in the reference, whose main discards run_*'s return value,
// the process exits with status 0 after an error too.
#if defined(PDAL_LIN_CP_AL)
#include "lin_cp_al.h"
#include "lin_cp_al_mps_parser.h"
using params_t = cp_al_parameters_t;
#define VARIANT_NAME "rhr_lin_cp_al"
#define VARIANT_REASON pdal_lin_cp_al_termination_reason_to_string
#define VARIANT_RESULT_FREE cp_al_result_free
#elif defined(PDAL_PLANTED_CP_AL)
#include "planted_cp_al.h"
#include "planted_cp_al_mps_parser.h"
using params_t = cp_al_parameters_t;
#define VARIANT_NAME "rhr_cp_al"
#define VARIANT_REASON pdal_planted_cp_al_termination_reason_to_string
#define VARIANT_RESULT_FREE cp_al_result_free
#elif defined(PDAL_PLANTED_FA_CP)
#include "planted_fa_cp.h"
#include "planted_fa_cp_mps_parser.h"
using params_t = fa_cp_parameters_t;
#define VARIANT_NAME "rhr_fa_cp"
#define VARIANT_REASON pdal_planted_fa_cp_termination_reason_to_string
#define VARIANT_RESULT_FREE fa_cp_result_free
#else
#error "build with -DPDAL_LIN_CP_AL, -DPDAL_PLANTED_CP_AL or -DPDAL_PLANTED_FA_CP"
#endif
#include "cli_io.hpp"

#include <getopt.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

namespace {

bool is_equality_problem(const qp_problem_t *p)
{   // validate_eqbox_problem (QP/LIN_CP_AL/src/cli.c:59-80, QP/PLANTED_CP_AL/src/cli.c:59-80,
    // QP/PLANTED_FA_CP/src/cli.c:95-116)
    const double tol = 1e-9;
    for (int i = 0; i < p->num_constraints; ++i) {
        const double lo = p->constraint_lower_bound[i], hi = p->constraint_upper_bound[i];
        if (!std::isfinite(lo) || !std::isfinite(hi)) {
            fprintf(stderr, "Error: EQBOX build requires finite equality rows; row %d is [%g, %g].\n", i, lo, hi);
            return false;
        }
        if (std::fabs(lo - hi) > tol * (1.0 + std::fmax(std::fabs(lo), std::fabs(hi)))) {
            fprintf(stderr, "Error: EQBOX build requires Ax=b; row %d has interval [%g, %g].\n", i, lo, hi);
            return false;
        }
    }
    return true;
}

#ifdef PDAL_PLANTED_FA_CP
struct Spelling {
    const char *name;
    int value;
};
// QP/PLANTED_FA_CP/src/cli.c:32-67
const Spelling kModes[] = {{"up", ADAPTIVE_AL_SIGMA_MODE_UP_ONLY},
                           {"up_only", ADAPTIVE_AL_SIGMA_MODE_UP_ONLY},
                           {"down", ADAPTIVE_AL_SIGMA_MODE_DOWN_ONLY},
                           {"down_only", ADAPTIVE_AL_SIGMA_MODE_DOWN_ONLY},
                           {"bidir", ADAPTIVE_AL_SIGMA_MODE_BIDIRECTIONAL},
                           {"bidirectional", ADAPTIVE_AL_SIGMA_MODE_BIDIRECTIONAL}};
const Spelling kRules[] = {{"trigger", ADAPTIVE_AL_SIGMA_RULE_TRIGGER},
                           {"residual_trigger", ADAPTIVE_AL_SIGMA_RULE_TRIGGER},
                           {"sqrt", ADAPTIVE_AL_SIGMA_RULE_SQRT},
                           {"sqrt_balance", ADAPTIVE_AL_SIGMA_RULE_SQRT},
                           {"bangbang", ADAPTIVE_AL_SIGMA_RULE_BANGBANG},
                           {"switch", ADAPTIVE_AL_SIGMA_RULE_BANGBANG}};

template <size_t N>
bool lookup(const Spelling (&table)[N], const char *s, int *out)
{
    for (const Spelling &e : table)
        if (strcmp(s, e.name) == 0) {
            *out = e.value;
            return true;
        }
    return false;
}
#endif

void usage(const char *prog)
{
    fprintf(stderr, "Usage: %s [OPTIONS] <mps_file> <output_dir>\n\n", prog);
    fprintf(stderr,
            "Arguments:\n"
            "  <mps_file>               Path to the input problem in MPS format (.mps .QPS or .mps.gz).\n"
            "  <output_dir>             Directory for <basename>_summary.txt, <basename>_primal_solution.txt,\n"
            "                           <basename>_dual_solution.txt.\n\n"
            "  This EQBOX build only accepts problems with Ax = b.\n\n"
            "Options:\n"
            "  -h, --help                          Display this help message.\n"
            "  -v, --verbose <int>                 Verbosity level: 0=silent, 1=summary, 2=detailed (default: 1).\n"
            "      --time_limit <seconds>          Time limit in seconds (default: 3600.0).\n");
    fprintf(stderr, "      --iter_limit <iterations>       Iteration limit (default: %d).\n", INT32_MAX);
    fprintf(stderr,
            "      --eps_opt <tolerance>           Relative optimality tolerance (default: 1e-4).\n"
            "      --eps_feas <tolerance>          Relative feasibility tolerance (default: 1e-4).\n"
            "      --eps_infeas_detect <tolerance> Infeasibility detection tolerance (default: 1e-10).\n"
            "  -f, --feasibility_polishing         Polish primal and dual feasibility after the solve.\n"
            "      --eps_feas_polish <tolerance>   Relative polishing tolerance (default: 1e-6).\n"
            "      --l_inf_ruiz_iter <int>         Iterations for L-inf Ruiz rescaling (default: 10).\n"
            "      --no_pock_chambolle             Disable Pock-Chambolle rescaling (default: enabled).\n"
            "      --pock_chambolle_alpha <float>  Value for Pock-Chambolle alpha (default: 1.0).\n"
            "      --no_bound_obj_rescaling        Disable bound objective rescaling (default: enabled).\n"
#if defined(PDAL_LIN_CP_AL)
            "      --no_al_qp                      Use the non-augmented baseline branch.\n"
            "      --sigma_init <float>            Initial sigma of the linearized CP-AL (0 = automatic, default).\n"
            "      --al_lambda <float>             Deprecated alias for --sigma_init.\n"
            "      --reflection_coefficient <float> Reflection parameter of the Halpern step (default: 0.0).\n"
#elif defined(PDAL_PLANTED_CP_AL)
            "      --no_al_qp                      Use the non-augmented baseline branch.\n"
            "      --sigma_init <float>            Initial sigma of the single-parameter CP-AL (0 = automatic, default).\n"
            "      --reflection_coefficient <float> Reflection coefficient of the Halpern step (default: 1.0).\n"
            "      --trajectory_csv <path>         Write iteration trajectory CSV for plotting.\n"
#else
            "      --no_al_qp                      Use the non-augmented baseline branch.\n"
            "      --al_sigma <float>              AL penalty parameter of the primal subproblem (default: 0.1).\n"
            "      --adaptive_al_sigma             Restart-only adaptive updates of al_sigma.\n"
            "      --adaptive_al_sigma_mode <m>    up, down or bidir (default: bidir).\n"
            "      --adaptive_al_sigma_rule <r>    trigger, sqrt or bangbang (default: trigger).\n"
            "      --al_sigma_min <float>          Lower bound of adaptive al_sigma (default: 5e-3).\n"
            "      --al_sigma_max <float>          Upper bound of adaptive al_sigma (default: 3e-1).\n"
            "      --al_sigma_up <float>           Multiplicative increase at restart (default: 1.5).\n"
            "      --al_sigma_down <float>         Multiplicative decrease at restart (default: 0.7).\n"
            "      --al_sigma_primal_trigger <float> Increase when rel_primal dominates by this ratio (default: 3.0).\n"
            "      --al_sigma_gap_trigger <float>  Decrease when rel_gap dominates by this ratio (default: 5.0).\n"
            "      --reflection_coefficient <float> Reflection coefficient of the Halpern step (default: 1.0).\n"
            "      --trajectory_csv <path>         Write iteration trajectory CSV for plotting.\n"
            "      --restart_artificial <float>    Artificial restart threshold (default: 0.36).\n"
            "      --restart_sufficient <float>    Sufficient reduction threshold (default: 0.2).\n"
            "      --restart_necessary <float>     Necessary reduction threshold (default: 0.8).\n"
            "      --restart_kp <float>            Primal-weight PID proportional gain (default: 0.99).\n"
            "      --restart_ki <float>            Primal-weight PID integral gain (default: 0.01).\n"
            "      --restart_kd <float>            Primal-weight PID derivative gain (default: 0.0).\n"
            "      --restart_i_smooth <float>      Integral smoothing of the primal-weight PID (default: 0.3).\n"
#endif
            "      --eval_freq <int>               Termination evaluation frequency (default: 200).\n"
            "      --sv_max_iter <int>             Max iterations for singular value estimation (default: 5000).\n"
            "      --sv_tol <float>                Tolerance for singular value estimation (default: 1e-4).\n"
            "      --opt_norm <norm_type>          l2 or linf (default: linf).\n");
}

int run(int argc, char *argv[])
{
    params_t params;
    set_default_parameters(&params);
    // option codes as in the reference runners (1001-1015 shared with FA_CP's)
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
#if defined(PDAL_LIN_CP_AL)
        {"sigma_init", required_argument, nullptr, 1016},
        {"al_lambda", required_argument, nullptr, 1017},
        {"reflection_coefficient", required_argument, nullptr, 1018},
#elif defined(PDAL_PLANTED_CP_AL)
        {"sigma_init", required_argument, nullptr, 1016},
        {"trajectory_csv", required_argument, nullptr, 1017},
        {"reflection_coefficient", required_argument, nullptr, 1018},
#else
        {"al_sigma", required_argument, nullptr, 1016},
        {"reflection_coefficient", required_argument, nullptr, 1017},
        {"adaptive_al_sigma", no_argument, nullptr, 1018},
        {"al_sigma_min", required_argument, nullptr, 1019},
        {"al_sigma_max", required_argument, nullptr, 1020},
        {"al_sigma_up", required_argument, nullptr, 1021},
        {"al_sigma_down", required_argument, nullptr, 1022},
        {"al_sigma_primal_trigger", required_argument, nullptr, 1023},
        {"al_sigma_gap_trigger", required_argument, nullptr, 1024},
        {"restart_artificial", required_argument, nullptr, 1025},
        {"restart_sufficient", required_argument, nullptr, 1026},
        {"restart_necessary", required_argument, nullptr, 1027},
        {"restart_kp", required_argument, nullptr, 1028},
        {"restart_ki", required_argument, nullptr, 1029},
        {"restart_kd", required_argument, nullptr, 1030},
        {"restart_i_smooth", required_argument, nullptr, 1031},
        {"adaptive_al_sigma_mode", required_argument, nullptr, 1032},
        {"adaptive_al_sigma_rule", required_argument, nullptr, 1033},
        {"trajectory_csv", required_argument, nullptr, 1034},
#endif
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
#if defined(PDAL_LIN_CP_AL)
        case 1016:
        case 1017: params.sigma_init = atof(optarg); break;
        case 1018: params.reflection_coefficient = atof(optarg); break;
#elif defined(PDAL_PLANTED_CP_AL)
        case 1016: params.sigma_init = atof(optarg); break;
        case 1017:
            params.trajectory_csv = optarg;
            break;
        case 1018: params.reflection_coefficient = atof(optarg); break;
#else
        case 1016: params.al_sigma = atof(optarg); break;
        case 1017: params.reflection_coefficient = atof(optarg); break;
        case 1018: params.adaptive_al_sigma = true; break;
        case 1019: params.al_sigma_min = atof(optarg); break;
        case 1020: params.al_sigma_max = atof(optarg); break;
        case 1021: params.al_sigma_increase_factor = atof(optarg); break;
        case 1022: params.al_sigma_decrease_factor = atof(optarg); break;
        case 1023: params.al_sigma_primal_trigger = atof(optarg); break;
        case 1024: params.al_sigma_gap_trigger = atof(optarg); break;
        case 1025: params.restart_params.artificial_restart_threshold = atof(optarg); break;
        case 1026: params.restart_params.sufficient_reduction_for_restart = atof(optarg); break;
        case 1027: params.restart_params.necessary_reduction_for_restart = atof(optarg); break;
        case 1028: params.restart_params.k_p = atof(optarg); break;
        case 1029: params.restart_params.k_i = atof(optarg); break;
        case 1030: params.restart_params.k_d = atof(optarg); break;
        case 1031: params.restart_params.i_smooth = atof(optarg); break;
        case 1032: {
            int v;
            if (!lookup(kModes, optarg, &v)) {
                fprintf(stderr, "Error: adaptive_al_sigma_mode must be one of 'up', 'down', or 'bidir'\n");
                return 1;
            }
            params.adaptive_al_sigma_mode = (adaptive_al_sigma_mode_t)v;
        } break;
        case 1033: {
            int v;
            if (!lookup(kRules, optarg, &v)) {
                fprintf(stderr, "Error: adaptive_al_sigma_rule must be one of 'trigger', 'sqrt', or 'bangbang'\n");
                return 1;
            }
            params.adaptive_al_sigma_rule = (adaptive_al_sigma_rule_t)v;
        } break;
        case 1034:
            params.trajectory_csv = optarg;
            break;
#endif
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
    if (!is_equality_problem(problem)) {
        qp_problem_free(problem);
        return 1;
    }
    auto *r = optimize(&params, problem);
    if (!r) {
        fprintf(stderr, "Solver failed.\n");
    } else {
        pdal::qpcli::save_summary(r, VARIANT_REASON(r->termination_reason), outdir, name, [](FILE *) {});
        pdal::qpcli::save_solution(r->primal_solution, problem->num_variables, outdir, name, "_primal_solution.txt");
        pdal::qpcli::save_solution(r->dual_solution, problem->num_constraints, outdir, name, "_dual_solution.txt");
        VARIANT_RESULT_FREE(r);
    }
    qp_problem_free(problem);
    return 0;
}

}  // namespace

int main(int argc, char *argv[])
{
    (void)run(argc, argv);
    return 0;
}
