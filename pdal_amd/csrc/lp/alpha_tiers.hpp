// Adaptive alpha tiers of the plain CP_AL path (the reference calls alpha
// "lambda"): at every alpha-mode restart, pick the next alpha from a small grid
// by how much the KKT residuals, the duality gap and the fixed-point error fell
// over the epoch that just ended.  Off unless CP_AL_ADAPTIVE_LAMBDA is non-zero
// or CP_AL_LAMBDA_PROFILE names a profile.
//
// Reference: LP/CP_AL/src/solver.cu:1059-1145 (grid parsing, sorting and
// de-duplication), :3166-3262 (set-up and profiles), :3264-3276 (tier ->
// alpha), :3278-3536 (update at a restart: the tier ladder, or the race over
// the tiers); the forced restart of the race in the main loop (:1590-1600).
//
// Host-only scalar logic; the Solver owns one instance and copies it into the
// positive-alpha trial's snapshot with the other scalars.
#pragma once

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "cp_al_internal.hpp"

namespace pdal {
namespace lp {

struct AlphaTiers {
    static constexpr int kMaxTiers = 16;  // CP_AL_MAX_LAMBDA_TIERS (internal_types.h:25)

    bool adaptive = false, ready = false, race = false;
    int num = 1, tier = 0, hold = 0, restarts = 0;
    double values[kMaxTiers] = {};
    int hold_epochs = 2, min_restarts = 0, min_iters = 0, force_restart_iters = 0, last_restart_iter = 0;
    // race bookkeeping
    int next = 1, best_tier = 0, exploit_left = 0, exploit_epochs = 3, cool_left = 0, cool_epochs = 0;
    int cycle_tier = 0;
    double best_score = -INFINITY, cycle_score = -INFINITY, bad_ratio = INFINITY;
    // plateau thresholds
    double min_time = 0.0, feas_plateau = 0.98, gap_plateau = 0.95, fp_plateau = 0.98;
    double feas_active = 1e-6, gap_active = 1e-6, fp_active = 1e-8, gap_dominance = 100.0;
    // the metrics at the start of the running epoch
    double feas0 = INFINITY, gap0 = INFINITY, fp0 = INFINITY, time0 = 0.0;

    double alpha_of(int t) const
    {   // lambda_from_tier, :3264-3276
        if (num <= 0) return 0.0;
        t = t < 0 ? 0 : t >= num ? num - 1 : t;
        return values[t];
    }

    // Sorted, de-duplicated grid with 0 always present (:1085-1145).
    void set_grid(const double *v, int count)
    {
        double g[kMaxTiers];
        int n = 0;
        if (count <= 0) {
            g[n++] = 0.0;
        } else {
            bool zero = false;
            for (int i = 0; i < count && n < kMaxTiers; ++i) {
                const double a = std::fmin(std::fmax(v[i], 0.0), 0.95);
                zero = zero || a <= 1e-15;
                g[n++] = a;
            }
            if (!zero && n < kMaxTiers) g[n++] = 0.0;
        }
        for (int i = 1; i < n; ++i) {  // insertion sort, ascending
            const double key = g[i];
            int j = i - 1;
            for (; j >= 0 && g[j] > key; --j) g[j + 1] = g[j];
            g[j + 1] = key;
        }
        num = 0;
        for (int i = 0; i < n; ++i)
            if (num == 0 || std::fabs(g[i] - values[num - 1]) > 1e-15) values[num++] = g[i];
        if (num == 0) values[num++] = 0.0;
        for (int i = num; i < kMaxTiers; ++i) values[i] = values[num - 1];
    }

    // Tokens separated by , ; : space or tab; each clipped to [0, 0.95] (:1059-1083).
    static int parse_grid(const char *raw, double *out, int max_out)
    {
        if (!raw || !raw[0] || max_out <= 0) return 0;
        char buf[512];
        strncpy(buf, raw, sizeof(buf) - 1);
        buf[sizeof(buf) - 1] = '\0';
        int n = 0;
        char *save = nullptr;
        for (char *tok = strtok_r(buf, ",;: \t", &save); tok && n < max_out; tok = strtok_r(nullptr, ",;: \t", &save)) {
            char *end = nullptr;
            const double v = strtod(tok, &end);
            if (end != tok && std::isfinite(v)) out[n++] = std::fmin(std::fmax(v, 0.0), 0.95);
        }
        return n;
    }

    // initialize_adaptive_lambda (:3166-3262).  Returns the starting alpha
    // when the controller is on, else `alpha` unchanged.
    double init(double alpha, double now)
    {
        *this = AlphaTiers{};
        const char *profile = getenv("CP_AL_LAMBDA_PROFILE");
        auto is = [&](const char *a, const char *b) {
            return profile && (strcmp(profile, a) == 0 || strcmp(profile, b) == 0);
        };
        const bool grid4 = is("grid4_race", "race4");
        const bool mp_race = is("mp_race", "safe_race");
        const bool mp_auto = is("mp_auto", "safe_auto");
        const bool small = grid4 || mp_race || mp_auto;
        adaptive = small || env_int("CP_AL_ADAPTIVE_LAMBDA", 0) != 0;

        double grid[kMaxTiers];
        int count = parse_grid(getenv("CP_AL_LAMBDA_GRID"), grid, kMaxTiers);
        if (count == 0 && (mp_race || mp_auto)) {
            static const double mp[] = {0.0, 0.005, 0.01, 0.02, 0.035, 0.05, 0.075, 0.10, 0.15, 0.20, 0.30, 0.40};
            count = (int)(sizeof(mp) / sizeof(mp[0]));
            for (int i = 0; i < count; ++i) grid[i] = mp[i];
        }
        if (count == 0) {
            grid[0] = env_double("CP_AL_LAMBDA_BASE", 0.0);
            grid[1] = env_double("CP_AL_LAMBDA_LOW", small ? 0.02 : 0.10);
            grid[2] = env_double("CP_AL_LAMBDA_MID", 0.10);
            grid[3] = env_double("CP_AL_LAMBDA_HIGH", small ? 0.20 : 0.10);
            count = 4;
        }
        set_grid(grid, count);

        hold_epochs = env_int("CP_AL_LAMBDA_HOLD_EPOCHS", mp_auto ? 1 : 2);
        min_restarts = env_int("CP_AL_LAMBDA_MIN_RESTARTS", (mp_race || mp_auto) ? 2 : grid4 ? 1 : 0);
        min_iters = env_int("CP_AL_LAMBDA_MIN_ITERS", (mp_race || mp_auto) ? 10000 : 0);
        force_restart_iters = env_int("CP_AL_LAMBDA_FORCE_RESTART_ITERS", mp_race ? 20000 : 0);
        race = grid4 || mp_race || env_int("CP_AL_LAMBDA_RACE", 0) != 0;
        exploit_epochs = env_int("CP_AL_LAMBDA_RACE_EXPLOIT_EPOCHS", mp_race ? 2 : grid4 ? 9999 : 3);
        cool_epochs = env_int("CP_AL_LAMBDA_RACE_COOLDOWN_EPOCHS", mp_race ? 3 : grid4 ? 1000000000 : 0);
        bad_ratio = env_double("CP_AL_LAMBDA_RACE_BAD_RATIO", mp_race ? 3.0 : grid4 ? 10.0 : INFINITY);
        min_time = env_double("CP_AL_LAMBDA_MIN_TIME", 0.0);
        feas_plateau = env_double("CP_AL_LAMBDA_FEAS_PLATEAU", 0.98);
        gap_plateau = env_double("CP_AL_LAMBDA_GAP_PLATEAU", 0.95);
        fp_plateau = env_double("CP_AL_LAMBDA_FP_PLATEAU", 0.98);
        feas_active = env_double("CP_AL_LAMBDA_FEAS_ACTIVE", 1e-6);
        gap_active = env_double("CP_AL_LAMBDA_GAP_ACTIVE", 1e-6);
        fp_active = env_double("CP_AL_LAMBDA_FP_ACTIVE", 1e-8);
        gap_dominance = env_double("CP_AL_LAMBDA_GAP_DOMINANCE", 100.0);

        hold_epochs = hold_epochs < 0 ? 0 : hold_epochs;
        min_restarts = min_restarts < 0 ? 0 : min_restarts;
        min_iters = min_iters < 0 ? 0 : min_iters;
        force_restart_iters = force_restart_iters < 0 ? 0 : force_restart_iters;
        exploit_epochs = exploit_epochs < 0 ? 0 : exploit_epochs;
        cool_epochs = cool_epochs < 0 ? 0 : cool_epochs;
        bad_ratio = std::fmax(bad_ratio, 1.0);
        min_time = std::fmax(min_time, 0.0);
        time0 = now;
        return adaptive ? values[0] : alpha;
    }

    // The race's artificial restart (:1590-1600).
    bool force_restart(int total) const
    {
        return adaptive && race && force_restart_iters > 0 && total > 0 &&
               total - last_restart_iter >= force_restart_iters;
    }

    // update_adaptive_lambda_on_restart (:3278-3536).  Returns the alpha of
    // the next epoch (`alpha` unchanged while the controller is off or only
    // records its first epoch).
    double on_restart(double alpha, double rel_p, double rel_d, double rel_gap, double fp, int total, double now,
                      bool verbose)
    {
        if (!adaptive) return alpha;
        const double eps = 1e-300;
        const double feas1 = std::fmax(rel_p, rel_d);
        const double gap1 = std::fmax(std::fabs(rel_gap), 0.0);
        const double fp1 = (std::isfinite(fp) && fp > 0.0) ? fp : INFINITY;
        restarts += 1;
        if (!ready) {
            feas0 = feas1;
            gap0 = gap1;
            fp0 = fp1;
            time0 = now;
            ready = true;
            if (verbose)
                printf("lambda-adapt init restart %d lambda %.3e feas %.3e gap %.3e fp %.3e\n", restarts, alpha,
                       feas1, gap1, fp1);
            return alpha;
        }
        const double feas_r = std::fmax(feas1, eps) / std::fmax(feas0, eps);
        const double gap_r = std::fmax(gap1, eps) / std::fmax(gap0, eps);
        const double fp_r = std::fmax(fp1, eps) / std::fmax(fp0, eps);
        const double dt = std::fmax(now - time0, 1e-9);
        const bool mature = restarts >= min_restarts && total >= min_iters && now >= min_time;
        const int old_tier = tier;
        const double old_alpha = alpha;

        if (race) {
            double score = 0.45 * std::log(std::fmax(feas0, eps) / std::fmax(feas1, eps)) +
                           0.35 * std::log(std::fmax(fp0, eps) / std::fmax(fp1, eps)) +
                           0.20 * std::log(std::fmax(gap0, eps) / std::fmax(gap1, eps));
            score /= dt;
            const bool finite = std::isfinite(score);
            const bool bad = old_tier > 0 && (feas_r > bad_ratio || gap_r > bad_ratio || fp_r > bad_ratio);
            if (!bad && finite && score > best_score) {
                best_score = score;
                best_tier = old_tier;
            }
            if (!bad && finite && score > cycle_score) {
                cycle_score = score;
                cycle_tier = old_tier;
            }
            const double worst = std::fmax(feas_r, std::fmax(gap_r, fp_r));
            const bool feas_lag = score > 5.0 && restarts <= 10 && feas_r > 1.20 && feas_r < 2.00 && gap_r < 1.20 &&
                                  fp_r < 1.20;
            const bool tradeoff = score > 5.0 && feas_r < 0.50 && std::fmax(gap_r, fp_r) > 3.00 && worst < bad_ratio;
            const bool compact = score < -10.0 && worst < 3.0 && feas_r < 1.10;
            const bool plateau = mature && ((std::fmax(feas0, feas1) > feas_active && feas_r > feas_plateau) ||
                                            (std::fmax(gap0, gap1) > gap_active && gap_r > gap_plateau) ||
                                            (std::fmax(fp0, fp1) > fp_active && fp_r > fp_plateau));
            const bool start = old_tier > 0 || plateau || feas_lag || tradeoff || compact;
            const double score_or_none = finite ? score : -INFINITY;
            int want = 0;
            if (bad) {
                next = 1;
                exploit_left = 0;
                cool_left = cool_epochs;
                cycle_tier = 0;
                cycle_score = -INFINITY;
            } else if (cool_left > 0) {
                cool_left -= 1;
                next = 1;
                exploit_left = 0;
                cycle_tier = 0;
                cycle_score = score_or_none;
            } else if (!start) {
                next = 1;
                exploit_left = 0;
                cycle_tier = 0;
                cycle_score = score_or_none;
            } else if (!mature) {
                cycle_tier = 0;
                cycle_score = score_or_none;
            } else if (next < num) {  // explore the next tier
                want = next;
                if (next == 1) {
                    cycle_tier = old_tier;
                    cycle_score = score_or_none;
                }
                next += 1;
                exploit_left = exploit_epochs;
            } else if (exploit_left > 0) {  // stay on the cycle's best tier
                want = cycle_tier;
                exploit_left -= 1;
            } else {  // start a new cycle from the best tier
                next = 1;
                want = cycle_tier;
                cycle_score = score_or_none;
            }
            tier = want;
            hold = 0;
            alpha = alpha_of(tier);
            feas0 = feas1;
            gap0 = gap1;
            fp0 = fp1;
            time0 = now;
            if (verbose || old_tier != tier || std::fabs(old_alpha - alpha) > 1e-15)
                printf("lambda-race restart %d score %.3e best(tier=%d score=%.3e) cycle(tier=%d score=%.3e) "
                       "tier %d->%d lambda %.3e->%.3e next=%d exploit=%d cool=%d bad=%d start=%d mature=%d "
                       "ratios(feas %.3e gap %.3e fp %.3e)\n",
                       restarts, score, best_tier, best_score, cycle_tier, cycle_score, old_tier, tier, old_alpha,
                       alpha, next, exploit_left, cool_left, (int)bad, (int)start, (int)mature, feas_r, gap_r, fp_r);
            return alpha;
        }

        const bool feas_pl = std::fmax(feas0, feas1) > feas_active && feas_r > feas_plateau;
        const bool gap_pl = std::fmax(gap0, gap1) > gap_active && gap_r > gap_plateau;
        const bool fp_pl = std::fmax(fp0, fp1) > fp_active && fp_r > fp_plateau;
        const bool gap_dom = gap1 > gap_dominance * std::fmax(feas1, eps);
        const int low = num > 1 ? 1 : 0;
        const int mid = num > 2 ? num / 2 : low;
        const int high = num > 0 ? num - 1 : 0;
        int want = 0;
        if (!mature) want = 0;
        else if (gap_pl && gap_dom) want = high;
        else if (feas_pl || (gap_pl && fp_pl)) want = mid;
        else if (gap_pl || fp_pl) want = low;
        if (want > 0) {
            tier = want;
            hold = hold_epochs;
        } else if (hold > 0 && tier > 0) {
            hold -= 1;
        } else {
            tier = 0;
            hold = 0;
        }
        alpha = alpha_of(tier);
        feas0 = feas1;
        gap0 = gap1;
        fp0 = fp1;
        time0 = now;
        if (verbose || old_tier != tier || std::fabs(old_alpha - alpha) > 1e-15)
            printf("lambda-adapt restart %d ratios(feas %.3e gap %.3e fp %.3e) "
                   "plateau(feas %d gap %d fp %d dom %d mature %d) tier %d->%d hold %d lambda %.3e->%.3e\n",
                   restarts, feas_r, gap_r, fp_r, (int)feas_pl, (int)gap_pl, (int)fp_pl, (int)gap_dom, (int)mature,
                   old_tier, tier, hold, old_alpha, alpha);
        return alpha;
    }
};

}  // namespace lp
}  // namespace pdal
