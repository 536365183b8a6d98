// MPS / MPS.gz reader of libcp_al.so (host, untimed).
//
// Conventions follow LP/CP_AL/src/mps_parser.c:402-900:
//  * a section header is a line with exactly one token starting with a letter
//    (ROWS, COLUMNS, RHS, RANGES, BOUNDS, OBJSENSE, ENDATA); anything else in
//    no section (NAME lines) is ignored;
//  * the first N row is the objective (else the first row); E/L/G rows become
//    constraints in file order, with [0,0] / (-inf,0] / [0,inf) until RHS;
//  * COLUMNS lines with an odd token count start a new column; 'MARKER' lines
//    are skipped; entries on unknown rows are dropped; objective entries add;
//  * RHS on the objective row sets objective_constant = -value;
//  * RANGES: G -> u = rhs+|R|, L -> l = rhs-|R|, E -> R>=0 ? u = rhs+R : l = rhs+R;
//  * BOUNDS LO UP FX FR MI PL BV (UP never touches the lower bound); default
//    bounds [0, inf);
//  * OBJSENSE MAX / MAXIMIZE negates the objective and its constant.
// Files whose name ends in ".gz" are read through zlib.
#include "mps_parser.h"

#include <zlib.h>

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#define PDAL_API extern "C" __attribute__((visibility("default")))

namespace {

class LineReader {
  public:
    explicit LineReader(const char *path)
    {
        const size_t len = strlen(path);
        gz_ = len > 3 && strcmp(path + len - 3, ".gz") == 0;
        if (gz_) g_ = gzopen(path, "rb");
        else f_ = fopen(path, "r");
    }
    ~LineReader()
    {
        if (g_) gzclose(g_);
        if (f_) fclose(f_);
    }
    bool ok() const { return gz_ ? g_ != nullptr : f_ != nullptr; }
    bool next(std::string &line)
    {
        line.clear();
        char buf[4096];
        for (;;) {
            const char *r = gz_ ? gzgets(g_, buf, sizeof buf) : fgets(buf, sizeof buf, f_);
            if (!r) return !line.empty();
            line += buf;
            if (!line.empty() && line.back() == '\n') return true;
        }
    }

  private:
    bool gz_ = false;
    gzFile g_ = nullptr;
    FILE *f_ = nullptr;
};

enum class Sec { None, Rows, Columns, Rhs, Ranges, Bounds, ObjSense, End };

}  // namespace

PDAL_API lp_problem_t *read_mps_file(const char *filename)
{
    if (!filename) return nullptr;
    LineReader rd(filename);
    if (!rd.ok()) {
        fprintf(stderr, "ERROR: Could not open file %s\n", filename);
        return nullptr;
    }
    struct PendingRow {
        char type;
        std::string name;
    };
    std::vector<PendingRow> pending;
    std::unordered_map<std::string, int> row_of, col_of;
    std::vector<char> row_type;
    std::vector<double> lo, hi, obj, vlo, vhi;
    std::vector<int> ti, tj;
    std::vector<double> tv;
    std::string obj_row, cur_col;
    bool have_obj = false, rows_done = false, maximize = false;
    double obj_const = 0.0;
    Sec sec = Sec::None;

    auto finalize_rows = [&]() {
        int obj_idx = -1;
        for (size_t i = 0; i < pending.size(); ++i)
            if (pending[i].type == 'N') {
                obj_idx = (int)i;
                break;
            }
        if (obj_idx == -1 && !pending.empty()) obj_idx = 0;
        if (obj_idx != -1) {
            obj_row = pending[obj_idx].name;
            have_obj = true;
        }
        for (size_t i = 0; i < pending.size(); ++i) {
            if ((int)i == obj_idx) continue;
            const char t = pending[i].type;
            if (t != 'E' && t != 'L' && t != 'G') continue;
            if (row_of.count(pending[i].name)) continue;
            row_of.emplace(pending[i].name, (int)row_type.size());
            row_type.push_back(t);
            lo.push_back(t == 'L' ? -INFINITY : 0.0);
            hi.push_back(t == 'G' ? INFINITY : 0.0);
        }
        rows_done = true;
    };

    std::string line;
    std::vector<std::string> tok;
    while (rd.next(line)) {
        if (line.empty() || line[0] == '*' || line[0] == '\n' || line[0] == '\r') continue;
        tok.clear();
        size_t p = 0;
        while (p < line.size() && tok.size() < 6) {
            while (p < line.size() && strchr(" \t\n\r", line[p])) ++p;
            if (p >= line.size()) break;
            size_t q = p;
            while (q < line.size() && !strchr(" \t\n\r", line[q])) ++q;
            tok.emplace_back(line.substr(p, q - p));
            p = q;
        }
        if (tok.empty()) continue;
        if (tok.size() == 1 && isalpha((unsigned char)tok[0][0])) {
            Sec next = Sec::None;
            const std::string &h = tok[0];
            if (h == "ROWS") next = Sec::Rows;
            else if (h == "COLUMNS") next = Sec::Columns;
            else if (h == "RHS") next = Sec::Rhs;
            else if (h == "RANGES") next = Sec::Ranges;
            else if (h == "BOUNDS") next = Sec::Bounds;
            else if (h == "OBJSENSE") next = Sec::ObjSense;
            else if (h == "ENDATA") next = Sec::End;
            if (sec == Sec::Rows && next != Sec::Rows && !rows_done) finalize_rows();
            sec = next;
            if (sec == Sec::End) break;
            continue;
        }
        const int nt = (int)tok.size();
        switch (sec) {
        case Sec::ObjSense:
            if (tok[0] == "MAX" || tok[0] == "MAXIMIZE") maximize = true;
            break;
        case Sec::Rows:
            if (nt >= 2) pending.push_back({tok[0][0], tok[1]});
            break;
        case Sec::Columns: {
            if (nt < 2) break;
            if (tok[1] == "'MARKER'") break;
            int start;
            if (nt % 2 != 0) {
                cur_col = tok[0];
                start = 1;
            } else {
                if (cur_col.empty()) {
                    fprintf(stderr, "ERROR: Column data found before any column name was defined.\n");
                    return nullptr;
                }
                start = 0;
            }
            auto it = col_of.find(cur_col);
            int j;
            if (it == col_of.end()) {
                j = (int)obj.size();
                col_of.emplace(cur_col, j);
                obj.push_back(0.0);
                vlo.push_back(0.0);
                vhi.push_back(INFINITY);
            } else {
                j = it->second;
            }
            for (int i = start; i + 1 < nt; i += 2) {
                const double v = atof(tok[i + 1].c_str());
                if (have_obj && tok[i] == obj_row) {
                    obj[j] += v;
                } else {
                    auto r = row_of.find(tok[i]);
                    if (r != row_of.end()) {
                        ti.push_back(r->second);
                        tj.push_back(j);
                        tv.push_back(v);
                    }
                }
            }
            break;
        }
        case Sec::Rhs:
            for (int i = 1; i + 1 < nt; i += 2) {
                const double v = atof(tok[i + 1].c_str());
                if (have_obj && tok[i] == obj_row) {
                    obj_const = -v;
                    continue;
                }
                auto r = row_of.find(tok[i]);
                if (r == row_of.end()) continue;
                const int k = r->second;
                if (row_type[k] == 'L') hi[k] = v;
                else if (row_type[k] == 'G') lo[k] = v;
                else lo[k] = hi[k] = v;
            }
            break;
        case Sec::Ranges:
            for (int i = 1; i + 1 < nt; i += 2) {
                const double R = atof(tok[i + 1].c_str());
                auto r = row_of.find(tok[i]);
                if (r == row_of.end()) continue;
                const int k = r->second;
                const char t = row_type[k];
                const double rhs = (t == 'L') ? hi[k] : lo[k];
                if (t == 'G') hi[k] = rhs + std::fabs(R);
                else if (t == 'L') lo[k] = rhs - std::fabs(R);
                else if (R >= 0) hi[k] = rhs + R;
                else lo[k] = rhs + R;
            }
            break;
        case Sec::Bounds: {
            if (nt < 3) break;
            auto c = col_of.find(tok[2]);
            if (c == col_of.end()) break;
            const int j = c->second;
            const double v = nt > 3 ? atof(tok[3].c_str()) : 0.0;
            const std::string &bt = tok[0];
            if (bt == "LO") vlo[j] = v;
            else if (bt == "UP") vhi[j] = v;
            else if (bt == "FX") vlo[j] = vhi[j] = v;
            else if (bt == "FR") {
                vlo[j] = -INFINITY;
                vhi[j] = INFINITY;
            } else if (bt == "MI") vlo[j] = -INFINITY;
            else if (bt == "PL") vhi[j] = INFINITY;
            else if (bt == "BV") {
                vlo[j] = 0.0;
                vhi[j] = 1.0;
            }
            break;
        }
        default: break;
        }
    }
    if (!rows_done) finalize_rows();

    const int m = (int)row_type.size(), n = (int)obj.size(), nnz = (int)tv.size();
    lp_problem_t *pr = (lp_problem_t *)calloc(1, sizeof(lp_problem_t));
    pr->num_variables = n;
    pr->num_constraints = m;
    pr->constraint_matrix_num_nonzeros = nnz;
    pr->objective_constant = maximize ? -obj_const : obj_const;
    auto dup = [](const std::vector<double> &v) {
        double *p = (double *)malloc(std::max<size_t>(v.size(), 1) * sizeof(double));
        if (!v.empty()) memcpy(p, v.data(), v.size() * sizeof(double));
        return p;
    };
    if (maximize)
        for (double &c : obj) c = -c;
    pr->objective_vector = dup(obj);
    pr->variable_lower_bound = dup(vlo);
    pr->variable_upper_bound = dup(vhi);
    pr->constraint_lower_bound = dup(lo);
    pr->constraint_upper_bound = dup(hi);
    int *rp = (int *)calloc((size_t)m + 1, sizeof(int));
    int *ci = (int *)malloc((size_t)std::max(nnz, 1) * sizeof(int));
    double *vv = (double *)malloc((size_t)std::max(nnz, 1) * sizeof(double));
    for (int k = 0; k < nnz; ++k) rp[ti[k] + 1]++;
    for (int i = 0; i < m; ++i) rp[i + 1] += rp[i];
    std::vector<int> pos(rp, rp + m + 1);
    for (int k = 0; k < nnz; ++k) {
        const int d = pos[ti[k]]++;
        ci[d] = tj[k];
        vv[d] = tv[k];
    }
    pr->constraint_matrix_row_pointers = rp;
    pr->constraint_matrix_col_indices = ci;
    pr->constraint_matrix_values = vv;
    return pr;
}
