// Free/fixed MPS and QPS reader shared by libcp_al.so and libfa_cp.so (host,
// untimed).  Produces triplets and dense vectors; each library builds its own
// problem struct from them.
//
// Conventions (LP/CP_AL/src/mps_parser.c:402-900, QP/FA_CP/src/mps_parser.c:371-907):
//  * a section header is a line with exactly one token starting with a letter
//    (ROWS, COLUMNS, RHS, RANGES, BOUNDS, OBJSENSE, QUADOBJ, QMATRIX, ENDATA);
//    any other one-token line (NAME, MAX under OBJSENSE, ...) leaves every
//    section and its lines are ignored;
//  * the first N row is the objective (else the first row); E/L/G rows become
//    constraints in file order, with [0,0] / (-inf,0] / [0,inf) until RHS;
//  * COLUMNS lines with an odd token count start a new column; 'MARKER' lines
//    are skipped; entries on unknown rows are dropped; objective entries add;
//  * RHS on the objective row sets objective_constant = -value;
//  * RANGES: G -> u = rhs+|R|, L -> l = rhs-|R|, E -> R>=0 ? u = rhs+R : l = rhs+R;
//  * BOUNDS LO UP FX FR MI PL BV (UP never touches the lower bound); default
//    bounds [0, inf);
//  * OBJSENSE MAX / MAXIMIZE (on a line with more than one token) negates the
//    objective, the quadratic entries and the constant;
//  * QUADOBJ (lower triangle) adds (i,j) and (j,i) for i != j, (i,i) once;
//    QMATRIX (full matrix) adds every entry as given; zero values are skipped,
//    entries naming an unknown column are skipped with a warning (the LP
//    reader does not know these sections).
// Files whose name ends in ".gz" are read through zlib.
#pragma once

#include <zlib.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

namespace pdal {
namespace mps {

struct MpsData {
    int m = 0, n = 0;
    std::vector<double> obj, vlo, vhi;  // n
    std::vector<double> lo, hi;         // m
    std::vector<int> ai, aj;            // constraint triplets in file order
    std::vector<double> av;
    std::vector<int> qi, qj;            // quadratic-objective triplets in file order
    std::vector<double> qv;
    double obj_const = 0.0;             // sign already applied for MAX
    bool maximize = false;
};

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
    LineReader(const LineReader &) = delete;
    LineReader &operator=(const LineReader &) = delete;
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

enum class Sec { None, Rows, Columns, Rhs, Ranges, Bounds, ObjSense, QuadObj, QMatrix, End };

// Returns false (after printing the reason) when the file cannot be read or
// parsed.  quad = false reads as the LP reader does: QUADOBJ / QMATRIX are
// unknown headers and their lines are ignored.
inline bool read_mps(const char *filename, MpsData &d, bool quad)
{
    d = MpsData{};
    if (!filename) return false;
    LineReader rd(filename);
    if (!rd.ok()) {
        fprintf(stderr, "ERROR: Could not open file %s\n", filename);
        return false;
    }
    struct PendingRow {
        char type;
        std::string name;
    };
    std::vector<PendingRow> pending;
    std::unordered_map<std::string, int> row_of, col_of;
    std::vector<char> row_type;
    std::string obj_row, cur_col;
    bool have_obj = false, rows_done = false;
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
            d.lo.push_back(t == 'L' ? -INFINITY : 0.0);
            d.hi.push_back(t == 'G' ? INFINITY : 0.0);
        }
        rows_done = true;
    };
    auto quad_entry = [&](const std::vector<std::string> &tok, bool fill_sym) {
        if (tok.size() < 3) return;
        const double v = atof(tok[2].c_str());
        if (v == 0.0) return;
        auto a = col_of.find(tok[0]), b = col_of.find(tok[1]);
        if (a == col_of.end() || b == col_of.end()) {
            fprintf(stderr, "Warning: Variable '%s' or '%s' not found in COLUMNS. Skipping Q-entry.\n",
                    tok[0].c_str(), tok[1].c_str());
            return;
        }
        const int i = a->second, j = b->second;
        d.qi.push_back(i);
        d.qj.push_back(j);
        d.qv.push_back(v);
        if (i != j && fill_sym) {
            d.qi.push_back(j);
            d.qj.push_back(i);
            d.qv.push_back(v);
        }
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
            else if (quad && h == "QUADOBJ") next = Sec::QuadObj;
            else if (quad && h == "QMATRIX") next = Sec::QMatrix;
            else if (h == "ENDATA") next = Sec::End;
            if (sec == Sec::Rows && next != Sec::Rows && !rows_done) finalize_rows();
            sec = next;
            if (sec == Sec::End) break;
            continue;
        }
        const int nt = (int)tok.size();
        switch (sec) {
        case Sec::ObjSense:
            if (tok[0] == "MAX" || tok[0] == "MAXIMIZE") d.maximize = true;
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
                    fprintf(stderr, "ERROR: Failed to parse MPS file.\n");
                    return false;
                }
                start = 0;
            }
            auto it = col_of.find(cur_col);
            int j;
            if (it == col_of.end()) {
                j = (int)d.obj.size();
                col_of.emplace(cur_col, j);
                d.obj.push_back(0.0);
                d.vlo.push_back(0.0);
                d.vhi.push_back(INFINITY);
            } else {
                j = it->second;
            }
            for (int i = start; i + 1 < nt; i += 2) {
                const double v = atof(tok[i + 1].c_str());
                if (have_obj && tok[i] == obj_row) {
                    d.obj[j] += v;
                } else {
                    auto r = row_of.find(tok[i]);
                    if (r != row_of.end()) {
                        d.ai.push_back(r->second);
                        d.aj.push_back(j);
                        d.av.push_back(v);
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
                if (row_type[k] == 'L') d.hi[k] = v;
                else if (row_type[k] == 'G') d.lo[k] = v;
                else d.lo[k] = d.hi[k] = v;
            }
            break;
        case Sec::Ranges:
            for (int i = 1; i + 1 < nt; i += 2) {
                const double R = atof(tok[i + 1].c_str());
                auto r = row_of.find(tok[i]);
                if (r == row_of.end()) continue;
                const int k = r->second;
                const char t = row_type[k];
                const double rhs = (t == 'L') ? d.hi[k] : d.lo[k];
                if (t == 'G') d.hi[k] = rhs + std::fabs(R);
                else if (t == 'L') d.lo[k] = rhs - std::fabs(R);
                else if (R >= 0) d.hi[k] = rhs + R;
                else d.lo[k] = rhs + R;
            }
            break;
        case Sec::Bounds: {
            if (nt < 3) break;
            auto c = col_of.find(tok[2]);
            if (c == col_of.end()) break;
            const int j = c->second;
            const double v = nt > 3 ? atof(tok[3].c_str()) : 0.0;
            const std::string &bt = tok[0];
            if (bt == "LO") d.vlo[j] = v;
            else if (bt == "UP") d.vhi[j] = v;
            else if (bt == "FX") d.vlo[j] = d.vhi[j] = v;
            else if (bt == "FR") {
                d.vlo[j] = -INFINITY;
                d.vhi[j] = INFINITY;
            } else if (bt == "MI") d.vlo[j] = -INFINITY;
            else if (bt == "PL") d.vhi[j] = INFINITY;
            else if (bt == "BV") {
                d.vlo[j] = 0.0;
                d.vhi[j] = 1.0;
            }
            break;
        }
        case Sec::QuadObj: quad_entry(tok, true); break;
        case Sec::QMatrix: quad_entry(tok, false); break;
        default: break;
        }
    }
    if (!rows_done) finalize_rows();
    d.m = (int)row_type.size();
    d.n = (int)d.obj.size();
    d.obj_const = d.maximize ? -obj_const : obj_const;
    if (d.maximize) {
        for (double &c : d.obj) c = -c;
        for (double &q : d.qv) q = -q;
    }
    return true;
}

// Triplets -> CSR over `rows` rows, entries of a row in triplet order
// (a stable counting sort); arrays from malloc, as the C structs expect.
inline void triplets_to_csr(int rows, const std::vector<int> &ti, const std::vector<int> &tj,
                            const std::vector<double> &tv, int **rp_out, int **ci_out, double **v_out)
{
    const size_t nnz = tv.size();
    int *rp = (int *)calloc((size_t)rows + 1, sizeof(int));
    int *ci = (int *)malloc(std::max<size_t>(nnz, 1) * sizeof(int));
    double *vv = (double *)malloc(std::max<size_t>(nnz, 1) * sizeof(double));
    for (size_t k = 0; k < nnz; ++k) rp[ti[k] + 1]++;
    for (int i = 0; i < rows; ++i) rp[i + 1] += rp[i];
    std::vector<int> pos(rp, rp + rows + 1);
    for (size_t k = 0; k < nnz; ++k) {
        const int dst = pos[ti[k]]++;
        ci[dst] = tj[k];
        vv[dst] = tv[k];
    }
    *rp_out = rp;
    *ci_out = ci;
    *v_out = vv;
}

inline double *malloc_copy(const std::vector<double> &v)
{
    double *p = (double *)malloc(std::max<size_t>(v.size(), 1) * sizeof(double));
    if (!v.empty()) memcpy(p, v.data(), v.size() * sizeof(double));
    return p;
}

}  // namespace mps
}  // namespace pdal
