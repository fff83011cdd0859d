// ORACLE — test infrastructure only.  Not part of the product: libganit_hip.so and the ganit CLI never
// link this file.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load it
// (as oracle/liboracle.so through ctypes), and only as the checker / CPU baseline.
//
// A plain C++/OpenMP restatement of the reference's restarted-PDHG path, written to reproduce its
// arithmetic step for step (same operation order, same OpenMP loop shapes, no FMA contraction) so
// that at a fixed OMP thread count its iteration trajectory matches the reference CPU binary:
//   * CSR build / transpose ............ src/lp.cpp:9-72
//   * MPS/QPS reader ................... src/mps.cpp:19-272
//   * scaling + norm estimation ........ src/scaling.cpp:8-155, src/scaling.hpp:8-23
//   * CPU backend ...................... src/backend_cpu.cpp:25-156
//   * driver (phase / evaluate / run) .. src/pdhg_impl.hpp:36-328
// Extensions with no reference code (BF16 matrix values, Kahan-compensated FP32 reductions, FP64 evaluation of
// reduced-precision checks) are marked where they appear; they are pinned by tests against closed forms and error
// bounds, and the FP64 evaluation by being the reference's own FP64 evaluation applied to exactly widened vectors.
//
// How each part is pinned (tests/test_oracle.py): the reference binary's golden JSON lines
// (tests/golden/reference_runs.jsonl) for end-to-end runs, brute-force / closed-form checks for the
// kernels, numpy.linalg.norm for the norm estimate, and invariants for the scaling.
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <numeric>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

namespace orc {

constexpr double kInf = std::numeric_limits<double>::infinity();

// ------------------------------------------------------------------ data (src/lp.cpp, include/ganit/lp.hpp)
struct Csr {
    int rows = 0, cols = 0;
    std::vector<int> ptr, idx;
    std::vector<double> val;
    int64_t nnz() const { return (int64_t)val.size(); }
};

struct LP {
    std::string name;
    int m = 0, n = 0;
    Csr A, Q;
    std::vector<double> c, l, u, lo, hi;
    double obj_const = 0.0;
    bool maximize = false;
    std::vector<std::string> row_names, col_names;
    bool has_q() const { return Q.nnz() > 0; }
};

// src/lp.cpp:9-53 — bucket by row, then per row: sort by column, sum duplicates, drop exact zeros.
Csr csr_from_triplets(int rows, int cols, const std::vector<int>& r, const std::vector<int>& c,
                      const std::vector<double>& v) {
    Csr a;
    a.rows = rows;
    a.cols = cols;
    std::vector<int> start(rows + 1, 0);
    for (size_t k = 0; k < v.size(); ++k) {
        if (r[k] < 0 || r[k] >= rows || c[k] < 0 || c[k] >= cols)
            throw std::runtime_error("csr_from_triplets: index out of range");
        start[r[k] + 1]++;
    }
    for (int i = 0; i < rows; ++i) start[i + 1] += start[i];
    std::vector<int> bidx(v.size());
    std::vector<double> bval(v.size());
    std::vector<int> fill(start.begin(), start.end() - 1);
    for (size_t k = 0; k < v.size(); ++k) {
        int p = fill[r[k]]++;
        bidx[p] = c[k];
        bval[p] = v[k];
    }
    a.ptr.assign(rows + 1, 0);
    std::vector<std::pair<int, double>> row;
    for (int i = 0; i < rows; ++i) {
        row.clear();
        for (int p = start[i]; p < start[i + 1]; ++p) row.emplace_back(bidx[p], bval[p]);
        // stable w.r.t. the reference only in the summed value order; std::sort on (col) keys with
        // equal columns leaves duplicate order unspecified, exactly as the reference does.
        std::sort(row.begin(), row.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
        size_t q = 0;
        while (q < row.size()) {
            const int col = row[q].first;
            double s = 0.0;
            while (q < row.size() && row[q].first == col) s += row[q++].second;
            if (s != 0.0) {
                a.idx.push_back(col);
                a.val.push_back(s);
            }
        }
        a.ptr[i + 1] = (int)a.idx.size();
    }
    return a;
}

// src/lp.cpp:55-72 — counting-sort transpose, stable in the source row order.
Csr transpose(const Csr& a) {
    Csr t;
    t.rows = a.cols;
    t.cols = a.rows;
    t.ptr.assign(a.cols + 1, 0);
    for (int j : a.idx) t.ptr[j + 1]++;
    for (int j = 0; j < a.cols; ++j) t.ptr[j + 1] += t.ptr[j];
    t.idx.resize(a.idx.size());
    t.val.resize(a.val.size());
    std::vector<int> next(t.ptr.begin(), t.ptr.end() - 1);
    for (int i = 0; i < a.rows; ++i)
        for (int p = a.ptr[i]; p < a.ptr[i + 1]; ++p) {
            const int q = next[a.idx[p]]++;
            t.idx[q] = i;
            t.val[q] = a.val[p];
        }
    return t;
}

// ------------------------------------------------------------------ MPS reader (src/mps.cpp:19-272)
static std::string read_text(const std::string& path) {
    const bool gz = path.size() > 3 && path.compare(path.size() - 3, 3, ".gz") == 0;
    if (gz) {  // mps.cpp:20-35: decompress through gzip -dc
        std::string cmd = "gzip -dc '" + path + "'";
        FILE* f = popen(cmd.c_str(), "r");
        if (!f) throw std::runtime_error("cannot run gzip on " + path);
        std::string out;
        char buf[1 << 16];
        size_t k;
        while ((k = fread(buf, 1, sizeof(buf), f)) > 0) out.append(buf, k);
        pclose(f);
        if (out.empty()) throw std::runtime_error("empty or unreadable file " + path);
        return out;
    }
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("cannot open " + path);
    std::ostringstream ss;
    ss << in.rdbuf();
    return ss.str();
}

static void tokens(const std::string& line, std::vector<std::string>& tok) {
    tok.clear();
    size_t i = 0;
    while (i < line.size()) {
        while (i < line.size() && std::isspace((unsigned char)line[i])) ++i;
        if (i >= line.size()) break;
        size_t j = i;
        while (j < line.size() && !std::isspace((unsigned char)line[j])) ++j;
        tok.emplace_back(line, i, j - i);
        i = j;
    }
}

static double parse_num(const std::string& s, int lineno) {
    char* end = nullptr;
    double v = std::strtod(s.c_str(), &end);
    if (end == s.c_str()) throw std::runtime_error("MPS line " + std::to_string(lineno) + ": bad number '" + s + "'");
    return v;
}

LP read_mps(const std::string& path) {
    enum Sec { None, Name, ObjSense, Rows, Columns, Rhs, Ranges, Bounds, QuadObj, QMatrix };
    const std::string text = read_text(path);
    LP lp;
    std::unordered_map<std::string, int> row_of, col_of;
    std::unordered_set<std::string> extra_n;
    std::string obj_name, last_col;
    std::vector<char> rtype, lower_set;
    std::vector<double> rhs, range;
    std::vector<int> tr, tc, qr, qc;
    std::vector<double> tv, qv;
    int cur = -1, lineno = 0;
    Sec sec = None;
    std::vector<std::string> tok;
    std::istringstream in(text);
    std::string line;
    auto col_index = [&](const std::string& nm) {
        auto it = col_of.find(nm);
        if (it == col_of.end())
            throw std::runtime_error("MPS: unknown column '" + nm + "' at line " + std::to_string(lineno));
        return it->second;
    };
    auto row_index = [&](const std::string& nm) -> int {  // -1: an ignored extra N row
        auto it = row_of.find(nm);
        if (it != row_of.end()) return it->second;
        if (extra_n.count(nm)) return -1;
        throw std::runtime_error("MPS: unknown row '" + nm + "' at line " + std::to_string(lineno));
    };
    while (std::getline(in, line)) {
        ++lineno;
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line.empty() || line[0] == '*') continue;
        tokens(line, tok);
        if (tok.empty()) continue;
        if (!std::isspace((unsigned char)line[0])) {  // section header (mps.cpp:108-128)
            const std::string& h = tok[0];
            if (h == "NAME") { sec = Name; if (tok.size() > 1) lp.name = tok[1]; }
            else if (h == "OBJSENSE") {
                sec = ObjSense;
                if (tok.size() > 1) lp.maximize = (tok[1] == "MAX" || tok[1] == "MAXIMIZE");
            }
            else if (h == "OBJSENCE") sec = ObjSense;  // misspelt header: sense only from the next line
            else if (h == "ROWS") sec = Rows;
            else if (h == "COLUMNS") sec = Columns;
            else if (h == "RHS") sec = Rhs;
            else if (h == "RANGES") sec = Ranges;
            else if (h == "BOUNDS") sec = Bounds;
            else if (h == "QUADOBJ") sec = QuadObj;
            else if (h == "QMATRIX" || h == "QSECTION") sec = QMatrix;
            else if (h == "ENDATA") break;
            else throw std::runtime_error("MPS: unknown section '" + h + "' at line " + std::to_string(lineno));
            continue;
        }
        if (sec == ObjSense) {
            lp.maximize = (tok[0] == "MAX" || tok[0] == "MAXIMIZE");
        } else if (sec == Rows) {
            if (tok.size() < 2) throw std::runtime_error("MPS: bad ROWS line " + std::to_string(lineno));
            const char t = (char)std::toupper(tok[0][0]);
            if (t == 'N') {  // first N row is the objective, later ones are ignored (mps.cpp:137-139)
                if (obj_name.empty()) obj_name = tok[1]; else extra_n.insert(tok[1]);
            } else {
                row_of[tok[1]] = lp.m++;
                lp.row_names.push_back(tok[1]);
                rtype.push_back(t);
                rhs.push_back(0.0);
            }
        } else if (sec == Columns) {
            if (tok.size() >= 3 && tok[1] == "'MARKER'") continue;  // integrality is irrelevant to PDHG
            if (tok[0] != last_col) {
                last_col = tok[0];
                cur = lp.n++;
                col_of[tok[0]] = cur;
                lp.col_names.push_back(tok[0]);
                lp.c.push_back(0.0);
                lp.l.push_back(0.0);
                lp.u.push_back(kInf);
                lower_set.push_back(0);
            }
            for (size_t k = 1; k + 1 < tok.size(); k += 2) {
                const double v = parse_num(tok[k + 1], lineno);
                if (tok[k] == obj_name) { lp.c[cur] += v; continue; }
                const int r = row_index(tok[k]);
                if (r < 0) continue;
                tr.push_back(r); tc.push_back(cur); tv.push_back(v);
            }
        } else if (sec == Rhs || sec == Ranges) {
            const size_t first = (tok.size() % 2 == 1) ? 1 : 0;  // optional set name (mps.cpp:181)
            for (size_t k = first; k + 1 < tok.size(); k += 2) {
                const double v = parse_num(tok[k + 1], lineno);
                if (tok[k] == obj_name) { if (sec == Rhs) lp.obj_const = -v; continue; }  // mps.cpp:184-187
                const int r = row_index(tok[k]);
                if (r < 0) continue;
                if (sec == Rhs) rhs[r] = v;
                else { if (range.empty()) range.assign(lp.m, std::nan("")); range[r] = v; }
            }
        } else if (sec == Bounds) {
            const std::string t = tok[0];
            const bool valued = (t == "UP" || t == "LO" || t == "FX" || t == "LI" || t == "UI");
            std::string cname;
            double v = 0.0;
            if (valued) {
                if (tok.size() >= 4) { cname = tok[2]; v = parse_num(tok[3], lineno); }
                else if (tok.size() == 3) { cname = tok[1]; v = parse_num(tok[2], lineno); }
                else throw std::runtime_error("MPS: bad BOUNDS line " + std::to_string(lineno));
            } else {
                cname = tok.size() >= 3 ? tok[2] : tok[1];
            }
            const int j = col_index(cname);
            if (t == "UP" || t == "UI") {  // negative UP on a default-zero lower frees it (mps.cpp:215-220)
                lp.u[j] = v;
                if (v < 0 && lp.l[j] == 0.0 && !lower_set[j]) lp.l[j] = -kInf;
            } else if (t == "LO" || t == "LI") { lp.l[j] = v; lower_set[j] = 1; }
            else if (t == "FX") { lp.l[j] = v; lp.u[j] = v; lower_set[j] = 1; }
            else if (t == "FR") { lp.l[j] = -kInf; lp.u[j] = kInf; lower_set[j] = 1; }
            else if (t == "MI") { lp.l[j] = -kInf; lower_set[j] = 1; }
            else if (t == "PL") { lp.u[j] = kInf; }
            else if (t == "BV") { lp.l[j] = 0; lp.u[j] = 1; lower_set[j] = 1; }
            else throw std::runtime_error("MPS: unknown bound type '" + t + "'");
        } else if (sec == QuadObj || sec == QMatrix) {
            if (tok.size() < 3) throw std::runtime_error("MPS: bad quadratic line " + std::to_string(lineno));
            const int a = col_index(tok[0]), b = col_index(tok[1]);
            const double v = parse_num(tok[2], lineno);
            qr.push_back(a); qc.push_back(b); qv.push_back(v);
            if (sec == QuadObj && a != b) { qr.push_back(b); qc.push_back(a); qv.push_back(v); }  // mps.cpp:239
        }
    }
    // Row bounds from type, rhs and ranges (mps.cpp:247-265).
    lp.lo.assign(lp.m, 0.0);
    lp.hi.assign(lp.m, 0.0);
    for (int i = 0; i < lp.m; ++i) {
        const double b = rhs[i];
        if (rtype[i] == 'E') { lp.lo[i] = b; lp.hi[i] = b; }
        else if (rtype[i] == 'L') { lp.lo[i] = -kInf; lp.hi[i] = b; }
        else if (rtype[i] == 'G') { lp.lo[i] = b; lp.hi[i] = kInf; }
        else throw std::runtime_error("MPS: unknown row type for " + lp.row_names[i]);
        if (!range.empty() && !std::isnan(range[i])) {
            const double r = range[i];
            if (rtype[i] == 'E') { if (r > 0) lp.hi[i] = b + r; else lp.lo[i] = b + r; }
            else if (rtype[i] == 'L') lp.lo[i] = b - std::fabs(r);
            else lp.hi[i] = b + std::fabs(r);
        }
    }
    lp.A = csr_from_triplets(lp.m, lp.n, tr, tc, tv);
    if (!qv.empty()) lp.Q = csr_from_triplets(lp.n, lp.n, qr, qc, qv);
    else { lp.Q.rows = lp.n; lp.Q.cols = lp.n; lp.Q.ptr.assign(lp.n + 1, 0); }
    return lp;
}

// ------------------------------------------------------------------ scaling (src/scaling.cpp:8-155)
struct ScaledLP {
    int m = 0, n = 0;
    Csr A, AT, Q;
    std::vector<double> c, l, u, lo, hi, R, C, Rres, Cres;
    double bscale = 1.0, cscale = 1.0, normQ = 0.0, obj_const = 0.0, bnorm = 0.0, cnorm = 0.0;
    bool maximize = false;
    bool has_q() const { return Q.nnz() > 0; }
};

double norm2(const std::vector<double>& v) {  // scaling.cpp:8-12, sequential sum of squares
    double s = 0.0;
    for (double x : v) s += x * x;
    return std::sqrt(s);
}

// scaling.cpp:15-24: one representative finite magnitude per row (max of |lo|, |hi| over finite ends).
double finite_bound_norm(const std::vector<double>& lo, const std::vector<double>& hi) {
    double s = 0.0;
    for (size_t i = 0; i < lo.size(); ++i) {
        const double a = std::isfinite(lo[i]) ? std::fabs(lo[i]) : 0.0;
        const double b = std::isfinite(hi[i]) ? std::fabs(hi[i]) : 0.0;
        const double v = std::max(a, b);
        s += v * v;
    }
    return std::sqrt(s);
}

// scaling.cpp:27-31: A <- diag(rs) A diag(cs); accumulate into R, C.
static void rescale(Csr& a, const std::vector<double>& rs, const std::vector<double>& cs, std::vector<double>& R,
                    std::vector<double>& C) {
    for (int i = 0; i < a.rows; ++i)
        for (int p = a.ptr[i]; p < a.ptr[i + 1]; ++p) a.val[p] *= rs[i] * cs[a.idx[p]];
    for (int i = 0; i < a.rows; ++i) R[i] *= rs[i];
    for (int j = 0; j < a.cols; ++j) C[j] *= cs[j];
}

static void inv_sqrt_or_one(std::vector<double>& v) {
    for (auto& t : v) t = t > 0 ? 1.0 / std::sqrt(t) : 1.0;
}

ScaledLP scale_lp(const LP& lp, int ruiz_iters, bool pock_chambolle) {
    ScaledLP s;
    s.m = lp.m;
    s.n = lp.n;
    s.A = lp.A;
    s.R.assign(lp.m, 1.0);
    s.C.assign(lp.n, 1.0);
    s.maximize = lp.maximize;
    const double sign = lp.maximize ? -1.0 : 1.0;
    std::vector<double> rs(lp.m), cs(lp.n);
    for (int it = 0; it < ruiz_iters; ++it) {  // scaling.cpp:45-57: inf-norm Ruiz
        std::fill(rs.begin(), rs.end(), 0.0);
        std::fill(cs.begin(), cs.end(), 0.0);
        for (int i = 0; i < lp.m; ++i)
            for (int p = s.A.ptr[i]; p < s.A.ptr[i + 1]; ++p) {
                const double a = std::fabs(s.A.val[p]);
                rs[i] = std::max(rs[i], a);
                cs[s.A.idx[p]] = std::max(cs[s.A.idx[p]], a);
            }
        inv_sqrt_or_one(rs);
        inv_sqrt_or_one(cs);
        rescale(s.A, rs, cs, s.R, s.C);
    }
    if (pock_chambolle) {  // scaling.cpp:58-70: alpha = 1 (row/col 1-norms)
        std::fill(rs.begin(), rs.end(), 0.0);
        std::fill(cs.begin(), cs.end(), 0.0);
        for (int i = 0; i < lp.m; ++i)
            for (int p = s.A.ptr[i]; p < s.A.ptr[i + 1]; ++p) {
                const double a = std::fabs(s.A.val[p]);
                rs[i] += a;
                cs[s.A.idx[p]] += a;
            }
        inv_sqrt_or_one(rs);
        inv_sqrt_or_one(cs);
        rescale(s.A, rs, cs, s.R, s.C);
    }
    s.AT = transpose(s.A);
    s.c.resize(lp.n); s.l.resize(lp.n); s.u.resize(lp.n);
    for (int j = 0; j < lp.n; ++j) {
        s.c[j] = sign * lp.c[j] * s.C[j];
        s.l[j] = lp.l[j] / s.C[j];
        s.u[j] = lp.u[j] / s.C[j];
    }
    s.lo.resize(lp.m); s.hi.resize(lp.m);
    for (int i = 0; i < lp.m; ++i) {
        s.lo[i] = lp.lo[i] * s.R[i];
        s.hi[i] = lp.hi[i] * s.R[i];
    }
    s.bscale = finite_bound_norm(s.lo, s.hi) + 1.0;  // scaling.cpp:87-95
    s.cscale = norm2(s.c) + 1.0;
    for (auto& v : s.c) v /= s.cscale;
    for (auto& v : s.l) v /= s.bscale;
    for (auto& v : s.u) v /= s.bscale;
    for (auto& v : s.lo) v /= s.bscale;
    for (auto& v : s.hi) v /= s.bscale;
    s.Rres.resize(lp.m); s.Cres.resize(lp.n);
    for (int i = 0; i < lp.m; ++i) s.Rres[i] = s.R[i] / s.bscale;
    for (int j = 0; j < lp.n; ++j) s.Cres[j] = s.C[j] / s.cscale;
    s.obj_const = sign * lp.obj_const;
    s.Q.rows = lp.n; s.Q.cols = lp.n; s.Q.ptr.assign(lp.n + 1, 0);
    if (lp.has_q()) {  // scaling.cpp:101-123: Q' = sign C Q C bscale/cscale, ||Q'|| by 60 power steps
        s.Q = lp.Q;
        for (int i = 0; i < lp.n; ++i)
            for (int p = s.Q.ptr[i]; p < s.Q.ptr[i + 1]; ++p)
                s.Q.val[p] *= sign * s.C[i] * s.C[s.Q.idx[p]] * s.bscale / s.cscale;
        std::vector<double> v(lp.n, 1.0), w(lp.n);
        const double nv = norm2(v);
        for (auto& t : v) t /= nv;
        double lam = 0.0;
        for (int it = 0; it < 60; ++it) {
            for (int i = 0; i < lp.n; ++i) {
                double t = 0.0;
                for (int p = s.Q.ptr[i]; p < s.Q.ptr[i + 1]; ++p) t += s.Q.val[p] * v[s.Q.idx[p]];
                w[i] = t;
            }
            const double nw = norm2(w);
            if (nw == 0.0) break;
            lam = nw;
            for (int i = 0; i < lp.n; ++i) v[i] = w[i] / nw;
        }
        s.normQ = lam * 1.01;
    }
    s.bnorm = finite_bound_norm(lp.lo, lp.hi);  // original data (scaling.cpp:124-125)
    s.cnorm = norm2(lp.c);
    return s;
}

// scaling.cpp:131-155: power iteration on A'^T A' from the normalised ones vector.
double estimate_norm(const ScaledLP& s, int iters = 50) {
    if (s.A.nnz() == 0) return 1.0;
    std::vector<double> v(s.n, 1.0), w(s.m), z(s.n);
    const double nv = norm2(v);
    for (auto& t : v) t /= nv;
    double lambda = 0.0;
    for (int it = 0; it < iters; ++it) {
        for (int i = 0; i < s.m; ++i) {
            double t = 0.0;
            for (int p = s.A.ptr[i]; p < s.A.ptr[i + 1]; ++p) t += s.A.val[p] * v[s.A.idx[p]];
            w[i] = t;
        }
        for (int j = 0; j < s.n; ++j) {
            double t = 0.0;
            for (int p = s.AT.ptr[j]; p < s.AT.ptr[j + 1]; ++p) t += s.AT.val[p] * w[s.AT.idx[p]];
            z[j] = t;
        }
        const double nz = norm2(z);
        if (nz == 0.0) break;
        lambda = nz;
        for (int j = 0; j < s.n; ++j) v[j] = z[j] / nz;
    }
    return std::sqrt(lambda) * 1.01;
}

// ------------------------------------------------------------------ BF16 (extension, no reference code)
// bfloat16 storage: the high 16 bits of an IEEE binary32, rounded to nearest-even from float.
struct bf16 {
    uint16_t bits = 0;
    bf16() = default;
    explicit bf16(double d) {
        const float f = (float)d;
        uint32_t u;
        std::memcpy(&u, &f, 4);
        if ((u & 0x7fffffffu) > 0x7f800000u) { bits = (uint16_t)((u >> 16) | 0x40); return; }  // NaN stays NaN
        const uint32_t rounding = 0x7fffu + ((u >> 16) & 1u);
        bits = (uint16_t)((u + rounding) >> 16);
    }
    explicit operator float() const {
        const uint32_t u = (uint32_t)bits << 16;
        float f;
        std::memcpy(&f, &u, 4);
        return f;
    }
    explicit operator double() const { return (double)(float)(*this); }
};

// ------------------------------------------------------------------ CPU backend (src/backend_cpu.cpp:25-156)
template <class M>
struct CsrV {
    int rows = 0;
    const int* ptr = nullptr;
    const int* idx = nullptr;
    std::vector<M> val;
    CsrV() = default;
    explicit CsrV(const Csr& a) : rows(a.rows), ptr(a.ptr.data()), idx(a.idx.data()) {
        val.reserve(a.val.size());
        for (double v : a.val) val.push_back(M(v));
    }
};

template <class R>
static std::vector<R> narrow(const std::vector<double>& v) { return std::vector<R>(v.begin(), v.end()); }

// widen a stored matrix value to the vector type: float(bf16) first for BF16 values
template <class R, class M>
static inline R widen(M v) { return R(v); }
template <>
inline float widen<float, bf16>(bf16 v) { return (float)v; }
template <>
inline double widen<double, bf16>(bf16 v) { return (double)v; }

// Kahan-compensated FP32 sum of f(i) over [0, k) (extension, no reference code): every term is rounded to
// float, each OpenMP thread accumulates its share in float with a running compensation (Kahan 1965; the error
// of one thread's sum is at most (2u + O(k u^2)) sum |term|, u = 2^-24), and the per-thread float partials are
// added in double.
template <class F>
double kahan_fp32_sum(int k, F f) {
    double total = 0.0;
#pragma omp parallel reduction(+ : total)
    {
        float sacc = 0.0f, comp = 0.0f;
#pragma omp for
        for (int i = 0; i < k; ++i) {
            const float yv = (float)f(i) - comp;
            const volatile float t = sacc + yv;
            comp = (t - sacc) - yv;
            sacc = t;
        }
        total += (double)sacc;
    }
    return total;
}

// Reductions accumulate in double as the reference does (backend_cpu.cpp:101-155).  With
// kahan = true they instead use kahan_fp32_sum.
template <class M, class R>
struct CpuBackend {
    using Vec = std::vector<R>;
    const ScaledLP& s;
    CsrV<M> A, AT, Q;
    Vec c, l, u, lo, hi, Rres, Cres;
    bool kahan = false;

    explicit CpuBackend(const ScaledLP& lp, bool kahan_ = false)
        : s(lp), A(lp.A), AT(lp.AT), Q(lp.Q), c(narrow<R>(lp.c)), l(narrow<R>(lp.l)), u(narrow<R>(lp.u)),
          lo(narrow<R>(lp.lo)), hi(narrow<R>(lp.hi)), Rres(narrow<R>(lp.Rres)), Cres(narrow<R>(lp.Cres)),
          kahan(kahan_) {}

    Vec zeros(int k) { return Vec(k, R(0)); }
    Vec from_host(const std::vector<double>& h) { return narrow<R>(h); }
    std::vector<double> to_host(const Vec& v) { return std::vector<double>(v.begin(), v.end()); }
    void copy(const Vec& src, Vec& dst) { dst = src; }
    void fill(Vec& v, double a) { std::fill(v.begin(), v.end(), R(a)); }

    static void spmv(const CsrV<M>& a, const Vec& x, Vec& y) {  // backend_cpu.cpp:55-62
#pragma omp parallel for schedule(dynamic, 256)
        for (int i = 0; i < a.rows; ++i) {
            R t = 0;
            for (int p = a.ptr[i]; p < a.ptr[i + 1]; ++p) t += widen<R, M>(a.val[p]) * x[a.idx[p]];
            y[i] = t;
        }
    }
    void spmv_A(const Vec& x, Vec& ax) { spmv(A, x, ax); }
    void spmv_AT(const Vec& y, Vec& aty) { spmv(AT, y, aty); }
    void spmv_Q(const Vec& x, Vec& qx) { spmv(Q, x, qx); }

    // backend_cpu.cpp:67-80: x+ = clamp(x - tau (c - aty + qx), l, u); xbar = 2x+ - x; sumx += x+
    void primal_step(const Vec& x, const Vec& aty, const Vec* qx, double tau_d, Vec& xn, Vec& xbar, Vec& sumx) {
        const R tau = R(tau_d);
        const R* q = qx ? qx->data() : nullptr;
        const int n = s.n;
#pragma omp parallel for
        for (int j = 0; j < n; ++j) {
            R g = c[j] - aty[j] + (q ? q[j] : R(0));
            R v = x[j] - tau * g;
            v = std::min(std::max(v, l[j]), u[j]);
            xn[j] = v;
            xbar[j] = R(2) * v - x[j];
            sumx[j] += v;
        }
    }
    // backend_cpu.cpp:81-92: q = y - sigma axbar; y+ = q + clamp(-q, sigma lo, sigma hi); sumy += y+
    void dual_step(const Vec& y, const Vec& axbar, double sigma_d, Vec& yn, Vec& sumy) {
        const R sigma = R(sigma_d);
        const int m = s.m;
#pragma omp parallel for
        for (int i = 0; i < m; ++i) {
            R q = y[i] - sigma * axbar[i];
            R t = std::min(std::max(-q, sigma * lo[i]), sigma * hi[i]);
            R v = q + t;
            yn[i] = v;
            sumy[i] += v;
        }
    }
    void sub(const Vec& a, const Vec& b, Vec& d) {
#pragma omp parallel for
        for (size_t i = 0; i < a.size(); ++i) d[i] = a[i] - b[i];
    }
    void scale_into(const Vec& src, double a, Vec& dst) {
#pragma omp parallel for
        for (size_t i = 0; i < src.size(); ++i) dst[i] = R(a * double(src[i]));
    }

    // Generic FP64 / Kahan-FP32 sum of f(i) over [0, k).
    template <class F>
    double sum(int k, F f) {
        if (!kahan) {
            double acc = 0.0;
#pragma omp parallel for reduction(+ : acc)
            for (int i = 0; i < k; ++i) acc += f(i);
            return acc;
        }
        return kahan_fp32_sum(k, f);
    }

    double primal_res2(const Vec& ax) {  // backend_cpu.cpp:101-112
        return sum(s.m, [&](int i) {
            const double a = ax[i];
            const double p = std::min(std::max(a, double(lo[i])), double(hi[i]));
            const double r = (a - p) / double(Rres[i]);
            return r * r;
        });
    }
    // backend_cpu.cpp:113-131: reduced cost z = cw c - aty + qx; a bounded side contributes z*bound to the
    // dual objective, an unbounded side contributes z to the residual.
    void dual_res(const Vec& aty, const Vec* qx, double& rd2, double& vobj, bool with_c = true) {
        const double cw = with_c ? 1.0 : 0.0;
        const R* q = qx ? qx->data() : nullptr;
        auto zf = [&](int j) { return cw * double(c[j]) - double(aty[j]) + (q ? double(q[j]) : 0.0); };
        if (!kahan) {
            double a2 = 0.0, ob = 0.0;
            const int n = s.n;
#pragma omp parallel for reduction(+ : a2, ob)
            for (int j = 0; j < n; ++j) {
                const double z = zf(j);
                const double lj = l[j], uj = u[j];
                double r = 0.0;
                if (z > 0) { if (std::isfinite(lj)) ob += z * lj; else r = z; }
                else if (z < 0) { if (std::isfinite(uj)) ob += z * uj; else r = z; }
                r /= double(Cres[j]);
                a2 += r * r;
            }
            rd2 = a2;
            vobj = ob;
            return;
        }
        rd2 = sum(s.n, [&](int j) {
            const double z = zf(j);
            double r = 0.0;
            if (z > 0) { if (!std::isfinite(double(l[j]))) r = z; }
            else if (z < 0) { if (!std::isfinite(double(u[j]))) r = z; }
            r /= double(Cres[j]);
            return r * r;
        });
        vobj = sum(s.n, [&](int j) {
            const double z = zf(j);
            if (z > 0 && std::isfinite(double(l[j]))) return z * double(l[j]);
            if (z < 0 && std::isfinite(double(u[j]))) return z * double(u[j]);
            return 0.0;
        });
    }
    double row_dual_obj(const Vec& y) {  // backend_cpu.cpp:132-142
        return sum(s.m, [&](int i) {
            const double yi = y[i];
            if (yi > 0 && std::isfinite(double(lo[i]))) return yi * double(lo[i]);
            if (yi < 0 && std::isfinite(double(hi[i]))) return yi * double(hi[i]);
            return 0.0;
        });
    }
    double dot(const Vec& a, const Vec& b) {
        return sum((int)a.size(), [&](int i) { return double(a[i]) * double(b[i]); });
    }
    double dot_c(const Vec& x) { return dot(c, x); }
    double diff2(const Vec& a, const Vec& b) {
        return sum((int)a.size(), [&](int i) { const double d = double(a[i]) - double(b[i]); return d * d; });
    }
};

// ------------------------------------------------------------------ driver (src/pdhg_impl.hpp)
enum Precision { kDouble = 0, kMixed = 1, kDynamic = 2 };
enum Status { kOptimal = 0, kPrimalInfeasible = 1, kIterationLimit = 2, kTimeLimit = 3, kNumericalError = 4 };
enum Exit { eConverged, eStalled, eIterLimit, eTimeLimit, eInfeasible, eNumerical };

struct Options {
    double tol = 1e-4;
    long max_iter = 10000000;
    double time_limit = 3600.0;
    int check_every = 64;
    int ruiz_iters = 10;
    bool pock_chambolle = true;
    int verbose = 0;
    double log_interval = 1.0;
    double infeas_tol = 1e-8;
    int precision = kDouble;
    int stall_checks = 12;
    int stall_checks_l2 = 24;
    int l1_matrix = 0;   // extension: 0 = FP32 matrix values in L1/L2, 1 = BF16
    int kahan = 0;       // extension: Kahan-compensated FP32 reductions in L1
    int eval_fp64 = 0;   // extension: L1/L2 checks evaluate the iterate in FP64 (FP64 matrix and problem data,
                         // each iterate entry widened, exactly)
};

struct KKT {  // pdhg_impl.hpp:36-42
    double pobj = 0, dobj = 0, rel_p = 0, rel_d = 0, rel_gap = 0, err = 0;
    bool finite() const {
        return std::isfinite(pobj) && std::isfinite(dobj) && std::isfinite(rel_p) && std::isfinite(rel_d);
    }
    bool converged(double tol) const { return rel_p <= tol && rel_d <= tol && rel_gap <= tol; }
};

struct Shared {
    const ScaledLP& s;
    const Options& opt;
    double eta, normA, obj_scale;
    std::chrono::steady_clock::time_point t_solve;
    double elapsed() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t_solve).count();
    }
};

struct State {  // pdhg_impl.hpp:55-63
    std::vector<double> x, y, xr, yr;
    double restart_err = -1.0, w = 1.0;
    long it = 0;
    int restarts = 0;
    KKT kkt;
};

// Per-check trace (extension for tests: lets a caller compare trajectories check by check).
struct TraceRow { long it; double err_cur, err_avg, w; int restarted; };
static std::vector<TraceRow>* g_trace = nullptr;

template <class B>
static KKT kkt_at(B& be, const Shared& sh, const typename B::Vec& x, const typename B::Vec& y,
                  typename B::Vec& ax, typename B::Vec& aty, typename B::Vec& qx) {
    // pdhg_impl.hpp:92-109 (and the identical body of pdhg_evaluate, :211-229)
    const ScaledLP& s = sh.s;
    const bool hasq = s.has_q();
    KKT k;
    be.spmv_A(x, ax);
    const double rp2 = be.primal_res2(ax);
    be.spmv_AT(y, aty);
    double xqx = 0.0;
    if (hasq) { be.spmv_Q(x, qx); xqx = be.dot(x, qx); }
    double rd2 = 0, vobj = 0;
    be.dual_res(aty, hasq ? &qx : nullptr, rd2, vobj);
    const double robj = be.row_dual_obj(y);
    k.pobj = (be.dot_c(x) + 0.5 * xqx) * sh.obj_scale + s.obj_const;
    k.dobj = (robj + vobj - 0.5 * xqx) * sh.obj_scale + s.obj_const;
    k.rel_p = std::sqrt(rp2) / (1.0 + s.bnorm);
    k.rel_d = std::sqrt(rd2) / (1.0 + s.cnorm);
    k.rel_gap = std::fabs(k.pobj - k.dobj) / (1.0 + std::fabs(k.pobj) + std::fabs(k.dobj));
    k.err = std::sqrt(k.rel_p * k.rel_p + k.rel_d * k.rel_d + k.rel_gap * k.rel_gap);
    return k;
}

// e64 (extension, a reduced-precision level only): every KKT evaluation of the phase -- the initial point, the
// iterate, the average and the infeasibility ray -- is the FP64 one (kkt_at with the FP64 backend) of the level's
// vectors widened to double, exactly; the level's own reductions are then used only for the restart distances,
// which widen each entry too.  Kahan-compensated level reductions do not apply with it.
template <class B>
static Exit phase(Shared& sh, State& st, double phase_tol, bool stall_exit, const char* tag, int stall_checks,
                  bool kahan, bool e64 = false) {
    const ScaledLP& s = sh.s;
    const Options& opt = sh.opt;
    B be(s, kahan && !e64);
    using Vec = typename B::Vec;
    const int m = s.m, n = s.n;
    const bool hasq = s.has_q();
    Vec x = be.from_host(st.x), xn = be.zeros(n), xbar = be.zeros(n);
    Vec y = be.from_host(st.y), yn = be.zeros(m);
    Vec aty = be.zeros(n), axbar = be.zeros(m);
    Vec sumx = be.zeros(n), sumy = be.zeros(m), xa = be.zeros(n), ya = be.zeros(m);
    Vec xr = be.from_host(st.xr.empty() ? st.x : st.xr), yr = be.from_host(st.yr.empty() ? st.y : st.yr);
    Vec ax_tmp = be.zeros(m), aty_tmp = be.zeros(n), ray = be.zeros(m);
    Vec qx = be.zeros(hasq ? n : 0), qx_tmp = be.zeros(hasq ? n : 0);
    double w = st.w;

    auto steps = [&](double& tau, double& sigma) {  // pdhg_impl.hpp:87-91
        sigma = sh.eta * w;
        tau = sh.eta / w;
        if (hasq) tau = std::min(tau, 0.998 / (0.5 * s.normQ + sigma * sh.normA * sh.normA));
    };
    // the FP64 evaluation (e64): FP64 backend and work vectors, the level's vectors widened into w1, w2
    CpuBackend<double, double> be64(s);
    std::vector<double> w1, w2, ax64(e64 ? m : 0), aty64(e64 ? n : 0), qx64(e64 && hasq ? n : 0);
    auto widen_into = [](const Vec& v, std::vector<double>& d) { d.assign(v.begin(), v.end()); };
    auto kkt_of = [&](const Vec& xv, const Vec& yv) {
        if (!e64) return kkt_at(be, sh, xv, yv, ax_tmp, aty_tmp, qx_tmp);
        widen_into(xv, w1);
        widen_into(yv, w2);
        return kkt_at(be64, sh, w1, w2, ax64, aty64, qx64);
    };
    auto infeasible = [&]() {  // pdhg_impl.hpp:110-118
        double rd2 = 0, vobj = 0, robj = 0;
        if (e64) {  // the ray y - yr formed in FP64 from the widened entries, then the FP64 passes
            w2.resize(m);
            for (int i = 0; i < m; ++i) w2[i] = double(y[i]) - double(yr[i]);
            be64.spmv_AT(w2, aty64);
            be64.dual_res(aty64, nullptr, rd2, vobj, false);
            robj = be64.row_dual_obj(w2);
        } else {
            be.sub(y, yr, ray);
            be.spmv_AT(ray, aty_tmp);
            be.dual_res(aty_tmp, nullptr, rd2, vobj, false);
            robj = be.row_dual_obj(ray);
        }
        const double dray = (robj + vobj) * sh.obj_scale;
        if (!(dray > 0)) return false;
        return std::sqrt(rd2) * (1.0 + s.bnorm) / dray <= opt.infeas_tol;
    };

    KKT k_restart = kkt_of(x, y);
    if (st.restart_err > 0) k_restart.err = std::max(k_restart.err, st.restart_err);
    double prev_cand = std::numeric_limits<double>::infinity();
    long it_since = 0;
    double best_err = k_restart.err;
    int no_progress = 0;
    Exit why = eIterLimit;
    KKT best;
    bool best_avg = false;
    double last_log = -1e9;

    while (true) {  // pdhg_impl.hpp:130-190
        double tau, sigma;
        steps(tau, sigma);
        be.spmv_AT(y, aty);
        if (hasq) be.spmv_Q(x, qx);
        be.primal_step(x, aty, hasq ? &qx : nullptr, tau, xn, xbar, sumx);
        be.spmv_A(xbar, axbar);
        be.dual_step(y, axbar, sigma, yn, sumy);
        std::swap(x, xn);
        std::swap(y, yn);
        ++st.it;
        ++it_since;
        if (st.it % opt.check_every != 0) continue;

        KKT kc = kkt_of(x, y);
        be.scale_into(sumx, 1.0 / (double)it_since, xa);
        be.scale_into(sumy, 1.0 / (double)it_since, ya);
        KKT ka = kkt_of(xa, ya);
        const double t = sh.elapsed();
        if (!kc.finite() && !ka.finite()) { why = eNumerical; best = kc; break; }
        const bool avg_better = ka.finite() && (!kc.finite() || ka.err < kc.err);
        const KKT cand = avg_better ? ka : kc;
        if (opt.verbose && t - last_log >= opt.log_interval) {
            std::printf("%9ld %8.2fs %4s %15.8e %15.8e %9.2e %9.2e %9.2e %9.2e\n", st.it, t, tag, cand.pobj,
                        cand.dobj, cand.rel_p, cand.rel_d, cand.rel_gap, w);
            last_log = t;
        }
        if (kc.converged(phase_tol)) { why = eConverged; best = kc; best_avg = false; break; }
        if (ka.converged(phase_tol)) { why = eConverged; best = ka; best_avg = true; break; }
        if (st.it >= opt.max_iter) { why = eIterLimit; best = cand; best_avg = avg_better; break; }
        if (t >= opt.time_limit) { why = eTimeLimit; best = cand; best_avg = avg_better; break; }
        if (it_since >= 4 * opt.check_every && infeasible()) { why = eInfeasible; best = cand; best_avg = avg_better; break; }
        if (cand.err < 0.8 * best_err) { best_err = cand.err; no_progress = 0; }
        else if (++no_progress >= stall_checks && stall_exit) { why = eStalled; best = cand; best_avg = avg_better; break; }

        const bool sufficient = cand.err <= 0.2 * k_restart.err;
        const bool necessary = cand.err <= 0.8 * k_restart.err && cand.err > prev_cand;
        const bool artificial = it_since >= 0.36 * (double)st.it;
        prev_cand = cand.err;
        const bool restart = sufficient || necessary || artificial;
        if (g_trace) g_trace->push_back({st.it, kc.err, ka.err, w, restart ? 1 : 0});
        if (restart) {  // pdhg_impl.hpp:177-189
            if (avg_better) { be.copy(xa, x); be.copy(ya, y); }
            const double dx = std::sqrt(be.diff2(x, xr)), dy = std::sqrt(be.diff2(y, yr));
            if (dx > 1e-10 && dy > 1e-10) w = std::exp(0.5 * std::log(dy / dx) + 0.5 * std::log(w));
            be.copy(x, xr);
            be.copy(y, yr);
            be.fill(sumx, 0.0);
            be.fill(sumy, 0.0);
            it_since = 0;
            k_restart = cand;
            prev_cand = std::numeric_limits<double>::infinity();
            ++st.restarts;
        }
    }
    st.x = be.to_host(best_avg ? xa : x);
    st.y = be.to_host(best_avg ? ya : y);
    st.xr = be.to_host(xr);
    st.yr = be.to_host(yr);
    st.restart_err = k_restart.err;
    st.w = w;
    st.kkt = best;
    return why;
}

static KKT evaluate64(Shared& sh, const State& st) {  // pdhg_impl.hpp:201-230
    CpuBackend<double, double> be(sh.s);
    using Vec = std::vector<double>;
    Vec x = st.x, y = st.y, ax(sh.s.m), aty(sh.s.n), qx(sh.s.has_q() ? sh.s.n : 0);
    return kkt_at(be, sh, x, y, ax, aty, qx);
}

struct Result {
    int status = kNumericalError;
    std::vector<double> x, y;
    double pobj = 0, dobj = 0, rel_p = 0, rel_d = 0, rel_gap = 0;
    long iterations = 0, fp32_iterations = 0, l2_iterations = 0;
    int restarts = 0;
    double setup_s = 0, solve_s = 0, norm_a = 0;
};

static int status_of(Exit e) {  // pdhg_impl.hpp:232-240
    switch (e) {
        case eConverged: return kOptimal;
        case eInfeasible: return kPrimalInfeasible;
        case eTimeLimit: return kTimeLimit;
        case eNumerical: return kNumericalError;
        default: return kIterationLimit;
    }
}

template <class ML>
static Exit run_level1(Shared& sh, State& st, const Options& opt) {
    return phase<CpuBackend<ML, float>>(sh, st, opt.tol, true, "L1", opt.stall_checks, opt.kahan != 0,
                                        opt.eval_fp64 != 0);
}
template <class ML>
static Exit run_level2(Shared& sh, State& st, const Options& opt) {
    return phase<CpuBackend<ML, double>>(sh, st, opt.tol, true, "L2", opt.stall_checks_l2, false, opt.eval_fp64 != 0);
}

Result run_pdhg(const LP& lp, const Options& opt) {  // pdhg_impl.hpp:243-328
    using Clock = std::chrono::steady_clock;
    const auto t0 = Clock::now();
    Result res;
    ScaledLP s = scale_lp(lp, opt.ruiz_iters, opt.pock_chambolle);
    const double normA = estimate_norm(s);
    Shared sh{s, opt, 0.998 / normA, normA, s.cscale * s.bscale, Clock::now()};
    res.norm_a = normA;
    State st;
    st.x.resize(s.n);
    for (int j = 0; j < s.n; ++j) st.x[j] = std::min(std::max(0.0, s.l[j]), s.u[j]);
    st.y.assign(s.m, 0.0);
    {
        const double cn = norm2(s.c), bn = finite_bound_norm(s.lo, s.hi);
        st.w = (cn > 1e-10 && bn > 1e-10) ? cn / bn : 1.0;
    }
    res.setup_s = std::chrono::duration<double>(Clock::now() - t0).count();
    sh.t_solve = Clock::now();
    Exit ex;
    long l1_iters = 0;
    auto finished = [&](Exit e) {
        const KKT k64 = evaluate64(sh, st);
        return (e == eConverged && k64.converged(opt.tol)) || e == eTimeLimit || e == eIterLimit || e == eInfeasible;
    };
    if (opt.precision != kDouble) {
        ex = opt.l1_matrix == 1 ? run_level1<bf16>(sh, st, opt) : run_level1<float>(sh, st, opt);
        l1_iters = st.it;
        bool done = finished(ex);
        if (!done && opt.precision == kDynamic) {
            ex = opt.l1_matrix == 1 ? run_level2<bf16>(sh, st, opt) : run_level2<float>(sh, st, opt);
            res.l2_iterations = st.it - l1_iters;
            done = finished(ex);
        }
        if (!done) ex = phase<CpuBackend<double, double>>(sh, st, opt.tol, false, "L3", 0, false);
    } else {
        ex = phase<CpuBackend<double, double>>(sh, st, opt.tol, false, "L3", 0, false);
    }
    const KKT fin = evaluate64(sh, st);
    int status = status_of(ex);
    if (status == kOptimal && !fin.converged(opt.tol * 1.0000001)) status = kNumericalError;
    res.solve_s = sh.elapsed();
    res.status = status;
    res.iterations = st.it;
    res.fp32_iterations = l1_iters;
    res.restarts = st.restarts;
    const double sign = s.maximize ? -1.0 : 1.0;
    res.x.resize(s.n);
    res.y.resize(s.m);
    for (int j = 0; j < s.n; ++j) res.x[j] = st.x[j] * s.C[j] * s.bscale;  // pdhg_impl.hpp:313-314
    for (int i = 0; i < s.m; ++i) res.y[i] = sign * st.y[i] * s.R[i] * s.cscale;
    res.pobj = sign * fin.pobj;
    res.dobj = sign * fin.dobj;
    res.rel_p = fin.rel_p;
    res.rel_d = fin.rel_d;
    res.rel_gap = fin.rel_gap;
    return res;
}

}  // namespace orc

// ================================================================== C API for tests (ctypes)
using namespace orc;

namespace {
thread_local std::string g_err;
struct Handle {
    LP lp;
    ScaledLP s;
    bool scaled = false;
};
Handle* H(void* h) { return static_cast<Handle*>(h); }
void ensure_scaled(Handle* h, int ruiz = 10, int pc = 1) {
    if (!h->scaled) { h->s = scale_lp(h->lp, ruiz, pc != 0); h->scaled = true; }
}
Csr csr_from(int rows, int cols, const int* ptr, const int* idx, const double* val) {
    Csr a;
    a.rows = rows;
    a.cols = cols;
    a.ptr.assign(ptr, ptr + rows + 1);
    a.idx.assign(idx, idx + ptr[rows]);
    a.val.assign(val, val + ptr[rows]);
    return a;
}
}  // namespace

extern "C" {

const char* oracle_last_error() { return g_err.c_str(); }
void oracle_set_threads(int k) { if (k > 0) omp_set_num_threads(k); }
int oracle_max_threads() { return omp_get_max_threads(); }

void* oracle_lp_read_mps(const char* path) {
    try {
        auto* h = new Handle;
        h->lp = read_mps(path);
        return h;
    } catch (const std::exception& e) { g_err = e.what(); return nullptr; }
}

// Problem from triplets (duplicates summed, zeros dropped, exactly as the reader does).  Q may be empty.
void* oracle_lp_from_triplets(int m, int n, long long nnz, const int* r, const int* c, const double* v,
                              const double* cost, const double* l, const double* u, const double* lo,
                              const double* hi, double obj_const, int maximize, long long qnnz, const int* qr,
                              const int* qc, const double* qv) {
    try {
        auto* h = new Handle;
        LP& lp = h->lp;
        lp.m = m; lp.n = n;
        lp.A = csr_from_triplets(m, n, std::vector<int>(r, r + nnz), std::vector<int>(c, c + nnz),
                                 std::vector<double>(v, v + nnz));
        if (qnnz > 0)
            lp.Q = csr_from_triplets(n, n, std::vector<int>(qr, qr + qnnz), std::vector<int>(qc, qc + qnnz),
                                     std::vector<double>(qv, qv + qnnz));
        else { lp.Q.rows = n; lp.Q.cols = n; lp.Q.ptr.assign(n + 1, 0); }
        lp.c.assign(cost, cost + n); lp.l.assign(l, l + n); lp.u.assign(u, u + n);
        lp.lo.assign(lo, lo + m); lp.hi.assign(hi, hi + m);
        lp.obj_const = obj_const;
        lp.maximize = maximize != 0;
        return h;
    } catch (const std::exception& e) { g_err = e.what(); return nullptr; }
}

void oracle_lp_free(void* h) { delete H(h); }

// dims[0..4] = m, n, nnz(A), nnz(Q), maximize
void oracle_lp_dims(void* h, long long* dims) {
    const LP& lp = H(h)->lp;
    dims[0] = lp.m; dims[1] = lp.n; dims[2] = lp.A.nnz(); dims[3] = lp.Q.nnz(); dims[4] = lp.maximize;
}
// Copies the original-space data out (arrays sized from oracle_lp_dims); any pointer may be null.
void oracle_lp_export(void* h, int* aptr, int* aidx, double* aval, int* qptr, int* qidx, double* qval, double* c,
                      double* l, double* u, double* lo, double* hi, double* obj_const) {
    const LP& lp = H(h)->lp;
    auto cp = [](auto* dst, const auto& src) { if (dst) std::copy(src.begin(), src.end(), dst); };
    cp(aptr, lp.A.ptr); cp(aidx, lp.A.idx); cp(aval, lp.A.val);
    cp(qptr, lp.Q.ptr); cp(qidx, lp.Q.idx); cp(qval, lp.Q.val);
    cp(c, lp.c); cp(l, lp.l); cp(u, lp.u); cp(lo, lp.lo); cp(hi, lp.hi);
    if (obj_const) *obj_const = lp.obj_const;
}

// Scaled problem.  scal[0..8] = bscale, cscale, normQ, obj_const, bnorm, cnorm, normA, nnz(AT), maximize.
void oracle_scaled_scalars(void* h, int ruiz, int pc, double* scal) {
    Handle* hh = H(h);
    hh->scaled = false;
    ensure_scaled(hh, ruiz, pc);
    const ScaledLP& s = hh->s;
    scal[0] = s.bscale; scal[1] = s.cscale; scal[2] = s.normQ; scal[3] = s.obj_const;
    scal[4] = s.bnorm; scal[5] = s.cnorm; scal[6] = estimate_norm(s); scal[7] = (double)s.AT.nnz();
    scal[8] = s.maximize;
}
// Any pointer may be null.  Matrices in CSR; AT is the stored transpose.
void oracle_scaled_export(void* h, int* aptr, int* aidx, double* aval, int* tptr, int* tidx, double* tval,
                          int* qptr, int* qidx, double* qval, double* c, double* l, double* u, double* lo,
                          double* hi, double* R, double* C, double* Rres, double* Cres) {
    Handle* hh = H(h);
    ensure_scaled(hh);
    const ScaledLP& s = hh->s;
    auto cp = [](auto* dst, const auto& src) { if (dst) std::copy(src.begin(), src.end(), dst); };
    cp(aptr, s.A.ptr); cp(aidx, s.A.idx); cp(aval, s.A.val);
    cp(tptr, s.AT.ptr); cp(tidx, s.AT.idx); cp(tval, s.AT.val);
    cp(qptr, s.Q.ptr); cp(qidx, s.Q.idx); cp(qval, s.Q.val);
    cp(c, s.c); cp(l, s.l); cp(u, s.u); cp(lo, s.lo); cp(hi, s.hi);
    cp(R, s.R); cp(C, s.C); cp(Rres, s.Rres); cp(Cres, s.Cres);
}

// opts: [tol, max_iter, time_limit, check_every, ruiz_iters, pock_chambolle, verbose, infeas_tol, precision,
//        stall_checks, stall_checks_l2, l1_matrix(0 f32 / 1 bf16), kahan, eval_fp64]
// out:  [status, pobj, dobj, rel_p, rel_d, rel_gap, iterations, fp32_iterations, l2_iterations, restarts,
//        setup_s, solve_s, normA]
int oracle_solve(void* h, const double* o, double* out, double* x, double* y) {
    try {
        Options opt;
        opt.tol = o[0]; opt.max_iter = (long)o[1]; opt.time_limit = o[2]; opt.check_every = (int)o[3];
        opt.ruiz_iters = (int)o[4]; opt.pock_chambolle = o[5] != 0; opt.verbose = (int)o[6];
        opt.infeas_tol = o[7]; opt.precision = (int)o[8]; opt.stall_checks = (int)o[9];
        opt.stall_checks_l2 = (int)o[10]; opt.l1_matrix = (int)o[11]; opt.kahan = (int)o[12];
        opt.eval_fp64 = (int)o[13];
        Result r = run_pdhg(H(h)->lp, opt);
        const double vals[13] = {(double)r.status, r.pobj, r.dobj, r.rel_p, r.rel_d, r.rel_gap,
                                 (double)r.iterations, (double)r.fp32_iterations, (double)r.l2_iterations,
                                 (double)r.restarts, r.setup_s, r.solve_s, r.norm_a};
        std::copy(vals, vals + 13, out);
        if (x) std::copy(r.x.begin(), r.x.end(), x);
        if (y) std::copy(r.y.begin(), r.y.end(), y);
        return 0;
    } catch (const std::exception& e) { g_err = e.what(); return 1; }
}

// Same as oracle_solve, also returning the per-check trace: rows of [it, err_cur, err_avg, w, restarted].
int oracle_solve_traced(void* h, const double* o, double* out, double* x, double* y, double* trace, int max_rows,
                        int* rows) {
    std::vector<TraceRow> tr;
    g_trace = &tr;
    const int rc = oracle_solve(h, o, out, x, y);
    g_trace = nullptr;
    const int k = std::min<int>((int)tr.size(), max_rows);
    for (int i = 0; i < k; ++i) {
        trace[5 * i + 0] = (double)tr[i].it; trace[5 * i + 1] = tr[i].err_cur; trace[5 * i + 2] = tr[i].err_avg;
        trace[5 * i + 3] = tr[i].w; trace[5 * i + 4] = tr[i].restarted;
    }
    *rows = (int)tr.size();
    return rc;
}

// ---- kernel-level restatements for GPU parity tests.  mdt: 0 f64, 1 f32, 2 bf16 matrix values; vdt: 0 f64,
// 1 f32 vectors.  Vectors cross the API as double and are narrowed to the vector type first, exactly as the
// backend's from_host does; results are widened back.
}  // extern "C"
template <class M, class R>
static void spmv_t(int rows, const int* ptr, const int* idx, const double* val, const double* x, int ncols,
                   double* y) {
    Csr a;
    a.rows = rows;
    a.ptr.assign(ptr, ptr + rows + 1);
    a.idx.assign(idx, idx + ptr[rows]);
    a.val.assign(val, val + ptr[rows]);
    CsrV<M> av(a);
    std::vector<R> xv(x, x + ncols), yv(rows);
    CpuBackend<M, R>::spmv(av, xv, yv);
    std::copy(yv.begin(), yv.end(), y);
}
extern "C" {
int oracle_spmv(int mdt, int vdt, int rows, int cols, const int* ptr, const int* idx, const double* val,
                const double* x, double* y) {
    if (mdt == 0 && vdt == 0) spmv_t<double, double>(rows, ptr, idx, val, x, cols, y);
    else if (mdt == 1 && vdt == 0) spmv_t<float, double>(rows, ptr, idx, val, x, cols, y);
    else if (mdt == 1 && vdt == 1) spmv_t<float, float>(rows, ptr, idx, val, x, cols, y);
    else if (mdt == 2 && vdt == 0) spmv_t<bf16, double>(rows, ptr, idx, val, x, cols, y);
    else if (mdt == 2 && vdt == 1) spmv_t<bf16, float>(rows, ptr, idx, val, x, cols, y);
    else return 1;
    return 0;
}

// The reductions' Kahan-compensated FP32 sum (kahan_fp32_sum) of a double array, on the current OpenMP
// thread count.
double oracle_kahan_sum(long long k, const double* v) {
    return kahan_fp32_sum((int)k, [&](int i) { return v[i]; });
}

// BF16 rounding of a double array (float first, then round-to-nearest-even), widened back to double.
void oracle_round_bf16(long long k, const double* in, double* out) {
    for (long long i = 0; i < k; ++i) out[i] = (double)bf16(in[i]);
}

}  // extern "C"
template <class R>
static void primal_t(int n, const double* x, const double* aty, const double* qx, const double* c, const double* l,
                     const double* u, double tau_d, double* xn, double* xbar, double* sumx) {
    const R tau = R(tau_d);
    for (int j = 0; j < n; ++j) {  // backend_cpu.cpp:72-79, element by element
        R g = R(c[j]) - R(aty[j]) + (qx ? R(qx[j]) : R(0));
        R v = R(x[j]) - tau * g;
        v = std::min(std::max(v, R(l[j])), R(u[j]));
        const R s = R(sumx[j]) + v;
        xn[j] = v;
        xbar[j] = R(2) * v - R(x[j]);
        sumx[j] = s;
    }
}
extern "C" {
int oracle_primal_step(int vdt, int n, const double* x, const double* aty, const double* qx, const double* c,
                       const double* l, const double* u, double tau, double* xn, double* xbar, double* sumx) {
    if (vdt == 0) primal_t<double>(n, x, aty, qx, c, l, u, tau, xn, xbar, sumx);
    else primal_t<float>(n, x, aty, qx, c, l, u, tau, xn, xbar, sumx);
    return 0;
}
}  // extern "C"
template <class R>
static void dual_t(int m, const double* y, const double* axbar, const double* lo, const double* hi, double sigma_d,
                   double* yn, double* sumy) {
    const R sigma = R(sigma_d);
    for (int i = 0; i < m; ++i) {  // backend_cpu.cpp:85-91
        R q = R(y[i]) - sigma * R(axbar[i]);
        R t = std::min(std::max(-q, sigma * R(lo[i])), sigma * R(hi[i]));
        R v = q + t;
        yn[i] = v;
        sumy[i] = R(sumy[i]) + v;
    }
}
extern "C" {
int oracle_dual_step(int vdt, int m, const double* y, const double* axbar, const double* lo, const double* hi,
                     double sigma, double* yn, double* sumy) {
    if (vdt == 0) dual_t<double>(m, y, axbar, lo, hi, sigma, yn, sumy);
    else dual_t<float>(m, y, axbar, lo, hi, sigma, yn, sumy);
    return 0;
}

// KKT of a scaled-space point (x, y) in FP64 (pdhg_evaluate).  out: [pobj, dobj, rel_p, rel_d, rel_gap, err,
// rp2, rd2, vobj, robj, cx, xqx]
int oracle_kkt(void* h, const double* x, const double* y, double* out) {
    try {
        Handle* hh = H(h);
        ensure_scaled(hh);
        const ScaledLP& s = hh->s;
        Options opt;
        const double normA = estimate_norm(s);
        Shared sh{s, opt, 0.998 / normA, normA, s.cscale * s.bscale, std::chrono::steady_clock::now()};
        CpuBackend<double, double> be(s);
        std::vector<double> xv(x, x + s.n), yv(y, y + s.m), ax(s.m), aty(s.n), qx(s.has_q() ? s.n : 0);
        KKT k = kkt_at(be, sh, xv, yv, ax, aty, qx);
        double rd2 = 0, vobj = 0;
        be.dual_res(aty, s.has_q() ? &qx : nullptr, rd2, vobj);
        const double vals[12] = {k.pobj, k.dobj, k.rel_p, k.rel_d, k.rel_gap, k.err, be.primal_res2(ax), rd2, vobj,
                                 be.row_dual_obj(yv), be.dot_c(xv), s.has_q() ? be.dot(xv, qx) : 0.0};
        std::copy(vals, vals + 12, out);
        return 0;
    } catch (const std::exception& e) { g_err = e.what(); return 1; }
}

}  // extern "C"
