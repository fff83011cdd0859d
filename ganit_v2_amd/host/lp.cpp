// CSR construction, transpose and the MPS/QPS reader and writer of the ganit host library.
#include "lp.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string_view>
#include <unordered_map>
#include <unordered_set>

#include <omp.h>

namespace ganit {

Csr empty_csr(int rows, int cols) {
    Csr a;
    a.rows = rows;
    a.cols = cols;
    a.ptr.assign(rows + 1, 0);
    return a;
}

// Counting sort into row buckets, then per row: order by column, merge equal columns in their input order,
// drop exact zeros.  (Behaviour of the reference's lp.cpp:9-53.)
Csr csr_from_triplets(int rows, int cols, const std::vector<int>& r, const std::vector<int>& c,
                      const std::vector<double>& v) {
    if (r.size() != v.size() || c.size() != v.size()) throw std::runtime_error("csr_from_triplets: size mismatch");
    std::vector<int> head(rows + 1, 0);
    for (size_t k = 0; k < v.size(); ++k) {
        if (r[k] < 0 || r[k] >= rows || c[k] < 0 || c[k] >= cols)
            throw std::runtime_error("csr_from_triplets: index out of range");
        ++head[r[k] + 1];
    }
    for (int i = 0; i < rows; ++i) head[i + 1] += head[i];
    std::vector<std::pair<int, double>> bucket(v.size());
    {
        std::vector<int> at(head.begin(), head.end() - 1);
        for (size_t k = 0; k < v.size(); ++k) bucket[at[r[k]]++] = {c[k], v[k]};
    }
    Csr a = empty_csr(rows, cols);
    // rows are independent: each is ordered and merged in place (in parallel), then copied out.  A row whose
    // columns already increase strictly (column-major input) skips the sort -- its sorted order is unique, so
    // the result is the same; with equal columns the sort runs as always.
    std::vector<int> kept(rows, 0);
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < rows; ++i) {
        auto b = bucket.begin() + head[i], e = bucket.begin() + head[i + 1];
        bool strict = true;
        for (auto q = b; q + 1 < e && strict; ++q) strict = q->first < (q + 1)->first;
        if (!strict) std::sort(b, e, [](const auto& x, const auto& y) { return x.first < y.first; });
        auto w = b;
        for (auto p = b; p != e;) {
            const int col = p->first;
            double s = 0.0;
            for (; p != e && p->first == col; ++p) s += p->second;
            if (s != 0.0) *w++ = {col, s};
        }
        kept[i] = (int)(w - b);
    }
    for (int i = 0; i < rows; ++i) a.ptr[i + 1] = a.ptr[i] + kept[i];
    a.idx.resize((size_t)a.ptr[rows]);
    a.val.resize((size_t)a.ptr[rows]);
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < rows; ++i) {
        const auto b = bucket.begin() + head[i];
        for (int k = 0; k < kept[i]; ++k) {
            a.idx[(size_t)a.ptr[i] + k] = b[k].first;
            a.val[(size_t)a.ptr[i] + k] = b[k].second;
        }
    }
    return a;
}

Csr transpose(const Csr& a) {
    Csr t = empty_csr(a.cols, a.rows);
    for (int j : a.idx) ++t.ptr[j + 1];
    for (int j = 0; j < a.cols; ++j) t.ptr[j + 1] += t.ptr[j];
    t.idx.resize(a.idx.size());
    t.val.resize(a.val.size());
    std::vector<int> at(t.ptr.begin(), t.ptr.end() - 1);
    for (int i = 0; i < a.rows; ++i)
        for (int p = a.ptr[i]; p < a.ptr[i + 1]; ++p) {
            const int q = at[a.idx[p]]++;
            t.idx[q] = i;
            t.val[q] = a.val[p];
        }
    return t;
}

int LP::num_int() const { return static_cast<int>(std::count(is_int.begin(), is_int.end(), 1)); }

// ---------------------------------------------------------------------------------------------- reader
namespace {

std::string load_file(const std::string& path) {
    const bool gz = path.size() > 3 && path.compare(path.size() - 3, 3, ".gz") == 0;
    std::string out;
    if (gz) {
        const std::string cmd = "gzip -dc '" + path + "'";
        FILE* f = popen(cmd.c_str(), "r");
        if (!f) throw std::runtime_error("cannot run gzip on " + path);
        char buf[1 << 16];
        size_t k;
        while ((k = fread(buf, 1, sizeof(buf), f)) > 0) out.append(buf, k);
        pclose(f);
        if (out.empty()) throw std::runtime_error("empty or unreadable file " + path);
        return out;
    }
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) throw std::runtime_error("cannot open " + path);
    std::fseek(f, 0, SEEK_END);
    const long len = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    out.resize(len > 0 ? (size_t)len : 0);
    if (len > 0 && std::fread(&out[0], 1, (size_t)len, f) != (size_t)len) {
        std::fclose(f);
        throw std::runtime_error("short read on " + path);
    }
    std::fclose(f);
    return out;
}

// Name -> index for the reader: open addressing over (hash, index) slots, power-of-two capacity, linear
// probing; names are views into the file buffer (alive for the whole parse), compared only on a hash match.
// A node-based map chases a pointer per probe; at millions of COLUMNS entries those misses dominated the parse.
class NameIndex {
  public:
    static uint64_t hash(std::string_view s) {
        uint64_t h = 1469598103934665603ull;  // FNV-1a, 64-bit
        for (const char ch : s) h = (h ^ (unsigned char)ch) * 1099511628211ull;
        return h ^ (h >> 29);
    }
    // the index stored for `s`, or -1
    int find(std::string_view s) const {
        if (keys_.empty()) return -1;
        const uint64_t h = hash(s);
        for (size_t i = h & mask_;; i = (i + 1) & mask_) {
            if (slot_[i].idx < 0) return -1;
            if (slot_[i].h == h && keys_[(size_t)slot_[i].key] == s) return slot_[i].idx;
        }
    }
    // room for `n` names without rehashing
    void reserve(size_t n) {
        while (2 * n > slot_.size()) grow();
        keys_.reserve(n);
    }
    // map `s` to `v` (replacing an existing mapping, as map[s] = v does)
    void assign(std::string_view s, int v) {
        if (2 * (keys_.size() + 1) > slot_.size()) grow();
        const uint64_t h = hash(s);
        for (size_t i = h & mask_;; i = (i + 1) & mask_) {
            if (slot_[i].idx < 0) {
                slot_[i] = {h, v, (int)keys_.size()};
                keys_.push_back(s);
                return;
            }
            if (slot_[i].h == h && keys_[(size_t)slot_[i].key] == s) {
                slot_[i].idx = v;
                return;
            }
        }
    }

  private:
    struct Slot {
        uint64_t h;
        int idx;  // -1: empty
        int key;
    };
    void grow() {
        std::vector<Slot> old;
        old.swap(slot_);
        slot_.assign(std::max<size_t>(64, 2 * old.size()), Slot{0, -1, 0});
        mask_ = slot_.size() - 1;
        for (const Slot& o : old) {
            if (o.idx < 0) continue;
            size_t i = o.h & mask_;
            while (slot_[i].idx >= 0) i = (i + 1) & mask_;
            slot_[i] = o;
        }
    }
    std::vector<Slot> slot_;
    std::vector<std::string_view> keys_;
    size_t mask_ = 0;
};

// The C locale's isspace set (space, \t, \n, \v, \f, \r), inline.
inline bool is_ws(char ch) { return ch == ' ' || (ch >= '\t' && ch <= '\r'); }

// Whitespace tokens of one line as (pointer, length) views into the file buffer.
struct Tok {
    const char* p;
    size_t n;
    bool is(const char* s) const { return std::strlen(s) == n && std::memcmp(p, s, n) == 0; }
    std::string str() const { return std::string(p, n); }
};

// Plain decimals ([+-]digits[.digits][(e|E)[+-]digits]) with at most 19 significant digits, a mantissa below
// 2^53 and a decimal exponent within [-22, 22]: mantissa and power of ten are both exact doubles, so one IEEE
// multiply or divide gives the correctly rounded value -- the same double strtod returns (Clinger's fast
// path).  Anything else returns false and goes to strtod.
bool fast_decimal(const char* p, size_t n, double& out) {
    static const double kPow10[23] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,  1e8,  1e9,  1e10, 1e11,
                                      1e12, 1e13, 1e14, 1e15, 1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22};
    size_t i = 0;
    bool neg = false;
    if (i < n && (p[i] == '+' || p[i] == '-')) neg = p[i++] == '-';
    uint64_t mant = 0;
    int digits = 0, frac = 0;
    bool any = false;
    for (; i < n && p[i] >= '0' && p[i] <= '9'; ++i, any = true) {
        if (mant == 0 && p[i] == '0') continue;  // leading zeros are not significant
        if (++digits > 19) return false;
        mant = mant * 10 + (uint64_t)(p[i] - '0');
    }
    if (i < n && p[i] == '.') {
        for (++i; i < n && p[i] >= '0' && p[i] <= '9'; ++i, any = true) {
            ++frac;
            if (mant == 0 && p[i] == '0') continue;
            if (++digits > 19) return false;
            mant = mant * 10 + (uint64_t)(p[i] - '0');
        }
    }
    if (!any) return false;
    int e10 = 0;
    if (i < n && (p[i] == 'e' || p[i] == 'E')) {
        ++i;
        bool eneg = false;
        if (i < n && (p[i] == '+' || p[i] == '-')) eneg = p[i++] == '-';
        if (i >= n || p[i] < '0' || p[i] > '9') return false;
        for (; i < n && p[i] >= '0' && p[i] <= '9'; ++i) {
            e10 = e10 * 10 + (p[i] - '0');
            if (e10 > 1000) return false;
        }
        if (eneg) e10 = -e10;
    }
    if (i != n) return false;
    e10 -= frac;
    if (mant > (uint64_t(1) << 53) || e10 < -22 || e10 > 22) return false;
    double v = (double)mant;
    v = e10 >= 0 ? v * kPow10[e10] : v / kPow10[-e10];
    out = neg ? -v : v;
    return true;
}

double to_num(const Tok& t, int line) {
    double v;
    if (fast_decimal(t.p, t.n, v)) return v;
    char tmp[128];
    const size_t k = std::min(t.n, sizeof(tmp) - 1);
    std::memcpy(tmp, t.p, k);
    tmp[k] = 0;
    char* end = nullptr;
    v = std::strtod(tmp, &end);
    if (end == tmp) throw std::runtime_error("MPS line " + std::to_string(line) + ": bad number '" + t.str() + "'");
    return v;
}

}  // namespace

LP read_mps(const std::string& path) {
    const std::string text = load_file(path);
    enum Sec { None, Name, ObjSense, Rows, Columns, Rhs, Ranges, Bounds, QuadObj, QMatrix };
    LP lp;
    // names are views into `text`, which outlives the parse
    NameIndex row_of, col_of;
    std::unordered_set<std::string_view> extra_n;
    std::string_view obj_name, last_col;
    std::vector<char> rtype, lower_set;
    std::vector<double> rhs, range;
    std::vector<int> tr, tc, qr, qc;
    std::vector<double> tv, qv;
    bool int_mode = false;
    int neg_up = 0, cur = -1, line_no = 0;
    Sec sec = None;
    std::vector<Tok> tok;
    auto view = [](const Tok& t) { return std::string_view(t.p, t.n); };
    auto col_index = [&](const Tok& t) {
        const int j = col_of.find(view(t));
        if (j < 0) throw std::runtime_error("MPS: unknown column '" + t.str() + "' at line " + std::to_string(line_no));
        return j;
    };
    auto row_index = [&](const Tok& t) -> int {  // -1 for a row of an ignored extra N row
        const std::string_view s = view(t);
        const int r = row_of.find(s);
        if (r >= 0) return r;
        if (extra_n.count(s)) return -1;
        throw std::runtime_error("MPS: unknown row '" + std::string(s) + "' at line " + std::to_string(line_no));
    };

    const char* p = text.data();
    const char* const end = p + text.size();

    // The COLUMNS section, parsed in parallel: it is cut at line starts into chunks that threads tokenise and
    // look up independently (the row index is read-only by then), each chunk recording its column runs,
    // MARKER lines and entries in file order; one sequential pass then numbers the columns, applies the
    // integer markers and adds objective entries in line order, and the triplets are copied out in file order
    // -- so the result, and the first error reported, are the sequential loop's.
    auto parse_columns = [&]() {
        const int64_t line0 = line_no;  // the COLUMNS header line
        const char* const s0 = p;
        // the section ends at the next header line (first character not whitespace, not a comment)
        const char* s1 = p;
        while (s1 < end) {
            const char* eol = static_cast<const char*>(std::memchr(s1, '\n', (size_t)(end - s1)));
            if (!eol) eol = end;
            if (eol > s1 && !is_ws(*s1) && *s1 != '*' && !(eol - s1 == 1 && *s1 == '\r')) break;
            s1 = eol + (eol < end ? 1 : 0);
        }
        struct Item {
            std::string_view name;  // empty: a MARKER line
            int marker;             // +1 INTORG, -1 INTEND, 0 other marker
            size_t e0, e1;          // entries [e0, e1) of the chunk
            int col;
        };
        struct Err {
            int64_t line = -1;  // chunk-local line index of the first error, -1: none
            bool bad_number = false;
            std::string what;  // the offending token
        };
        struct Chunk {
            const char *a, *b;
            std::vector<Item> items;
            std::vector<int> row;  // kObjRow: an objective entry
            std::vector<double> val;
            std::vector<size_t> obj;  // positions of the objective entries, in order
            size_t ntrip = 0;         // entries that are not objective entries
            int64_t lines = 0;
            Err err;
        };
        constexpr int kObjRow = -2;
        const int nthreads = std::max(1, omp_get_max_threads());
        const size_t bytes = (size_t)(s1 - s0);
        const int nchunks = bytes < (1u << 20) ? 1 : nthreads * 4;
        std::vector<Chunk> ch((size_t)nchunks);
        {
            const char* a = s0;
            for (int k = 0; k < nchunks; ++k) {
                const char* b = k + 1 == nchunks ? s1 : s0 + bytes * (size_t)(k + 1) / (size_t)nchunks;
                if (b < a) b = a;
                if (k + 1 < nchunks) {  // a chunk ends after a newline
                    const char* nl = static_cast<const char*>(std::memchr(b, '\n', (size_t)(s1 - b)));
                    b = nl ? nl + 1 : s1;
                }
                ch[(size_t)k].a = a;
                ch[(size_t)k].b = b;
                a = b;
            }
        }
#pragma omp parallel for schedule(dynamic, 1)
        for (int k = 0; k < nchunks; ++k) {
            Chunk& c = ch[(size_t)k];
            c.row.reserve((size_t)(c.b - c.a) / 24);
            c.val.reserve((size_t)(c.b - c.a) / 24);
            std::vector<Tok> tk;
            std::string_view prev;
            for (const char* q = c.a; q < c.b;) {
                const char* eol = static_cast<const char*>(std::memchr(q, '\n', (size_t)(c.b - q)));
                if (!eol) eol = c.b;
                const char* line = q;
                size_t len = (size_t)(eol - q);
                q = eol + (eol < c.b ? 1 : 0);
                const int64_t ln = c.lines++;
                if (c.err.line >= 0) continue;
                if (len && line[len - 1] == '\r') --len;
                if (len == 0 || line[0] == '*') continue;
                tk.clear();
                for (size_t i = 0; i < len;) {
                    while (i < len && is_ws(line[i])) ++i;
                    if (i >= len) break;
                    size_t j = i;
                    while (j < len && !is_ws(line[j])) ++j;
                    tk.push_back({line + i, j - i});
                    i = j;
                }
                if (tk.empty()) continue;
                if (tk.size() >= 3 && tk[1].is("'MARKER'")) {
                    c.items.push_back({std::string_view(), tk[2].is("'INTORG'") ? 1 : tk[2].is("'INTEND'") ? -1 : 0,
                                       c.row.size(), c.row.size(), -1});
                    continue;
                }
                const std::string_view name(tk[0].p, tk[0].n);
                if (c.items.empty() || c.items.back().name.empty() || c.items.back().name != name)
                    c.items.push_back({name, 0, c.row.size(), c.row.size(), -1});
                for (size_t t = 1; t + 1 < tk.size(); t += 2) {
                    double v;
                    if (!fast_decimal(tk[t + 1].p, tk[t + 1].n, v)) {
                        char tmp[128];
                        const size_t n = std::min(tk[t + 1].n, sizeof(tmp) - 1);
                        std::memcpy(tmp, tk[t + 1].p, n);
                        tmp[n] = 0;
                        char* e = nullptr;
                        v = std::strtod(tmp, &e);
                        if (e == tmp) {
                            c.err = {ln, true, tk[t + 1].str()};
                            break;
                        }
                    }
                    const std::string_view rn(tk[t].p, tk[t].n);
                    int r;
                    if (rn == obj_name) r = kObjRow;
                    else if ((r = row_of.find(rn)) < 0) {
                        if (!extra_n.count(rn)) {
                            c.err = {ln, false, std::string(rn)};
                            break;
                        }
                        continue;  // an extra N row: ignored
                    }
                    if (r == kObjRow) c.obj.push_back(c.row.size());
                    else ++c.ntrip;
                    c.row.push_back(r);
                    c.val.push_back(v);
                }
                c.items.back().e1 = c.row.size();
            }
        }
        // sequential: column numbering, markers, objective sums, first error -- in file order
        int64_t lbase = line0;
        size_t ntrip = 0, nitems = 0;
        for (const Chunk& c : ch) nitems += c.items.size();
        col_of.reserve((size_t)lp.n + nitems);
        lp.col_names.reserve((size_t)lp.n + nitems);
        for (Chunk& c : ch) {
            size_t ob = 0;  // next objective entry of the chunk
            for (Item& it : c.items) {
                if (it.name.empty()) {
                    if (it.marker > 0) int_mode = true;
                    else if (it.marker < 0) int_mode = false;
                    continue;
                }
                if (it.name != last_col || cur < 0) {
                    last_col = it.name;
                    cur = lp.n++;
                    col_of.assign(last_col, cur);
                    lp.col_names.push_back(std::string(last_col));
                    lp.c.push_back(0.0);
                    lp.l.push_back(0.0);
                    lp.u.push_back(kInf);
                    lp.is_int.push_back(int_mode ? 1 : 0);
                    lower_set.push_back(0);
                }
                it.col = cur;
                for (; ob < c.obj.size() && c.obj[ob] < it.e1; ++ob) lp.c[(size_t)cur] += c.val[c.obj[ob]];
            }
            ntrip += c.ntrip;
            if (c.err.line >= 0) {  // the sequential loop's messages
                const std::string at = std::to_string(lbase + 1 + c.err.line);
                if (c.err.bad_number) throw std::runtime_error("MPS line " + at + ": bad number '" + c.err.what + "'");
                throw std::runtime_error("MPS: unknown row '" + c.err.what + "' at line " + at);
            }
            lbase += c.lines;
        }
        // triplets in file order
        std::vector<size_t> off(ch.size() + 1, 0);
        for (size_t k = 0; k < ch.size(); ++k) off[k + 1] = off[k] + ch[k].ntrip;
        const size_t t0 = tr.size();
        tr.resize(t0 + ntrip);
        tc.resize(t0 + ntrip);
        tv.resize(t0 + ntrip);
#pragma omp parallel for schedule(dynamic, 1)
        for (int k = 0; k < nchunks; ++k) {
            const Chunk& c = ch[(size_t)k];
            size_t o = t0 + off[(size_t)k];
            for (const Item& it : c.items)
                for (size_t e = it.e0; e < it.e1; ++e)
                    if (c.row[e] != kObjRow) {
                        tr[o] = c.row[e];
                        tc[o] = it.col;
                        tv[o] = c.val[e];
                        ++o;
                    }
        }
        p = s1;
        line_no = (int)lbase;
    };

    while (p < end) {
        const char* eol = static_cast<const char*>(std::memchr(p, '\n', (size_t)(end - p)));
        if (!eol) eol = end;
        const char* line = p;
        size_t len = (size_t)(eol - p);
        p = eol + (eol < end ? 1 : 0);
        ++line_no;
        if (len && line[len - 1] == '\r') --len;
        if (len == 0 || line[0] == '*') continue;
        tok.clear();
        for (size_t i = 0; i < len;) {
            while (i < len && is_ws(line[i])) ++i;
            if (i >= len) break;
            size_t j = i;
            while (j < len && !is_ws(line[j])) ++j;
            tok.push_back({line + i, j - i});
            i = j;
        }
        if (tok.empty()) continue;
        if (!is_ws(line[0])) {  // section header
            const Tok& h = tok[0];
            if (h.is("NAME")) { sec = Name; if (tok.size() > 1) lp.name = tok[1].str(); }
            else if (h.is("OBJSENSE")) {
                sec = ObjSense;
                if (tok.size() > 1) lp.maximize = tok[1].is("MAX") || tok[1].is("MAXIMIZE");
            }
            else if (h.is("ROWS")) sec = Rows;
            else if (h.is("COLUMNS")) {
                sec = Columns;
                parse_columns();  // the whole section; p and line_no end at the next header line
            }
            else if (h.is("RHS")) sec = Rhs;
            else if (h.is("RANGES")) sec = Ranges;
            else if (h.is("BOUNDS")) sec = Bounds;
            else if (h.is("QUADOBJ")) sec = QuadObj;
            else if (h.is("QMATRIX") || h.is("QSECTION")) sec = QMatrix;
            else if (h.is("ENDATA")) break;
            else if (h.is("OBJSENCE")) sec = ObjSense;
            else throw std::runtime_error("MPS: unknown section '" + h.str() + "' at line " + std::to_string(line_no));
            continue;
        }
        switch (sec) {
            case ObjSense:
                lp.maximize = tok[0].is("MAX") || tok[0].is("MAXIMIZE");
                break;
            case Rows: {
                if (tok.size() < 2) throw std::runtime_error("MPS: bad ROWS line " + std::to_string(line_no));
                const char t = (char)std::toupper((unsigned char)tok[0].p[0]);
                if (t == 'N') {
                    if (obj_name.empty()) obj_name = view(tok[1]); else extra_n.insert(view(tok[1]));
                } else {
                    row_of.assign(view(tok[1]), lp.m++);
                    lp.row_names.push_back(tok[1].str());
                    rtype.push_back(t);
                    rhs.push_back(0.0);
                }
                break;
            }
            case Rhs:
            case Ranges: {
                const size_t first = (tok.size() % 2 == 1) ? 1 : 0;  // optional set name
                for (size_t k = first; k + 1 < tok.size(); k += 2) {
                    const double v = to_num(tok[k + 1], line_no);
                    if (view(tok[k]) == obj_name) {
                        if (sec == Rhs) lp.obj_const = -v;  // objective RHS is minus the constant
                        continue;
                    }
                    const int r = row_index(tok[k]);
                    if (r < 0) continue;
                    if (sec == Rhs) rhs[r] = v;
                    else {
                        if (range.empty()) range.assign(lp.m, std::nan(""));
                        range[r] = v;
                    }
                }
                break;
            }
            case Bounds: {
                const Tok& t = tok[0];
                const bool valued = t.is("UP") || t.is("LO") || t.is("FX") || t.is("LI") || t.is("UI");
                const Tok* cname;
                double v = 0.0;
                if (valued) {
                    if (tok.size() >= 4) { cname = &tok[2]; v = to_num(tok[3], line_no); }
                    else if (tok.size() == 3) { cname = &tok[1]; v = to_num(tok[2], line_no); }
                    else throw std::runtime_error("MPS: bad BOUNDS line " + std::to_string(line_no));
                } else {
                    cname = tok.size() >= 3 ? &tok[2] : &tok[1];
                }
                const int j = col_index(*cname);
                if (t.is("UP") || t.is("UI")) {
                    lp.u[j] = v;
                    if (v < 0 && lp.l[j] == 0.0 && !lower_set[j]) { lp.l[j] = -kInf; ++neg_up; }
                    if (t.is("UI")) lp.is_int[j] = 1;
                } else if (t.is("LO") || t.is("LI")) {
                    lp.l[j] = v;
                    lower_set[j] = 1;
                    if (t.is("LI")) lp.is_int[j] = 1;
                } else if (t.is("FX")) { lp.l[j] = v; lp.u[j] = v; lower_set[j] = 1; }
                else if (t.is("FR")) { lp.l[j] = -kInf; lp.u[j] = kInf; lower_set[j] = 1; }
                else if (t.is("MI")) { lp.l[j] = -kInf; lower_set[j] = 1; }
                else if (t.is("PL")) { lp.u[j] = kInf; }
                else if (t.is("BV")) { lp.l[j] = 0; lp.u[j] = 1; lp.is_int[j] = 1; lower_set[j] = 1; }
                else throw std::runtime_error("MPS: unknown bound type '" + t.str() + "'");
                break;
            }
            case QuadObj:
            case QMatrix: {
                if (tok.size() < 3) throw std::runtime_error("MPS: bad quadratic line " + std::to_string(line_no));
                const int a = col_index(tok[0]), b = col_index(tok[1]);
                const double v = to_num(tok[2], line_no);
                qr.push_back(a); qc.push_back(b); qv.push_back(v);
                if (sec == QuadObj && a != b) { qr.push_back(b); qc.push_back(a); qv.push_back(v); }
                break;
            }
            default:
                break;
        }
    }
    lp.lo.assign(lp.m, 0.0);
    lp.hi.assign(lp.m, 0.0);
    for (int i = 0; i < lp.m; ++i) {
        const double b = rhs[i];
        switch (rtype[i]) {
            case 'E': lp.lo[i] = b; lp.hi[i] = b; break;
            case 'L': lp.lo[i] = -kInf; lp.hi[i] = b; break;
            case 'G': lp.lo[i] = b; lp.hi[i] = kInf; break;
            default: throw std::runtime_error("MPS: unknown row type for " + lp.row_names[i]);
        }
        if (!range.empty() && !std::isnan(range[i])) {
            const double r = range[i];
            if (rtype[i] == 'E') { if (r > 0) lp.hi[i] = b + r; else lp.lo[i] = b + r; }
            else if (rtype[i] == 'L') lp.lo[i] = b - std::fabs(r);
            else lp.hi[i] = b + std::fabs(r);
        }
    }
    lp.A = csr_from_triplets(lp.m, lp.n, tr, tc, tv);
    lp.Q = qv.empty() ? empty_csr(lp.n, lp.n) : csr_from_triplets(lp.n, lp.n, qr, qc, qv);
    if (neg_up)
        std::cerr << "[mps] warning: " << neg_up << " negative UP bounds with zero lower bound; lower set to -inf\n";
    return lp;
}

// ---------------------------------------------------------------------------------------------- writer
void write_mps(const LP& lp, const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) throw std::runtime_error("cannot write " + path);
    auto rname = [&](int i) { return i < (int)lp.row_names.size() ? lp.row_names[i] : "R" + std::to_string(i); };
    auto cname = [&](int j) { return j < (int)lp.col_names.size() ? lp.col_names[j] : "C" + std::to_string(j); };
    std::fprintf(f, "NAME %s\n", lp.name.empty() ? "GANIT" : lp.name.c_str());
    if (lp.maximize) std::fprintf(f, "OBJSENSE\n    MAX\n");
    std::fprintf(f, "ROWS\n N OBJ\n");
    std::vector<char> type(lp.m);
    std::vector<double> rhs(lp.m, 0.0), rng(lp.m, 0.0);
    std::vector<char> has_rng(lp.m, 0);
    for (int i = 0; i < lp.m; ++i) {
        const double lo = lp.lo[i], hi = lp.hi[i];
        if (lo == hi) { type[i] = 'E'; rhs[i] = lo; }
        else if (std::isinf(lo) && lo < 0) { type[i] = 'L'; rhs[i] = hi; }      // hi may be +inf: a free row
        else if (std::isinf(hi) && hi > 0) { type[i] = 'G'; rhs[i] = lo; }
        else { type[i] = 'G'; rhs[i] = lo; rng[i] = hi - lo; has_rng[i] = 1; }  // G row: hi = lo + |range|
        std::fprintf(f, " %c %s\n", type[i], rname(i).c_str());
    }
    std::fprintf(f, "COLUMNS\n");
    const Csr at = transpose(lp.A);
    for (int j = 0; j < lp.n; ++j) {
        const std::string cn = cname(j);
        // one objective entry per column, even if zero, so that every column is declared
        std::fprintf(f, "    %s OBJ %.17g\n", cn.c_str(), lp.c[j]);
        for (int p = at.ptr[j]; p < at.ptr[j + 1]; ++p)
            std::fprintf(f, "    %s %s %.17g\n", cn.c_str(), rname(at.idx[p]).c_str(), at.val[p]);
    }
    std::fprintf(f, "RHS\n");
    if (lp.obj_const != 0.0) std::fprintf(f, "    RHS OBJ %.17g\n", -lp.obj_const);
    for (int i = 0; i < lp.m; ++i)
        if (rhs[i] != 0.0) std::fprintf(f, "    RHS %s %.17g\n", rname(i).c_str(), rhs[i]);
    bool any_rng = false;
    for (int i = 0; i < lp.m; ++i) any_rng |= has_rng[i] != 0;
    if (any_rng) {
        std::fprintf(f, "RANGES\n");
        for (int i = 0; i < lp.m; ++i)
            if (has_rng[i]) std::fprintf(f, "    RNG %s %.17g\n", rname(i).c_str(), rng[i]);
    }
    std::fprintf(f, "BOUNDS\n");
    for (int j = 0; j < lp.n; ++j) {
        const std::string cn = cname(j);
        const double l = lp.l[j], u = lp.u[j];
        if (l == u) { std::fprintf(f, " FX BND %s %.17g\n", cn.c_str(), l); continue; }
        if (std::isinf(l) && std::isinf(u)) { std::fprintf(f, " FR BND %s\n", cn.c_str()); continue; }
        if (std::isinf(l)) std::fprintf(f, " MI BND %s\n", cn.c_str());
        else if (l != 0.0 || u < 0.0) std::fprintf(f, " LO BND %s %.17g\n", cn.c_str(), l);
        if (!std::isinf(u)) std::fprintf(f, " UP BND %s %.17g\n", cn.c_str(), u);
    }
    if (lp.has_q()) {
        std::fprintf(f, "QUADOBJ\n");
        for (int i = 0; i < lp.n; ++i)
            for (int p = lp.Q.ptr[i]; p < lp.Q.ptr[i + 1]; ++p)
                if (lp.Q.idx[p] <= i)
                    std::fprintf(f, "    %s %s %.17g\n", cname(i).c_str(), cname(lp.Q.idx[p]).c_str(), lp.Q.val[p]);
    }
    std::fprintf(f, "ENDATA\n");
    std::fclose(f);
}

}  // namespace ganit
