/*
 * Inputs-file reader for the FerroX hot path: a ParmParse-compatible subset.
 *
 * Replaces amrex::ParmParse + InitializeFerroXNamespace (Source/FerroX.cpp:273-572),
 * c_GeometryProperties::ParseBasicDomainInput (Source/Input/GeometryProperties/
 * GeometryProperties.cpp:87-165) and the boundary-token parsing of
 * Source/Input/BoundaryConditions/BoundaryConditions.cpp:52-214.
 *
 * Format: "key = v1 v2 ...", '#' starts a comment, a later definition of a key
 * replaces an earlier one, double-quoted values may span lines.  Numeric
 * values may be arithmetic expressions (+ - * / ^ **, parentheses, e-notation,
 * sqrt/exp/log/sin/cos/abs, pi) over names defined as my_constants.<name>,
 * following the parser the reference uses for getWithParser keys
 * (Source/Utils/SelectWarpXUtils/WarpXUtil.cpp:101-315).  Unknown keys are
 * accepted and ignored, as ParmParse does.
 */
#include "../../include/ferrox_mi.h"

#include <cctype>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

namespace {

const double UNSET = NAN;
const int IUNSET = INT_MIN;

void set_err(char* err, int errlen, const std::string& m) {
    if (err && errlen > 0) {
        std::snprintf(err, (size_t)errlen, "%s", m.c_str());
    }
}

/* ---------------- expression evaluator (recursive descent) ---------------- */
struct Expr {
    const std::string& s;
    size_t p = 0;
    const std::map<std::string, double>& consts;
    bool ok = true;
    std::string why;
    Expr(const std::string& str, const std::map<std::string, double>& c) : s(str), consts(c) {}
    void ws() { while (p < s.size() && std::isspace((unsigned char)s[p])) ++p; }
    double parse() {
        double v = sum();
        ws();
        if (p != s.size()) { ok = false; why = "trailing characters in '" + s + "'"; }
        return v;
    }
    double sum() {
        double v = prod();
        for (;;) {
            ws();
            if (p < s.size() && (s[p] == '+' || s[p] == '-')) { char o = s[p++]; double r = prod(); v = (o == '+') ? v + r : v - r; }
            else return v;
        }
    }
    double prod() {
        double v = unary();
        for (;;) {
            ws();
            if (p + 1 < s.size() && s[p] == '*' && s[p + 1] == '*') return v;   /* handled in power */
            if (p < s.size() && (s[p] == '*' || s[p] == '/')) { char o = s[p++]; double r = unary(); v = (o == '*') ? v * r : v / r; }
            else return v;
        }
    }
    double unary() {
        ws();
        if (p < s.size() && s[p] == '-') { ++p; return -unary(); }
        if (p < s.size() && s[p] == '+') { ++p; return unary(); }
        return power();
    }
    double power() {
        double b = atom();
        ws();
        if (p < s.size() && s[p] == '^') { ++p; return std::pow(b, unary()); }
        if (p + 1 < s.size() && s[p] == '*' && s[p + 1] == '*') { p += 2; return std::pow(b, unary()); }
        return b;
    }
    double atom() {
        ws();
        if (p >= s.size()) { ok = false; why = "unexpected end of '" + s + "'"; return 0; }
        if (s[p] == '(') { ++p; double v = sum(); ws(); if (p < s.size() && s[p] == ')') ++p; else { ok = false; why = "missing ')'"; } return v; }
        if (std::isdigit((unsigned char)s[p]) || s[p] == '.') {
            const char* b = s.c_str() + p;
            char* e = nullptr;
            double v = std::strtod(b, &e);
            if (e == b) { ok = false; why = "bad number"; return 0; }
            p += (size_t)(e - b);
            return v;
        }
        if (std::isalpha((unsigned char)s[p]) || s[p] == '_') {
            size_t b = p;
            while (p < s.size() && (std::isalnum((unsigned char)s[p]) || s[p] == '_' || s[p] == '.')) ++p;
            std::string name = s.substr(b, p - b);
            ws();
            if (p < s.size() && s[p] == '(') {
                ++p; double a = sum(); ws();
                if (p < s.size() && s[p] == ')') ++p; else { ok = false; why = "missing ')'"; }
                if (name == "sqrt") return std::sqrt(a);
                if (name == "exp") return std::exp(a);
                if (name == "log") return std::log(a);
                if (name == "sin") return std::sin(a);
                if (name == "cos") return std::cos(a);
                if (name == "abs" || name == "fabs") return std::fabs(a);
                ok = false; why = "unknown function " + name; return 0;
            }
            if (name == "pi") return 3.14159265358979323846;
            auto it = consts.find(name);
            if (it != consts.end()) return it->second;
            ok = false; why = "unknown name '" + name + "'"; return 0;
        }
        ok = false; why = std::string("unexpected '") + s[p] + "'";
        return 0;
    }
};

/* ------------------ functions of the cell centre (x, y, z) ---------------- */
/* Compiled once, evaluated per cell: the operator subset of the AMReX parser
 * the reference builds its geometry functions with (makeParser,
 * Utils/SelectWarpXUtils/WarpXUtil.cpp:101-168).  Precedence, lowest first:
 * or, and, comparisons, + -, * /, unary +-, ^ ** (right associative). */
struct FnNode {
    int op;          /* 0 const, 1 x, 2 y, 3 z, 10+ binary, 40+ unary fn, 60+ binary fn, 70 if */
    double v;
    int a, b, c;
};
struct Fn {
    std::vector<FnNode> nodes;
    int root = -1;
    double eval(int i, double x, double y, double z) const {
        const FnNode& n = nodes[i];
        switch (n.op) {
        case 0: return n.v;
        case 1: return x;
        case 2: return y;
        case 3: return z;
        case 10: return eval(n.a, x, y, z) + eval(n.b, x, y, z);
        case 11: return eval(n.a, x, y, z) - eval(n.b, x, y, z);
        case 12: return eval(n.a, x, y, z) * eval(n.b, x, y, z);
        case 13: return eval(n.a, x, y, z) / eval(n.b, x, y, z);
        case 14: return std::pow(eval(n.a, x, y, z), eval(n.b, x, y, z));
        case 20: return eval(n.a, x, y, z) < eval(n.b, x, y, z) ? 1.0 : 0.0;
        case 21: return eval(n.a, x, y, z) > eval(n.b, x, y, z) ? 1.0 : 0.0;
        case 22: return eval(n.a, x, y, z) <= eval(n.b, x, y, z) ? 1.0 : 0.0;
        case 23: return eval(n.a, x, y, z) >= eval(n.b, x, y, z) ? 1.0 : 0.0;
        case 24: return eval(n.a, x, y, z) == eval(n.b, x, y, z) ? 1.0 : 0.0;
        case 25: return eval(n.a, x, y, z) != eval(n.b, x, y, z) ? 1.0 : 0.0;
        case 30: return (eval(n.a, x, y, z) != 0.0 && eval(n.b, x, y, z) != 0.0) ? 1.0 : 0.0;
        case 31: return (eval(n.a, x, y, z) != 0.0 || eval(n.b, x, y, z) != 0.0) ? 1.0 : 0.0;
        case 39: return -eval(n.a, x, y, z);
        case 40: return std::sqrt(eval(n.a, x, y, z));
        case 41: return std::exp(eval(n.a, x, y, z));
        case 42: return std::log(eval(n.a, x, y, z));
        case 43: return std::log10(eval(n.a, x, y, z));
        case 44: return std::sin(eval(n.a, x, y, z));
        case 45: return std::cos(eval(n.a, x, y, z));
        case 46: return std::tan(eval(n.a, x, y, z));
        case 47: return std::asin(eval(n.a, x, y, z));
        case 48: return std::acos(eval(n.a, x, y, z));
        case 49: return std::atan(eval(n.a, x, y, z));
        case 50: return std::sinh(eval(n.a, x, y, z));
        case 51: return std::cosh(eval(n.a, x, y, z));
        case 52: return std::tanh(eval(n.a, x, y, z));
        case 53: return std::fabs(eval(n.a, x, y, z));
        case 54: return std::floor(eval(n.a, x, y, z));
        case 55: return std::ceil(eval(n.a, x, y, z));
        case 60: return std::pow(eval(n.a, x, y, z), eval(n.b, x, y, z));
        case 61: return std::atan2(eval(n.a, x, y, z), eval(n.b, x, y, z));
        case 62: return std::fmin(eval(n.a, x, y, z), eval(n.b, x, y, z));
        case 63: return std::fmax(eval(n.a, x, y, z), eval(n.b, x, y, z));
        case 64: return std::fmod(eval(n.a, x, y, z), eval(n.b, x, y, z));
        case 70: return eval(n.a, x, y, z) != 0.0 ? eval(n.b, x, y, z) : eval(n.c, x, y, z);
        }
        return 0.0;
    }
    double operator()(double x, double y, double z) const { return eval(root, x, y, z); }
};

struct FnParser {
    const std::string& s;
    const std::map<std::string, double>& consts;
    Fn& f;
    size_t p = 0;
    bool ok = true;
    std::string why;
    FnParser(const std::string& str, const std::map<std::string, double>& c, Fn& out) : s(str), consts(c), f(out) {}
    int node(int op, double v = 0, int a = -1, int b = -1, int c = -1) { f.nodes.push_back({op, v, a, b, c}); return (int)f.nodes.size() - 1; }
    void ws() { while (p < s.size() && std::isspace((unsigned char)s[p])) ++p; }
    bool lit(const char* t) {
        ws();
        const size_t n = std::strlen(t);
        if (s.compare(p, n, t) == 0) {
            if (std::isalpha((unsigned char)t[0]) && p + n < s.size() && (std::isalnum((unsigned char)s[p + n]) || s[p + n] == '_')) return false;
            p += n;
            return true;
        }
        return false;
    }
    int fail(const std::string& m) { if (ok) { ok = false; why = m; } return node(0, 0.0); }
    int parse() { int r = orx(); ws(); if (p != s.size()) fail("trailing characters at '" + s.substr(p, 16) + "'"); f.root = r; return r; }
    int orx() { int a = andx(); while (ok && (lit("||") || lit("or"))) a = node(31, 0, a, andx()); return a; }
    int andx() { int a = cmp(); while (ok && (lit("&&") || lit("and"))) a = node(30, 0, a, cmp()); return a; }
    int cmp() {
        int a = sum();
        if (lit("<=")) return node(22, 0, a, sum());
        if (lit(">=")) return node(23, 0, a, sum());
        if (lit("==")) return node(24, 0, a, sum());
        if (lit("!=")) return node(25, 0, a, sum());
        if (lit("<")) return node(20, 0, a, sum());
        if (lit(">")) return node(21, 0, a, sum());
        return a;
    }
    int sum() {
        int a = prod();
        for (;;) {
            ws();
            if (p < s.size() && s[p] == '+') { ++p; a = node(10, 0, a, prod()); }
            else if (p < s.size() && s[p] == '-') { ++p; a = node(11, 0, a, prod()); }
            else return a;
        }
    }
    int prod() {
        int a = unary();
        for (;;) {
            ws();
            if (p + 1 < s.size() && s[p] == '*' && s[p + 1] == '*') return a;
            if (p < s.size() && s[p] == '*') { ++p; a = node(12, 0, a, unary()); }
            else if (p < s.size() && s[p] == '/') { ++p; a = node(13, 0, a, unary()); }
            else return a;
        }
    }
    int unary() {
        ws();
        if (p < s.size() && s[p] == '-') { ++p; return node(39, 0, unary()); }
        if (p < s.size() && s[p] == '+') { ++p; return unary(); }
        return power();
    }
    int power() {
        int b = atom();
        ws();
        if (p < s.size() && s[p] == '^') { ++p; return node(14, 0, b, unary()); }
        if (p + 1 < s.size() && s[p] == '*' && s[p + 1] == '*') { p += 2; return node(14, 0, b, unary()); }
        return b;
    }
    int atom() {
        ws();
        if (p >= s.size()) return fail("unexpected end of expression");
        if (s[p] == '(') { ++p; int v = orx(); ws(); if (p < s.size() && s[p] == ')') ++p; else fail("missing ')'"); return v; }
        if (std::isdigit((unsigned char)s[p]) || s[p] == '.') {
            const char* b = s.c_str() + p;
            char* e = nullptr;
            double v = std::strtod(b, &e);
            if (e == b) return fail("bad number");
            p += (size_t)(e - b);
            return node(0, v);
        }
        if (std::isalpha((unsigned char)s[p]) || s[p] == '_') {
            size_t b = p;
            while (p < s.size() && (std::isalnum((unsigned char)s[p]) || s[p] == '_' || s[p] == '.')) ++p;
            const std::string name = s.substr(b, p - b);
            ws();
            if (p < s.size() && s[p] == '(') {
                ++p;
                std::vector<int> args;
                ws();
                if (p < s.size() && s[p] == ')') ++p;
                else {
                    for (;;) {
                        args.push_back(orx());
                        ws();
                        if (p < s.size() && s[p] == ',') { ++p; continue; }
                        if (p < s.size() && s[p] == ')') { ++p; break; }
                        return fail("missing ')' after arguments of " + name);
                    }
                }
                static const char* f1[] = {"sqrt", "exp", "log", "log10", "sin", "cos", "tan", "asin", "acos", "atan",
                                           "sinh", "cosh", "tanh", "abs", "floor", "ceil"};
                for (int q = 0; q < 16; ++q)
                    if (name == f1[q]) return args.size() == 1 ? node(40 + q, 0, args[0]) : fail(name + " takes 1 argument");
                if (name == "fabs") return args.size() == 1 ? node(53, 0, args[0]) : fail("fabs takes 1 argument");
                static const char* f2[] = {"pow", "atan2", "min", "max", "fmod"};
                for (int q = 0; q < 5; ++q)
                    if (name == f2[q]) return args.size() == 2 ? node(60 + q, 0, args[0], args[1]) : fail(name + " takes 2 arguments");
                if (name == "if") return args.size() == 3 ? node(70, 0, args[0], args[1], args[2]) : fail("if takes 3 arguments");
                return fail("unknown function " + name);
            }
            if (name == "x") return node(1);
            if (name == "y") return node(2);
            if (name == "z") return node(3);
            if (name == "pi") return node(0, 3.14159265358979323846);
            auto it = consts.find(name);
            if (it != consts.end()) return node(0, it->second);
            return fail("unknown name '" + name + "'");
        }
        return fail(std::string("unexpected '") + s[p] + "'");
    }
};

bool compile_fn(const std::string& s, const std::map<std::string, double>& consts, Fn& f, std::string& why) {
    f.nodes.clear();
    FnParser P(s, consts, f);
    P.parse();
    if (!P.ok) { why = P.why; return false; }
    return true;
}

/* --------------------------- key/value table ------------------------------ */
struct Table {
    std::map<std::string, std::vector<std::string>> kv;   /* last definition wins */
};

/* Split a value string into whitespace-separated tokens, keeping "dir(0.0)"
 * and quoted strings whole. */
std::vector<std::string> tokens(const std::string& v) {
    std::vector<std::string> out;
    size_t i = 0;
    while (i < v.size()) {
        while (i < v.size() && std::isspace((unsigned char)v[i])) ++i;
        if (i >= v.size()) break;
        if (v[i] == '"') {
            size_t j = v.find('"', i + 1);
            if (j == std::string::npos) j = v.size();
            out.push_back(v.substr(i + 1, j - i - 1));
            i = j + 1;
            continue;
        }
        size_t j = i;
        int depth = 0;
        while (j < v.size() && (depth > 0 || !std::isspace((unsigned char)v[j]))) {
            if (v[j] == '(') ++depth;
            if (v[j] == ')') --depth;
            ++j;
        }
        out.push_back(v.substr(i, j - i));
        i = j;
    }
    return out;
}

bool parse_text(const std::string& text, Table& t, std::string& why) {
    /* strip comments outside quotes, join quoted multi-line values */
    std::string clean;
    bool inq = false;
    for (size_t i = 0; i < text.size(); ++i) {
        char c = text[i];
        if (c == '"') inq = !inq;
        if (c == '#' && !inq) { while (i < text.size() && text[i] != '\n') ++i; clean += '\n'; continue; }
        if (c == '\n' && inq) c = ' ';
        clean += c;
    }
    std::istringstream is(clean);
    std::string line;
    int ln = 0;
    while (std::getline(is, line)) {
        ++ln;
        size_t eq = line.find('=');
        std::string key = line.substr(0, eq == std::string::npos ? line.size() : eq);
        while (!key.empty() && std::isspace((unsigned char)key.back())) key.pop_back();
        size_t b = 0;
        while (b < key.size() && std::isspace((unsigned char)key[b])) ++b;
        key = key.substr(b);
        if (key.empty()) continue;
        if (eq == std::string::npos) { why = "line " + std::to_string(ln) + ": expected 'key = value'"; return false; }
        t.kv[key] = tokens(line.substr(eq + 1));
    }
    return true;
}

struct Applier {
    fx_inputs* in;
    std::map<std::string, double> consts;
    std::string why;

    bool num(const std::string& tok, double& v) {
        Expr e(tok, consts);
        v = e.parse();
        if (!e.ok) { why = e.why; return false; }
        return true;
    }
    bool dbl(const std::vector<std::string>& vals, double* dst, int n, const std::string& key) {
        if ((int)vals.size() < n) { why = key + ": expected " + std::to_string(n) + " value(s)"; return false; }
        for (int i = 0; i < n; ++i) if (!num(vals[i], dst[i])) { why = key + ": " + why; return false; }
        return true;
    }
    bool integer(const std::vector<std::string>& vals, int* dst, int n, const std::string& key) {
        double tmp[3];
        if (!dbl(vals, tmp, n, key)) return false;
        for (int i = 0; i < n; ++i) {
            if (tmp[i] != std::floor(tmp[i])) { why = key + ": expected an integer"; return false; }
            dst[i] = (int)tmp[i];
        }
        return true;
    }
    bool bc(const std::vector<std::string>& vals, int* dst, const std::string& key) {
        if (vals.size() < 3) { why = key + ": expected 3 tokens"; return false; }
        for (int d = 0; d < 3; ++d) {
            const std::string& s = vals[d];
            if (s == "per") dst[d] = FX_BC_PERIODIC;
            else if (s.rfind("dir", 0) == 0) dst[d] = FX_BC_DIRICHLET;   /* dir(value): value parsed but never applied (ElectrostaticSolver.cpp:574-596) */
            else if (s.rfind("neu", 0) == 0) dst[d] = FX_BC_NEUMANN;
            else { why = key + ": unknown boundary token '" + s + "'"; return false; }
        }
        return true;
    }

    /* a function of (x,y,z): the value text with my_constants replaced by
     * their values (checked by compiling it) */
    bool func(const std::vector<std::string>& vals, char* dst, size_t cap, const std::string& key) {
        std::string src;
        for (size_t q = 0; q < vals.size(); ++q) src += (q ? " " : "") + vals[q];
        std::string out;
        for (size_t i = 0; i < src.size();) {
            if (std::isalpha((unsigned char)src[i]) || src[i] == '_') {
                size_t b = i;
                while (i < src.size() && (std::isalnum((unsigned char)src[i]) || src[i] == '_' || src[i] == '.')) ++i;
                const std::string name = src.substr(b, i - b);
                auto it = consts.find(name);
                if (it != consts.end()) {
                    char buf[64];
                    std::snprintf(buf, sizeof buf, "(%.17g)", it->second);
                    out += buf;
                } else {
                    out += name;
                }
            } else if (std::isdigit((unsigned char)src[i]) || src[i] == '.') {
                /* a number, with its exponent, copied whole */
                const char* b = src.c_str() + i;
                char* e = nullptr;
                (void)std::strtod(b, &e);
                const size_t n = e > b ? (size_t)(e - b) : 1;
                out += src.substr(i, n);
                i += n;
            } else {
                out += src[i++];
            }
        }
        Fn f;
        std::string w;
        if (!compile_fn(out, {}, f, w)) { why = key + ": " + w; return false; }
        if (out.size() + 1 > cap) { why = key + ": function text longer than " + std::to_string(cap - 1) + " characters"; return false; }
        std::memcpy(dst, out.c_str(), out.size() + 1);
        return true;
    }

    bool apply(const std::string& key, const std::vector<std::string>& v) {
        fx_inputs& I = *in;
#define D1(name) if (key == #name) return dbl(v, &I.name, 1, key);
#define D3(name) if (key == #name) return dbl(v, I.name, 3, key);
#define I1(name) if (key == #name) return integer(v, &I.name, 1, key);
#define I3(name) if (key == #name) return integer(v, I.name, 3, key);
        if (key == "domain.n_cell") return integer(v, I.n_cell, 3, key);
        if (key == "domain.prob_lo") return dbl(v, I.prob_lo, 3, key);
        if (key == "domain.prob_hi") return dbl(v, I.prob_hi, 3, key);
        if (key == "domain.is_periodic") return integer(v, I.is_periodic, 3, key);
        if (key == "domain.max_grid_size") return integer(v, I.max_grid_size, 3, key);
        if (key == "domain.blocking_factor" || key == "domain.coord_sys") return true;
        if (key == "domain.embedded_boundary") {
            int eb = 0; if (!integer(v, &eb, 1, key)) return false;
            if (eb) { why = "domain.embedded_boundary = 1 is not supported (embedded boundaries are out of scope)"; return false; }
            return true;
        }
        if (key == "tphase_geom.tphase_geom_function(x,y,z)") return func(v, I.tphase_geom_function, sizeof(I.tphase_geom_function), key);
        if (key == "angle_alpha.alpha_function(x,y,z)") return func(v, I.angle_alpha_function, sizeof(I.angle_alpha_function), key);
        if (key == "angle_beta.beta_function(x,y,z)") return func(v, I.angle_beta_function, sizeof(I.angle_beta_function), key);
        if (key == "angle_theta.theta_function(x,y,z)") return func(v, I.angle_theta_function, sizeof(I.angle_theta_function), key);
        if (key == "boundary.lo") return bc(v, I.bc_lo, key);
        if (key == "boundary.hi") return bc(v, I.bc_hi, key);
        I1(TimeIntegratorOrder) I1(prob_type) I1(is_polarization_scalar) I1(nsteps) I1(plot_int)
        I1(plot_Phi) I1(plot_PoissonRHS) I1(plot_E) I1(plot_holes) I1(plot_electrons) I1(plot_charge) I1(plot_epsilon) I1(plot_mask) I1(plot_tphase) I1(plot_alpha) I1(plot_beta) I1(plot_theta) I1(plot_PhiDiff)
        D1(dt) D1(delta) I1(random_seed) I1(voltage_sweep) I1(inc_step) I1(num_Vapp_max)
        D1(Phi_Bc_lo) D1(Phi_Bc_hi) D1(Phi_Bc_inc) D1(Phi_Bc_hi_max) D1(phi_tolerance)
        I1(include_Landau) I1(include_Grad) I1(include_Elec)
        I3(P_BC_flag_lo) I3(P_BC_flag_hi) D1(lambda)
        D1(epsilon_0) D1(epsilonX_fe) D1(epsilonX_fe_tphase) D1(epsilonZ_fe) D1(epsilon_de) D1(epsilon_si)
        D1(alpha) D1(beta) D1(gamma) D1(alpha_12) D1(alpha_112) D1(alpha_123) D1(BigGamma)
        D1(g11) D1(g44) D1(g12) D1(g44_p)
        D3(FE_lo) D3(FE_hi) D3(DE_lo) D3(DE_hi) D3(SC_lo) D3(SC_hi) D3(Channel_lo) D3(Channel_hi) D3(Remnant_P)
        D1(acceptor_doping) D1(donor_doping) I1(use_Fermi_Dirac) I1(use_work_function) D1(metal_work_function)
        I1(Coordinate_Transformation) I1(use_Euler_angles)
        if (key == "use_sundials") {
            int u = 0; if (!integer(v, &u, 1, key)) return false;
            if (u) { why = "use_sundials = 1 is not supported (SUNDIALS integrators are out of scope)"; return false; }
            return true;
        }
        /* solver controls added by this build */
        D1(mg_tol_rel) I1(mg_max_iter) I1(mg_nu1) I1(mg_nu2) I1(mg_warm_start) I1(outer_max_iter)
        I1(phi_ghost_mode) I1(precision) I1(bias_ramp_steps) D1(bias_ramp_start) D1(bias_ramp_end) I1(verbose)
        I1(mg_fused_sweep) I1(mg_block_levels) I1(mg_stream_min_cells) D1(mg_coarsen_ratio) I1(mg_exact_tail)
        I1(mg_refine_nu1) I1(mg_refine_nu2) I1(poisson_direct) I1(poisson_direct_check) I1(tdgl_fused_rhs) I1(tdgl_stream) I1(poisson_pcg)
#undef D1
#undef D3
#undef I1
#undef I3
        return true;   /* unknown keys are ignored, as ParmParse does */
    }
};

bool apply_table(Table& t, fx_inputs* in, std::string& why) {
    Applier a{in, {}, {}};
    /* my_constants first (they may refer to each other in file order; a simple
     * fixed-point pass resolves forward references) */
    for (int pass = 0; pass < 4; ++pass)
        for (auto& kv : t.kv)
            if (kv.first.rfind("my_constants.", 0) == 0 && !kv.second.empty()) {
                double v;
                if (a.num(kv.second[0], v)) a.consts[kv.first.substr(13)] = v;
            }
    for (auto& kv : t.kv) {
        if (kv.first.rfind("my_constants.", 0) == 0) continue;
        if (!a.apply(kv.first, kv.second)) { why = a.why; return false; }
    }
    return true;
}

}  // namespace

extern "C" {

const char* fx_version(void) { return "ferrox-mi355x 0.1"; }

int fx_inputs_size(void) { return (int)sizeof(fx_inputs); }

int fx_inputs_default(fx_inputs* in) {
    if (!in) return FX_EINVAL;
    std::memset(in, 0, sizeof(*in));
    for (int d = 0; d < 3; ++d) {
        in->n_cell[d] = IUNSET;
        in->prob_lo[d] = UNSET; in->prob_hi[d] = UNSET;
        in->is_periodic[d] = 0;                       /* GeometryProperties.cpp:100 */
        in->max_grid_size[d] = 128;                   /* :98 */
        in->bc_lo[d] = -1; in->bc_hi[d] = -1;          /* from is_periodic when absent */
        in->P_BC_flag_lo[d] = 0; in->P_BC_flag_hi[d] = 0;
        in->FE_lo[d] = UNSET; in->FE_hi[d] = UNSET;
        in->DE_lo[d] = UNSET; in->DE_hi[d] = UNSET;    /* prob_lo-1 / prob_hi+1, FerroX.cpp:445-456 */
        in->SC_lo[d] = UNSET; in->SC_hi[d] = UNSET;
        in->Channel_lo[d] = UNSET; in->Channel_hi[d] = UNSET;
    }
    in->Remnant_P[0] = 0.0; in->Remnant_P[1] = 0.0; in->Remnant_P[2] = 0.002;   /* :462-464 */
    in->TimeIntegratorOrder = IUNSET;
    in->prob_type = IUNSET;
    in->is_polarization_scalar = 1;                    /* :333 */
    in->nsteps = 10;                                   /* :365 */
    in->plot_int = -1;                                 /* :370 */
    in->plot_Phi = 1; in->plot_PoissonRHS = 1; in->plot_E = 1; in->plot_holes = 1; in->plot_electrons = 1; in->plot_charge = 1; in->plot_epsilon = 1; in->plot_mask = 1; in->plot_tphase = 1; in->plot_alpha = 1; in->plot_beta = 1; in->plot_theta = 1; in->plot_PhiDiff = 1;   /* :300-325 */
    in->dt = UNSET;
    in->delta = 1.e-6;                                 /* :376 */
    in->random_seed = 1;                               /* :409 */
    in->voltage_sweep = 0;                             /* :388 */
    in->inc_step = -1;                                 /* :391 */
    in->num_Vapp_max = 1;                              /* :412 */
    in->Phi_Bc_lo = 0.; in->Phi_Bc_hi = 0.; in->Phi_Bc_inc = 0.; in->Phi_Bc_hi_max = 0.;
    in->phi_tolerance = 1.e-7;                         /* :406 */
    in->include_Landau = 1; in->include_Grad = 1; in->include_Elec = 1;
    in->lambda = UNSET;
    in->epsilon_0 = UNSET; in->epsilonX_fe = UNSET; in->epsilonX_fe_tphase = UNSET;
    in->epsilonZ_fe = UNSET; in->epsilon_de = UNSET; in->epsilon_si = UNSET;
    in->alpha = UNSET; in->beta = UNSET; in->gamma = UNSET;
    in->alpha_12 = UNSET; in->alpha_112 = UNSET; in->alpha_123 = UNSET;
    in->BigGamma = UNSET; in->g11 = UNSET; in->g44 = UNSET; in->g12 = UNSET; in->g44_p = UNSET;
    in->acceptor_doping = 0.0; in->donor_doping = 0.0;
    in->use_Fermi_Dirac = 1;                           /* :556 */
    in->use_work_function = 0;
    in->metal_work_function = 4.85;
    in->Coordinate_Transformation = 0;
    in->use_Euler_angles = 0;
    in->mg_tol_rel = 1.e-10;                           /* ElectrostaticSolver.cpp:719 */
    in->mg_max_iter = 200;
    in->mg_nu1 = 2; in->mg_nu2 = 2;
    in->mg_warm_start = 0;
    in->outer_max_iter = 1000;
    in->phi_ghost_mode = FX_GHOST_AMREX;
    in->precision = FX_PREC_FP64;
    in->bias_ramp_steps = 0;
    in->bias_ramp_start = 0.0; in->bias_ramp_end = 0.0;
    in->verbose = 0;
    in->mg_fused_sweep = 1;
    in->mg_block_levels = 1;
    in->mg_stream_min_cells = 1 << 21;
    in->mg_coarsen_ratio = 1.9;
    in->mg_exact_tail = 1;
    in->poisson_direct = 1; in->poisson_direct_check = 0; in->tdgl_fused_rhs = 0; in->tdgl_stream = 1;
    in->poisson_pcg = 1;
    in->mg_refine_nu1 = 1; in->mg_refine_nu2 = 2;   /* measured on the 256x256x128 workload: 281.6 against 266.5 steps/s with (2, 2) */
    return FX_OK;
}

/* Derived defaults and the required-key check (pp.get aborts in the
 * reference when a required key is missing).  Idempotent. */
int fx_inputs_finalize(fx_inputs* in, char* err, int errlen) {
    if (!in) return FX_EINVAL;
    struct { const char* name; double v; } req[] = {
        {"dt", in->dt}, {"lambda", in->lambda}, {"epsilon_0", in->epsilon_0}, {"epsilonX_fe", in->epsilonX_fe},
        {"epsilonZ_fe", in->epsilonZ_fe}, {"epsilon_de", in->epsilon_de}, {"epsilon_si", in->epsilon_si},
        {"alpha", in->alpha}, {"beta", in->beta}, {"gamma", in->gamma}, {"alpha_12", in->alpha_12},
        {"alpha_112", in->alpha_112}, {"alpha_123", in->alpha_123}, {"BigGamma", in->BigGamma},
        {"g11", in->g11}, {"g44", in->g44}, {"g12", in->g12}, {"g44_p", in->g44_p},
        {"domain.prob_lo", in->prob_lo[0]}, {"domain.prob_hi", in->prob_hi[0]},
        {"FE_lo", in->FE_lo[0]}, {"FE_hi", in->FE_hi[0]},
    };
    for (auto& r : req)
        if (std::isnan(r.v)) { set_err(err, errlen, std::string("missing required key ") + r.name); return FX_EINVAL; }
    if (in->n_cell[0] == IUNSET) { set_err(err, errlen, "missing required key domain.n_cell"); return FX_EINVAL; }
    if (in->TimeIntegratorOrder == IUNSET) { set_err(err, errlen, "missing required key TimeIntegratorOrder"); return FX_EINVAL; }
    if (in->prob_type == IUNSET) { set_err(err, errlen, "missing required key prob_type"); return FX_EINVAL; }
    for (int d = 0; d < 3; ++d) {
        if (std::isnan(in->DE_lo[d])) in->DE_lo[d] = in->prob_lo[d] - 1.0;
        if (std::isnan(in->SC_lo[d])) in->SC_lo[d] = in->prob_lo[d] - 1.0;
        if (std::isnan(in->Channel_lo[d])) in->Channel_lo[d] = in->prob_lo[d] - 1.0;
        if (std::isnan(in->DE_hi[d])) in->DE_hi[d] = in->prob_hi[d] + 1.0;
        if (std::isnan(in->SC_hi[d])) in->SC_hi[d] = in->prob_hi[d] + 1.0;
        if (std::isnan(in->Channel_hi[d])) in->Channel_hi[d] = in->prob_hi[d] + 1.0;
        if (in->bc_lo[d] < 0) in->bc_lo[d] = in->is_periodic[d] ? FX_BC_PERIODIC : FX_BC_DIRICHLET;
        if (in->bc_hi[d] < 0) in->bc_hi[d] = in->is_periodic[d] ? FX_BC_PERIODIC : FX_BC_DIRICHLET;
        if ((in->bc_lo[d] == FX_BC_PERIODIC) != (in->is_periodic[d] != 0) ||
            (in->bc_hi[d] == FX_BC_PERIODIC) != (in->is_periodic[d] != 0)) {
            set_err(err, errlen, "boundary.lo/hi 'per' tokens must match domain.is_periodic");
            return FX_EINVAL;
        }
        if (in->n_cell[d] < 1) { set_err(err, errlen, "domain.n_cell must be positive"); return FX_EINVAL; }
        if (!(in->prob_hi[d] > in->prob_lo[d])) { set_err(err, errlen, "domain.prob_hi must exceed domain.prob_lo"); return FX_EINVAL; }
        if (in->P_BC_flag_lo[d] < 0 || in->P_BC_flag_lo[d] > 4 || in->P_BC_flag_hi[d] < 0 || in->P_BC_flag_hi[d] > 4) {
            set_err(err, errlen, "P_BC_flag_lo/hi must be in 0..4");
            return FX_EINVAL;
        }
    }
    if (std::isnan(in->epsilonX_fe_tphase)) in->epsilonX_fe_tphase = in->epsilonX_fe;   /* FerroX.cpp:343-345 */
    if (in->prob_type < 1 || in->prob_type > 3) { set_err(err, errlen, "Invalid prob_type (1, 2 or 3)"); return FX_EINVAL; }
    if (in->TimeIntegratorOrder != 1 && in->TimeIntegratorOrder != 2) {
        set_err(err, errlen, "TimeIntegratorOrder must be 1 (forward Euler) or 2 (Heun)");
        return FX_EINVAL;
    }
    /* the V-cycles rely on their first pre-smoothing sweep to start every
     * coarse level from zero, so at least one is needed */
    if (in->mg_nu1 < 1 || in->mg_nu2 < 0 || in->mg_refine_nu1 < 1 || in->mg_refine_nu2 < 0 || in->mg_max_iter < 1 ||
        in->outer_max_iter < 1) {
        set_err(err, errlen, "bad multigrid controls (mg_nu1 and mg_refine_nu1 must be at least 1)");
        return FX_EINVAL;
    }
    if (in->tdgl_stream < 0 || in->tdgl_stream > 3) {
        set_err(err, errlen, "tdgl_stream must be 0 (gathered), 1 (streamed), 2 (streamed, face cells by list) "
                             "or 3 (streamed, face cells by plane)");
        return FX_EINVAL;
    }
    return FX_OK;
}

int fx_inputs_parse_file(const char* path, fx_inputs* in, char* err, int errlen) {
    if (!path || !in) return FX_EINVAL;
    std::ifstream f(path);
    if (!f) { set_err(err, errlen, std::string("cannot open ") + path); return FX_EINVAL; }
    std::stringstream ss;
    ss << f.rdbuf();
    Table t;
    std::string why;
    if (!parse_text(ss.str(), t, why) || !apply_table(t, in, why)) { set_err(err, errlen, why); return FX_EINVAL; }
    return FX_OK;
}

int fx_inputs_set(fx_inputs* in, const char* kv, char* err, int errlen) {
    if (!in || !kv) return FX_EINVAL;
    Table t;
    std::string why;
    if (!parse_text(kv, t, why) || !apply_table(t, in, why)) { set_err(err, errlen, why); return FX_EINVAL; }
    return FX_OK;
}

int fx_eval_cell_function(const fx_inputs* in, const char* expr, double* out, char* err, int errlen) {
    if (!in || !expr || !out) return FX_EINVAL;
    Fn f;
    std::string why;
    if (!compile_fn(expr, {}, f, why)) { set_err(err, errlen, why); return FX_EINVAL; }
    const int nx = in->n_cell[0], ny = in->n_cell[1], nz = in->n_cell[2];
    double dx[3];
    for (int d = 0; d < 3; ++d) dx[d] = (in->prob_hi[d] - in->prob_lo[d]) / in->n_cell[d];
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i)
                out[(size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k)] =
                    f(in->prob_lo[0] + (i + 0.5) * dx[0], in->prob_lo[1] + (j + 0.5) * dx[1], in->prob_lo[2] + (k + 0.5) * dx[2]);
    return FX_OK;
}

}  // extern "C"
