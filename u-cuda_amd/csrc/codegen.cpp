// codegen.cpp — parser, expression trees, KRS step text and HIP/C emitters.
//
// Parsing follows the reference grammar (U-CUDA/codegen.cpp:59-234) so the same
// input text means the same mathematics:
//   expr  := term (('+'|'-') term)*
//   term  := power (('*'|'/'|'\cdot'|implicit product in LaTeX) power)*
//   power := unary ['^' power]        (LaTeX: '^' takes a braced argument)
//   unary := '-' unary | '+' unary | primary
// so a leading minus binds tighter than '^' ("-x^2" is (-x)^2), as in the
// reference. Printing follows U-CUDA/codegen.cpp:236-291 (precedence-driven
// parentheses, integer powers 2..4 expanded to products, integers printed as
// "N.0", other numbers as the shortest round-trip "%.{p}g").
#include "codegen.hpp"

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <regex>
#include <memory>
#include <set>
#include <sstream>
#include <stdexcept>

namespace ucuda {
namespace {

// ------------------------------------------------------------------ trees --
// Raw: text printed as is (the harmonic inputs hc[k] / hs[k] of the
// device source), never parsed, differentiated or evaluated.
enum class K { Num, Var, Par, Time, Const, Call, Add, Sub, Mul, Div, Pow, Neg, Raw };

struct Node;
using P = std::shared_ptr<const Node>;
struct Node {
    K k;
    double num = 0.0;      // Num
    int idx = 0;           // Var: state index, Par: a[] index (1-based)
    std::string name;      // Call / Const / unresolved symbol
    P a, b;                // operands
    std::vector<P> args;   // Call
};

P num(double v) { auto n = std::make_shared<Node>(); n->k = K::Num; n->num = v; return n; }
P bin(K k, P a, P b) { auto n = std::make_shared<Node>(); n->k = k; n->a = std::move(a); n->b = std::move(b); return n; }
P neg(P a) { auto n = std::make_shared<Node>(); n->k = K::Neg; n->a = std::move(a); return n; }
P call(const std::string& f, std::vector<P> args) {
    auto n = std::make_shared<Node>(); n->k = K::Call; n->name = f; n->args = std::move(args); return n;
}
P sym(const std::string& s) { auto n = std::make_shared<Node>(); n->k = K::Const; n->name = s; return n; }

const std::set<std::string>& plain_functions() {
    static const std::set<std::string> f = {
        "sin", "cos", "tan", "asin", "acos", "atan", "sinh", "cosh", "tanh",
        "exp", "log", "log2", "log10", "sqrt", "cbrt", "fabs", "abs",
        "pow", "atan2", "fmod", "floor", "ceil"};
    return f;
}
const std::map<std::string, std::string>& latex_functions() {
    static const std::map<std::string, std::string> m = {
        {"sin", "sin"}, {"cos", "cos"}, {"tan", "tan"}, {"arcsin", "asin"},
        {"arccos", "acos"}, {"arctan", "atan"}, {"sinh", "sinh"}, {"cosh", "cosh"},
        {"tanh", "tanh"}, {"exp", "exp"}, {"ln", "log"}, {"log", "log"}, {"sqrt", "sqrt"}};
    return m;
}
const std::map<std::string, std::string>& latex_letters() {
    static const std::map<std::string, std::string> m = {
        {"alpha", "alpha"}, {"beta", "beta"}, {"gamma", "gamma"}, {"delta", "delta"},
        {"epsilon", "epsilon"}, {"varepsilon", "epsilon"}, {"zeta", "zeta"}, {"eta", "eta"},
        {"theta", "theta"}, {"vartheta", "theta"}, {"iota", "iota"}, {"kappa", "kappa"},
        {"lambda", "lambda"}, {"mu", "mu"}, {"nu", "nu"}, {"xi", "xi"}, {"pi", "pi"},
        {"rho", "rho"}, {"sigma", "sigma"}, {"tau", "tau"}, {"phi", "phi"}, {"varphi", "phi"},
        {"chi", "chi"}, {"psi", "psi"}, {"omega", "omega"}, {"Gamma", "Gamma"},
        {"Delta", "Delta"}, {"Theta", "Theta"}, {"Lambda", "Lambda"}, {"Sigma", "Sigma"},
        {"Phi", "Phi"}, {"Psi", "Psi"}, {"Omega", "Omega"}};
    return m;
}

// ----------------------------------------------------------------- parser --
class Parser {
public:
    Parser(const std::string& text, bool latex) : s_(text), latex_(latex) {}

    P parse() {
        P e = expr();
        skip();
        if (pos_ != s_.size()) fail("trailing input at pos " + std::to_string(pos_));
        return e;
    }

private:
    const std::string s_;
    size_t pos_ = 0;
    bool latex_;

    [[noreturn]] void fail(const std::string& m) const { throw std::runtime_error(m); }

    void skip() {
        for (;;) {
            while (pos_ < s_.size() && std::isspace((unsigned char)s_[pos_])) ++pos_;
            // LaTeX spacing commands carry no meaning in an expression.
            if (latex_ && pos_ + 1 < s_.size() && s_[pos_] == '\\' &&
                (s_[pos_ + 1] == ',' || s_[pos_ + 1] == ';' || s_[pos_ + 1] == '!' ||
                 s_[pos_ + 1] == ' ' || s_[pos_ + 1] == ':')) {
                pos_ += 2;
                continue;
            }
            break;
        }
    }
    char peek() { skip(); return pos_ < s_.size() ? s_[pos_] : '\0'; }
    bool accept(char c) { if (peek() == c) { ++pos_; return true; } return false; }
    bool starts(const char* w) const { return s_.compare(pos_, std::strlen(w), w) == 0; }

    std::string command_name() {
        size_t j = pos_;
        while (j < s_.size() && std::isalpha((unsigned char)s_[j])) ++j;
        std::string c = s_.substr(pos_, j - pos_);
        pos_ = j;
        return c;
    }
    double number() {
        size_t j = pos_;
        while (j < s_.size() && (std::isdigit((unsigned char)s_[j]) || s_[j] == '.')) ++j;
        if (j < s_.size() && (s_[j] == 'e' || s_[j] == 'E')) {
            ++j;
            if (j < s_.size() && (s_[j] == '+' || s_[j] == '-')) ++j;
            while (j < s_.size() && std::isdigit((unsigned char)s_[j])) ++j;
        }
        std::string t = s_.substr(pos_, j - pos_);
        char* end = nullptr;
        double v = std::strtod(t.c_str(), &end);
        if (t.empty() || end == t.c_str()) fail("bad number: '" + t + "'");
        pos_ = j;
        return v;
    }
    std::string identifier() {
        size_t j = pos_;
        while (j < s_.size() && (std::isalnum((unsigned char)s_[j]) || s_[j] == '_')) ++j;
        std::string id = s_.substr(pos_, j - pos_);
        pos_ = j;
        return id;
    }

    P expr() {
        P n = term();
        for (;;) {
            char c = peek();
            if (c == '+') { ++pos_; n = bin(K::Add, n, term()); }
            else if (c == '-') { ++pos_; n = bin(K::Sub, n, term()); }
            else return n;
        }
    }
    bool primary_follows() {
        skip();
        if (pos_ >= s_.size()) return false;
        char c = s_[pos_];
        if (c == '(' || c == '{' || c == '.' || c == '_') return true;
        if (std::isalnum((unsigned char)c)) return true;
        if (c == '\\') return !(starts("\\cdot") || starts("\\right"));
        return false;
    }
    P term() {
        P n = power();
        for (;;) {
            if (latex_) {
                skip();
                if (starts("\\cdot")) { pos_ += 5; n = bin(K::Mul, n, power()); continue; }
            }
            char c = peek();
            if (c == '*') { ++pos_; n = bin(K::Mul, n, power()); }
            else if (c == '/') { ++pos_; n = bin(K::Div, n, power()); }
            else if (latex_ && primary_follows()) { n = bin(K::Mul, n, power()); }
            else return n;
        }
    }
    P power() {
        P base = unary();
        if (peek() == '^') {
            ++pos_;
            return bin(K::Pow, base, latex_ ? braced() : power());
        }
        return base;
    }
    P unary() {
        char c = peek();
        if (c == '-') { ++pos_; return neg(unary()); }
        if (c == '+') { ++pos_; return unary(); }
        return primary();
    }
    // LaTeX argument: {expr} or a single token.
    P braced() {
        skip();
        if (pos_ < s_.size() && s_[pos_] == '{') {
            ++pos_;
            P e = expr();
            if (!accept('}')) fail("expected }");
            return e;
        }
        return single();
    }
    P single() {
        skip();
        char c = pos_ < s_.size() ? s_[pos_] : '\0';
        if (std::isdigit((unsigned char)c) || c == '.') return num(number());
        if (latex_ && c == '\\') {
            ++pos_;
            std::string cmd = command_name();
            auto it = latex_letters().find(cmd);
            if (it != latex_letters().end()) return sym(it->second);
            fail("unexpected latex command \\" + cmd + " as single arg");
        }
        if (std::isalpha((unsigned char)c) || c == '_') return sym(identifier());
        return primary();
    }
    P primary() {
        skip();
        char c = pos_ < s_.size() ? s_[pos_] : '\0';
        if (c == '(') {
            ++pos_;
            P e = expr();
            if (!accept(')')) fail("expected )");
            return e;
        }
        if (latex_) {
            if (starts("\\left")) {
                pos_ += 5;
                skip();
                if (pos_ < s_.size() && (s_[pos_] == '(' || s_[pos_] == '[')) ++pos_;
                P e = expr();
                skip();
                if (starts("\\right")) {
                    pos_ += 6;
                    skip();
                    if (pos_ < s_.size() && (s_[pos_] == ')' || s_[pos_] == ']')) ++pos_;
                }
                return e;
            }
            if (c == '{') {
                ++pos_;
                P e = expr();
                if (!accept('}')) fail("expected }");
                return e;
            }
            if (c == '\\') {
                ++pos_;
                std::string cmd = command_name();
                if (cmd == "frac") {
                    P n = braced();
                    P d = braced();
                    return bin(K::Div, n, d);
                }
                auto f = latex_functions().find(cmd);
                if (f != latex_functions().end()) {
                    P exponent;
                    if (peek() == '^') { ++pos_; exponent = braced(); }
                    P fn = call(f->second, {braced()});
                    return exponent ? bin(K::Pow, fn, exponent) : fn;
                }
                auto l = latex_letters().find(cmd);
                if (l != latex_letters().end()) return sym(l->second);
                fail("unknown latex command \\" + cmd);
            }
        }
        if (std::isdigit((unsigned char)c) || c == '.') return num(number());
        if (std::isalpha((unsigned char)c) || c == '_') {
            std::string id = identifier();
            if (peek() == '(') {
                if (!plain_functions().count(id)) fail("unknown function: " + id);
                ++pos_;
                std::vector<P> args;
                if (peek() != ')') {
                    args.push_back(expr());
                    while (accept(',')) args.push_back(expr());
                }
                if (!accept(')')) fail("expected ) after " + id);
                return call(id, std::move(args));
            }
            return sym(id);
        }
        fail(std::string("unexpected char '") + c + "' at pos " + std::to_string(pos_));
    }
};

// Replaces symbol names by state / parameter / time references.
P resolve(const P& n, const System& s) {
    switch (n->k) {
    case K::Const: {
        for (size_t i = 0; i < s.vars.size(); ++i)
            if (s.vars[i] == n->name) { auto r = std::make_shared<Node>(); r->k = K::Var; r->idx = (int)i; r->name = n->name; return r; }
        for (size_t j = 0; j < s.params.size(); ++j)
            if (s.params[j] == n->name) { auto r = std::make_shared<Node>(); r->k = K::Par; r->idx = 1 + (int)j; r->name = n->name; return r; }
        if (!s.time_symbol.empty() && n->name == s.time_symbol) { auto r = std::make_shared<Node>(); r->k = K::Time; return r; }
        if (n->name == "pi") return n;
        throw std::runtime_error("unknown symbol: " + n->name);
    }
    case K::Num: case K::Var: case K::Par: case K::Time: case K::Raw: return n;
    case K::Neg: return neg(resolve(n->a, s));
    case K::Call: {
        std::vector<P> a;
        for (auto& x : n->args) a.push_back(resolve(x, s));
        return call(n->name, std::move(a));
    }
    default: return bin(n->k, resolve(n->a, s), resolve(n->b, s));
    }
}

std::vector<P> parse_system(const System& s) {
    if (s.vars.size() != s.rhs.size()) throw std::runtime_error("vars/rhs size mismatch");
    std::vector<P> out;
    for (const auto& r : s.rhs) out.push_back(resolve(Parser(r, s.latex).parse(), s));
    return out;
}

// --------------------------------------------------------------- printing --
int precedence(K k) {
    switch (k) {
    case K::Add: case K::Sub: return 1;
    case K::Mul: case K::Div: return 2;
    case K::Neg: return 3;
    case K::Pow: return 4;
    default: return 5;
    }
}

// "N.0" for integers below 1e12, otherwise the shortest "%.{p}g" that
// round-trips, with ".0" appended when it has neither '.' nor an exponent.
std::string format_number(double v) {
    if (v == (double)(long long)v) {
        long long iv = (long long)v;
        if (iv > -1000000000000LL && iv < 1000000000000LL) return std::to_string(iv) + ".0";
    }
    char buf[64];
    for (int p = 1; p <= 17; ++p) {
        std::snprintf(buf, sizeof buf, "%.*g", p, v);
        if (std::strtod(buf, nullptr) == v) break;
    }
    std::string r(buf);
    if (r.find('.') == std::string::npos && r.find('e') == std::string::npos &&
        r.find("inf") == std::string::npos && r.find("nan") == std::string::npos)
        r += ".0";
    return r;
}

struct Style {
    std::string state = "X";       // state array name
    std::string params = "a";      // parameter array name
    std::string time = "t";        // time variable name
    bool typed_literals = false;   // wrap literals as ((T)v)
    bool c_functions = false;      // abs -> fabs for C
};

void print(const P& n, const Style& st, std::ostream& o);
void print_child(const P& n, const Style& st, std::ostream& o, int min_prec) {
    if (precedence(n->k) < min_prec) { o << "("; print(n, st, o); o << ")"; }
    else print(n, st, o);
}
void print(const P& n, const Style& st, std::ostream& o) {
    switch (n->k) {
    case K::Num:
        if (st.typed_literals) o << "((T)" << format_number(n->num) << ")";
        else o << format_number(n->num);
        break;
    case K::Var: o << st.state << "[" << n->idx << "]"; break;
    case K::Par: o << st.params << "[" << n->idx << "]"; break;
    case K::Time: o << st.time; break;
    case K::Const: o << (st.typed_literals ? "((T)3.14159265358979323846)" : n->name); break;
    case K::Raw: o << n->name; break;
    case K::Neg: o << "-"; print_child(n->a, st, o, precedence(K::Neg)); break;
    case K::Add: print_child(n->a, st, o, 1); o << " + "; print_child(n->b, st, o, 1); break;
    case K::Sub: print_child(n->a, st, o, 1); o << " - "; print_child(n->b, st, o, 2); break;
    case K::Mul: print_child(n->a, st, o, 2); o << " * "; print_child(n->b, st, o, 2); break;
    case K::Div: print_child(n->a, st, o, 2); o << " / "; print_child(n->b, st, o, 3); break;
    case K::Pow: {
        if (n->b->k == K::Num) {
            double e = n->b->num;
            int ie = (int)e;
            if (e == (double)ie && ie >= 2 && ie <= 4) {
                o << "(";
                for (int i = 0; i < ie; ++i) { if (i) o << " * "; print_child(n->a, st, o, 2); }
                o << ")";
                break;
            }
        }
        o << "pow("; print(n->a, st, o); o << ", "; print(n->b, st, o); o << ")";
        break;
    }
    case K::Call: {
        std::string f = n->name;
        if (f == "abs") f = "fabs";
        o << f << "(";
        for (size_t i = 0; i < n->args.size(); ++i) { if (i) o << ", "; print(n->args[i], st, o); }
        o << ")";
        break;
    }
    }
}
std::string to_text(const P& n, const Style& st) { std::ostringstream o; print(n, st, o); return o.str(); }

// ------------------------------------------------ differentiation (exact) --
bool is_num(const P& n, double v) { return n->k == K::Num && n->num == v; }
bool is_zero(const P& n) { return is_num(n, 0.0); }

P s_add(P a, P b) {
    if (is_zero(a)) return b;
    if (is_zero(b)) return a;
    if (a->k == K::Num && b->k == K::Num) return num(a->num + b->num);
    if (b->k == K::Neg) return bin(K::Sub, a, b->a);
    return bin(K::Add, a, b);
}
P s_neg(P a) {
    if (a->k == K::Num) return num(-a->num);
    if (a->k == K::Neg) return a->a;
    return neg(a);
}
P s_sub(P a, P b) {
    if (is_zero(b)) return a;
    if (is_zero(a)) return s_neg(b);
    if (a->k == K::Num && b->k == K::Num) return num(a->num - b->num);
    if (b->k == K::Neg) return bin(K::Add, a, b->a);
    return bin(K::Sub, a, b);
}
P s_mul(P a, P b) {
    if (is_zero(a) || is_zero(b)) return num(0.0);
    if (is_num(a, 1.0)) return b;
    if (is_num(b, 1.0)) return a;
    if (is_num(a, -1.0)) return s_neg(b);
    if (is_num(b, -1.0)) return s_neg(a);
    if (a->k == K::Num && b->k == K::Num) return num(a->num * b->num);
    if (a->k == K::Neg && b->k == K::Neg) return s_mul(a->a, b->a);
    if (a->k == K::Neg) return s_neg(s_mul(a->a, b));
    if (b->k == K::Neg) return s_neg(s_mul(a, b->a));
    return bin(K::Mul, a, b);
}
P s_div(P a, P b) {
    if (is_zero(a)) return num(0.0);
    if (is_num(b, 1.0)) return a;
    if (a->k == K::Neg) return s_neg(s_div(a->a, b));
    return bin(K::Div, a, b);
}
P f1(const char* f, P x) { return call(f, {std::move(x)}); }

// d n / d X[j]
P diff(const P& n, int j) {
    switch (n->k) {
    case K::Num: case K::Par: case K::Time: case K::Const: case K::Raw: return num(0.0);
    case K::Var: return num(n->idx == j ? 1.0 : 0.0);
    case K::Neg: return s_neg(diff(n->a, j));
    case K::Add: return s_add(diff(n->a, j), diff(n->b, j));
    case K::Sub: return s_sub(diff(n->a, j), diff(n->b, j));
    case K::Mul: return s_add(s_mul(diff(n->a, j), n->b), s_mul(n->a, diff(n->b, j)));
    case K::Div: {
        P da = diff(n->a, j), db = diff(n->b, j);
        if (is_zero(db)) return s_div(da, n->b);
        // (da*b - a*db) / (b*b)
        return s_div(s_sub(s_mul(da, n->b), s_mul(n->a, db)), bin(K::Mul, n->b, n->b));
    }
    case K::Pow: {
        P du = diff(n->a, j);
        if (n->b->k == K::Num) {
            double e = n->b->num;
            if (is_zero(du)) return num(0.0);
            P d;
            if (e == 2.0) d = s_mul(num(2.0), n->a);
            else if (e == 1.0) d = num(1.0);
            else d = s_mul(num(e), bin(K::Pow, n->a, num(e - 1.0)));
            return s_mul(d, du);
        }
        P dv = diff(n->b, j);
        // d(u^v) = u^v * (dv*log(u) + v*du/u)
        P inner = s_add(s_mul(dv, f1("log", n->a)), s_div(s_mul(n->b, du), n->a));
        if (is_zero(inner)) return num(0.0);
        return s_mul(n, inner);
    }
    case K::Call: {
        const std::string& f = n->name;
        if (n->args.size() == 2) {
            P u = n->args[0], v = n->args[1];
            P du = diff(u, j), dv = diff(v, j);
            if (f == "pow") return diff(bin(K::Pow, u, v), j);
            if (f == "fmod") return s_sub(du, s_mul(f1("trunc", s_div(u, v)), dv));
            if (f == "atan2") {  // d atan2(u,v) = (v du - u dv)/(u^2+v^2)
                if (is_zero(du) && is_zero(dv)) return num(0.0);
                return s_div(s_sub(s_mul(v, du), s_mul(u, dv)), bin(K::Add, bin(K::Mul, u, u), bin(K::Mul, v, v)));
            }
            throw std::runtime_error("cannot differentiate " + f);
        }
        if (n->args.size() != 1) throw std::runtime_error("bad arity for " + f);
        P u = n->args[0];
        P du = diff(u, j);
        if (is_zero(du)) return num(0.0);
        P g;
        if (f == "sin") g = f1("cos", u);
        else if (f == "cos") g = s_neg(f1("sin", u));
        else if (f == "tan") { P c = f1("cos", u); g = s_div(num(1.0), bin(K::Mul, c, c)); }
        else if (f == "asin") g = s_div(num(1.0), f1("sqrt", bin(K::Sub, num(1.0), bin(K::Mul, u, u))));
        else if (f == "acos") g = s_neg(s_div(num(1.0), f1("sqrt", bin(K::Sub, num(1.0), bin(K::Mul, u, u)))));
        else if (f == "atan") g = s_div(num(1.0), bin(K::Add, num(1.0), bin(K::Mul, u, u)));
        else if (f == "sinh") g = f1("cosh", u);
        else if (f == "cosh") g = f1("sinh", u);
        else if (f == "tanh") { P t = f1("tanh", u); g = bin(K::Sub, num(1.0), bin(K::Mul, t, t)); }
        else if (f == "exp") g = n;
        else if (f == "log") g = s_div(num(1.0), u);
        else if (f == "log2") g = s_div(num(1.4426950408889634), u);
        else if (f == "log10") g = s_div(num(0.4342944819032518), u);
        else if (f == "sqrt") g = s_div(num(0.5), n);
        else if (f == "cbrt") g = s_div(num(1.0), bin(K::Mul, num(3.0), bin(K::Mul, n, n)));
        else if (f == "fabs" || f == "abs") g = call("copysign", {num(1.0), u});
        else if (f == "floor" || f == "ceil") return num(0.0);
        else throw std::runtime_error("cannot differentiate " + f);
        return s_mul(g, du);
    }
    }
    return num(0.0);
}

// ------------------------------------------------------------ op counting --
void count(const P& n, int& flops, int& calls) {
    switch (n->k) {
    case K::Num: case K::Var: case K::Par: case K::Time: case K::Const: case K::Raw: return;
    case K::Neg: count(n->a, flops, calls); return;  // sign flips are free modifiers
    case K::Call: ++calls; for (auto& a : n->args) count(a, flops, calls); return;
    case K::Pow:
        count(n->a, flops, calls);
        if (n->b->k == K::Num && n->b->num == (int)n->b->num && n->b->num >= 2 && n->b->num <= 4) {
            flops += (int)n->b->num - 1;
            // the base is printed once per factor
            int f2 = 0, c2 = 0; count(n->a, f2, c2); flops += f2 * ((int)n->b->num - 1); calls += c2 * ((int)n->b->num - 1);
        } else { ++calls; count(n->b, flops, calls); }
        return;
    default: ++flops; count(n->a, flops, calls); count(n->b, flops, calls); return;
    }
}

double eval(const P& n, const double* X, const double* a, double t) {
    switch (n->k) {
    case K::Num: return n->num;
    case K::Var: return X[n->idx];
    case K::Par: return a[n->idx];
    case K::Time: return t;
    case K::Const: return 3.14159265358979323846;
    case K::Raw: throw std::runtime_error("raw text has no value");
    case K::Neg: return -eval(n->a, X, a, t);
    case K::Add: return eval(n->a, X, a, t) + eval(n->b, X, a, t);
    case K::Sub: return eval(n->a, X, a, t) - eval(n->b, X, a, t);
    case K::Mul: return eval(n->a, X, a, t) * eval(n->b, X, a, t);
    case K::Div: return eval(n->a, X, a, t) / eval(n->b, X, a, t);
    case K::Pow: {
        double u = eval(n->a, X, a, t);
        if (n->b->k == K::Num) {
            double e = n->b->num; int ie = (int)e;
            if (e == (double)ie && ie >= 2 && ie <= 4) { double r = u; for (int i = 1; i < ie; ++i) r = r * u; return r; }
        }
        return std::pow(u, eval(n->b, X, a, t));
    }
    case K::Call: {
        const std::string& f = n->name;
        double u = eval(n->args[0], X, a, t);
        if (n->args.size() == 2) {
            double v = eval(n->args[1], X, a, t);
            if (f == "pow") return std::pow(u, v);
            if (f == "atan2") return std::atan2(u, v);
            if (f == "fmod") return std::fmod(u, v);
            if (f == "copysign") return std::copysign(u, v);
        }
        if (f == "sin") return std::sin(u); if (f == "cos") return std::cos(u); if (f == "tan") return std::tan(u);
        if (f == "asin") return std::asin(u); if (f == "acos") return std::acos(u); if (f == "atan") return std::atan(u);
        if (f == "sinh") return std::sinh(u); if (f == "cosh") return std::cosh(u); if (f == "tanh") return std::tanh(u);
        if (f == "exp") return std::exp(u); if (f == "log") return std::log(u); if (f == "log2") return std::log2(u);
        if (f == "log10") return std::log10(u); if (f == "sqrt") return std::sqrt(u); if (f == "cbrt") return std::cbrt(u);
        if (f == "fabs" || f == "abs") return std::fabs(u); if (f == "floor") return std::floor(u);
        if (f == "ceil") return std::ceil(u); if (f == "trunc") return std::trunc(u);
        throw std::runtime_error("unknown function " + f);
    }
    }
    return 0.0;
}

// ----------------------------------------------------------- KRS schemes --
// Text layout of each scheme follows U-CUDA/codegen.cpp:579-606 so that
// callers comparing step bodies (and the CPU oracle, which compiles them)
// see the same statements.
std::vector<std::string> rhs_text(const System& s, const std::string& state) {
    Style st; st.state = state;
    std::vector<std::string> out;
    for (auto& t : parse_system(s)) out.push_back(to_text(t, st));
    return out;
}

std::string krs_euler(const System& s) {
    int N = (int)s.vars.size();
    auto f = rhs_text(s, "X");
    std::ostringstream o;
    o << "    numb X1[" << N << "];\n    int i;\n";
    for (int i = 0; i < N; ++i) o << "    X1[" << i << "] = X[" << i << "] + h * (" << f[i] << ");\n";
    o << "    for (i = 0; i < " << N << "; i++)\n        X[i] = X1[i];\n";
    return o.str();
}
std::string krs_euler_cromer(const System& s) {
    auto f = rhs_text(s, "X");
    std::ostringstream o;
    for (size_t i = 0; i < f.size(); ++i) o << "    X[" << i << "] = X[" << i << "] + h * (" << f[i] << ");\n";
    return o.str();
}
std::string krs_midpoint(const System& s) {
    int N = (int)s.vars.size();
    auto f0 = rhs_text(s, "X"), f1v = rhs_text(s, "X1");
    std::ostringstream o;
    o << "    numb X1[" << N << "];\n";
    for (int i = 0; i < N; ++i) o << "    X1[" << i << "] = X[" << i << "] + 0.5 * h * (" << f0[i] << ");\n";
    for (int i = 0; i < N; ++i) o << "    X[" << i << "] = X[" << i << "] + h * (" << f1v[i] << ");\n";
    return o.str();
}
std::string krs_rk4(const System& s) {
    int N = (int)s.vars.size();
    auto k = rhs_text(s, "X1");
    std::ostringstream o;
    o << "    numb X1[" << N << "];\n    numb k[" << N << "][4];\n    int N = " << N << ";\n    int i, j;\n";
    o << "    for (i = 0; i < N; i++) {\n        X1[i] = X[i];\n    }\n";
    o << "    for (j = 0; j < 4; j++) {\n";
    for (int i = 0; i < N; ++i) o << "        k[" << i << "][j] = (" << k[i] << ");\n";
    o << "        if (j == 3) {\n            for (i = 0; i < N; i++) {\n"
         "            X[i] = X[i] + h * (k[i][0] + 2 * k[i][1] + 2 * k[i][2] + k[i][3]) / 6;\n"
         "            }\n        }\n";
    o << "        else if (j == 2) {\n            for (i = 0; i < N; i++) {\n"
         "            X1[i] = X[i] + h * k[i][j];\n            }\n        }\n";
    o << "        else {\n            for (i = 0; i < N; i++) {\n"
         "                X1[i] = X[i] + 0.5 * h * k[i][j];\n            }\n        }\n    }\n";
    return o.str();
}


// Dormand-Prince 8(7) 13-stage step (Prince & Dormand 1981, RK8(7)13M) in
// the reference's text form (U-CUDA/codegen.cpp:611-675): the coefficient
// tables are printed exactly as the reference prints them (13 significant
// digits), the 8th-order solution y replaces X, the 7th-order z is formed
// and discarded.
std::string krs_dopri78(const System& s) {
    static const char* const kTables =
        "    static const numb M[13][12] = {\n"
        "        {0,0,0,0,0,0,0,0,0,0,0,0},\n"
        "        {0.05555555555556,0,0,0,0,0,0,0,0,0,0,0},\n"
        "        {0.02083333333333,0.0625,0,0,0,0,0,0,0,0,0,0},\n"
        "        {0.03125,0,0.09375,0,0,0,0,0,0,0,0,0},\n"
        "        {0.3125,0,-1.171875,1.171875,0,0,0,0,0,0,0,0},\n"
        "        {0.0375,0,0,0.1875,0.15,0,0,0,0,0,0,0},\n"
        "        {0.04791013711111,0,0,0.1122487127778,-0.02550567377778,0.01284682388889,0,0,0,0,0,0},\n"
        "        {0.01691798978729,0,0,0.387848278486,0.0359773698515,0.1969702142157,-0.1727138523405,0,0,0,0,0},\n"
        "        {0.06909575335919,0,0,-0.6342479767289,-0.1611975752246,0.1386503094588,0.9409286140358,0.2116363264819,0,0,0,0},\n"
        "        {0.183556996839,0,0,-2.468768084316,-0.2912868878163,-0.02647302023312,2.847838764193,0.2813873314699,0.1237448998633,0,0,0},\n"
        "        {-1.215424817396,0,0,16.67260866595,0.9157418284168,-6.056605804357,-16.00357359416,14.8493030863,-13.37157573529,5.13418264818,0,0},\n"
        "        {0.2588609164383,0,0,-4.774485785489,-0.435093013777,-3.049483332072,5.577920039936,6.155831589861,-5.062104586737,2.193926173181,0.1346279986593,0},\n"
        "        {0.8224275996265,0,0,-11.65867325728,-0.7576221166909,0.7139735881596,12.07577498689,-2.12765911392,1.990166207049,-0.234286471544,0.1758985777079,0}\n"
        "    };\n"
        "    static const numb B[2][13] = {\n"
        "        {0.04174749114153,0,0,0,0,-0.05545232861124,0.2393128072012,0.7035106694034,-0.7597596138145,0.6605630309223,0.1581874825101,-0.2381095387529,0.25},\n"
        "        {0.02955321367635,0,0,0,0,-0.8286062764878,0.3112409000511,2.4673451906,-2.546941651842,1.443548583677,0.07941559588113,0.04444444444444,0}\n"
        "    };\n";
    const int N = (int)s.vars.size();
    const auto k = rhs_text(s, "X1");
    std::ostringstream o;
    o << kTables;
    for (const char* arr : {"X1", "X2", "y", "z"}) o << "    numb " << arr << "[" << N << "];\n";
    o << "    numb k[" << N << "][13];\n    int N = " << N << ";\n    int i, j, l;\n";
    o << "    for (i = 0; i < N; ++i) X1[i] = X[i];\n    for (i = 0; i < 13; ++i) {\n";
    for (int v = 0; v < N; ++v) o << "        k[" << v << "][i] = (" << k[v] << ");\n";
    o << "        if (i != 12) {\n"
         "            for (l = 0; l < N; ++l) X2[l] = 0;\n"
         "            for (j = 0; j < i + 1; ++j)\n"
         "                for (l = 0; l < N; ++l)\n"
         "                    X2[l] += M[i + 1][j] * k[l][j];\n"
         "            for (l = 0; l < N; ++l)\n"
         "                X1[l] = X[l] + h * X2[l];\n"
         "        }\n    }\n";
    // weighted sums of the stages: row 0 -> y (order 8), row 1 -> z (order 7)
    const char* const target[2] = {"y", "z"};
    for (int row = 0; row < 2; ++row) {
        o << "    for (l = 0; l < N; ++l) X2[l] = 0;\n"
             "    for (i = 0; i < 13; ++i)\n"
             "        for (l = 0; l < N; ++l)\n"
             "            X2[l] += B[" << row << "][i] * k[l][i];\n"
             "    for (l = 0; l < N; ++l)\n"
             "        " << target[row] << "[l] = X[l] + h * X2[l];\n";
    }
    o << "    (void)z[0];\n    for (l = 0; l < N; ++l) X[l] = y[l];\n";
    return o.str();
}

// ------------------------------------------------------- CD (composition) --
// Composition of an explicit half step h1 = h*a[0] in forward variable order
// and a diagonally implicit half step h2 = h*(1 - a[0]) in reverse order
// (U-CUDA/codegen.cpp:677-754). Each right-hand side is printed in plain form
// (names as written, powers as x^n) and classified per variable by the
// reference's textual rules: absent -> explicit update; linear (no x^n with
// n >= 2, no x*x, x not inside a known function call) -> the analytic
// solution of x = x0 + h2*(sign*coef*x + rest); otherwise four fixed-point
// iterations from the explicit value. The rules work on the text, so the
// same std::regex patterns are used to reproduce the same statements.
void print_plain(const P& n, std::ostream& o);
void print_plain_child(const P& n, std::ostream& o, int min_prec) {
    if (precedence(n->k) < min_prec) { o << "("; print_plain(n, o); o << ")"; }
    else print_plain(n, o);
}
void print_plain(const P& n, std::ostream& o) {
    switch (n->k) {
    case K::Num: o << format_number(n->num); break;
    case K::Neg: o << "-"; print_plain_child(n->a, o, precedence(K::Neg)); break;
    case K::Add: print_plain_child(n->a, o, 1); o << " + "; print_plain_child(n->b, o, 1); break;
    case K::Sub: print_plain_child(n->a, o, 1); o << " - "; print_plain_child(n->b, o, 2); break;
    case K::Mul: print_plain_child(n->a, o, 2); o << " * "; print_plain_child(n->b, o, 2); break;
    case K::Div: print_plain_child(n->a, o, 2); o << " / "; print_plain_child(n->b, o, 3); break;
    case K::Pow: print_plain_child(n->a, o, 4); o << "^"; print_plain_child(n->b, o, 4); break;
    case K::Call:
        o << n->name << "(";
        for (size_t i = 0; i < n->args.size(); ++i) { if (i) o << ", "; print_plain(n->args[i], o); }
        o << ")";
        break;
    default: o << n->name; break;  // unresolved symbol
    }
}

std::string trim(const std::string& s) {
    const size_t a = s.find_first_not_of(" \t\n\r");
    if (a == std::string::npos) return "";
    return s.substr(a, s.find_last_not_of(" \t\n\r") - a + 1);
}

// regex helpers: a pattern that fails to compile counts as "no match"
bool re_search(const std::string& text, const std::string& pattern, std::smatch* m = nullptr) {
    try {
        std::smatch local;
        return std::regex_search(text, m ? *m : local, std::regex(pattern));
    } catch (const std::regex_error&) { return false; }
}
bool re_full(const std::string& text, const std::string& pattern, std::smatch* m = nullptr) {
    try {
        std::smatch local;
        return std::regex_match(text, m ? *m : local, std::regex(pattern));
    } catch (const std::regex_error&) { return false; }
}
std::string re_replace(const std::string& text, const std::string& pattern, const std::string& with) {
    try { return std::regex_replace(text, std::regex(pattern), with); } catch (const std::regex_error&) { return text; }
}

enum class Dependence { None, Linear, Nonlinear };

Dependence cd_dependence(const std::string& expr, const std::string& var) {
    if (!re_search(expr, "\\b" + var + "\\b")) return Dependence::None;
    std::smatch m;
    if (re_search(expr, "\\b" + var + "\\s*\\^\\s*(\\d+)", &m) && std::stoi(m[1].str()) >= 2) return Dependence::Nonlinear;
    if (re_search(expr, "\\b" + var + "\\s*\\*\\s*" + var + "\\b")) return Dependence::Nonlinear;
    static const char* const kFunctions[] = {"abs", "acos", "asin", "atan", "cos", "cosh", "exp", "fabs",
                                             "ln", "log", "pow", "sin", "sinh", "sqrt", "tan", "tanh"};
    for (const char* f : kFunctions)
        if (re_search(expr, std::string(f) + "\\s*\\([^)]*\\b" + var + "\\b[^)]*\\)")) return Dependence::Nonlinear;
    return Dependence::Linear;
}

// Splits at every '+' / '-' after the first character (parentheses are not
// tracked), except the sign of a number's exponent.
std::vector<std::string> cd_terms(const std::string& expr) {
    std::vector<std::string> out;
    size_t start = 0;
    auto push = [&](size_t from, size_t to) {
        std::string t = trim(expr.substr(from, to - from));
        if (!t.empty()) out.push_back(t);
    };
    for (size_t i = 1; i < expr.size(); ++i) {
        if ((expr[i] == '+' || expr[i] == '-') && expr[i - 1] != 'e' && expr[i - 1] != 'E') {
            push(start, i);
            start = i;
        }
    }
    if (start < expr.size()) push(start, expr.size());
    return out;
}

void signed_body(const std::string& token, std::string& sign, std::string& body) {
    sign = "+";
    body = token;
    if (!token.empty() && (token[0] == '-' || token[0] == '+')) {
        if (token[0] == '-') sign = "-";
        body = trim(token.substr(1));
    }
}

std::string join_rest(const std::vector<std::string>& parts) {
    std::string r;
    for (const auto& p : parts) r += (r.empty() ? "" : " ") + p;
    if (!r.empty() && r[0] == '+') r = trim(r.substr(1));
    return r;
}

struct LinearForm { std::string sign = "+", coef = "1", rest; };

// outer * (t1 +- t2 ...): the variable's sign inside, every other term
// multiplied by `outer`
LinearForm cd_distribute(const std::string& inner_in, const std::string& var, const std::string& outer) {
    std::string inner = inner_in;
    if (!inner.empty() && inner[0] != '+' && inner[0] != '-') inner = "+" + inner;
    LinearForm f;
    f.coef = outer;
    std::vector<std::string> rest;
    for (const auto& tok : cd_terms(inner)) {
        std::string sign, body;
        signed_body(tok, sign, body);
        if (re_full(body, "^" + var + "$")) { f.sign = sign; continue; }
        rest.push_back((sign == "-" ? "- " : "+ ") + outer + " * " + body);
    }
    f.rest = join_rest(rest);
    return f;
}

LinearForm cd_linear_form(const std::string& expr_in, const std::string& var) {
    const std::string expr = trim(expr_in);
    std::smatch m;
    // (left) * (right), the variable only on the right
    if (re_full(expr, R"(^\(([^)]+)\)\s*\*\s*\(([^)]+)\)$)", &m)) {
        const std::string left = m[1].str(), right = m[2].str();
        if (re_search(right, "\\b" + var + "\\b") && !re_search(left, "\\b" + var + "\\b"))
            return cd_distribute(right, var, "(" + left + ")");
    }
    // name * (inner) with the variable inside
    if (re_full(expr, R"(^([a-zA-Z_]\w*)\s*\*\s*\((.+)\)$)", &m)) {
        const std::string outer = m[1].str(), inner = m[2].str();
        if (re_search(inner, "\\b" + var + "\\b")) return cd_distribute(inner, var, outer);
    }
    // a sum of terms: the first term coef*var / var*coef / var is the linear one
    std::string e = expr;
    if (!e.empty() && e[0] != '+' && e[0] != '-') e = "+" + e;
    LinearForm f;
    bool found = false;
    std::vector<std::string> rest;
    for (const auto& tok : cd_terms(e)) {
        std::string sign, body;
        signed_body(tok, sign, body);
        bool hit = false;
        if (re_full(body, "^([a-zA-Z_]\\w*)\\s*\\*\\s*" + var + "$", &m)) { hit = true; f.coef = m[1].str(); }
        else if (re_full(body, "^" + var + "\\s*\\*\\s*([a-zA-Z_]\\w*)$", &m)) { hit = true; f.coef = m[1].str(); }
        else if (re_full(body, "^" + var + "$")) { hit = true; f.coef = "1"; }
        if (hit && !found) { found = true; f.sign = sign; }
        else rest.push_back(tok);
    }
    f.rest = join_rest(rest);
    return f;
}

// plain text -> C: x^n -> pow(x, n), ln( -> log(, names -> X[i] / a[1+j]
std::string cd_code(const std::string& expr, const System& s) {
    std::string c = expr;
    for (const auto& v : s.vars) c = re_replace(c, "\\b" + v + "\\s*\\^\\s*(\\d+)", "pow(" + v + ", $1)");
    c = re_replace(c, R"((X\[\d+\])\s*\^\s*(\d+))", "pow($1, $2)");
    c = re_replace(c, "\\bln\\s*\\(", "log(");
    for (size_t i = 0; i < s.vars.size(); ++i) c = re_replace(c, "\\b" + s.vars[i] + "\\b", "X[" + std::to_string(i) + "]");
    for (size_t j = 0; j < s.params.size(); ++j)
        c = re_replace(c, "\\b" + s.params[j] + "\\b", "a[" + std::to_string(1 + (int)j) + "]");
    return c;
}

std::string krs_cd(const System& s) {
    if (s.vars.size() != s.rhs.size()) throw std::runtime_error("vars/rhs size mismatch");
    const int N = (int)s.vars.size();
    if (N < 2) throw std::runtime_error("CD method requires N >= 2");
    std::vector<std::string> plain(N);
    for (int i = 0; i < N; ++i) {
        std::ostringstream p;
        print_plain(Parser(s.rhs[i], s.latex).parse(), p);
        plain[i] = p.str();
    }
    std::ostringstream o;
    o << "    numb h1 = h * a[0];\n    numb h2 = h * (1 - a[0]);\n";
    for (int i = 0; i < N; ++i) o << "    X[" << i << "] = X[" << i << "] + h1 * (" << cd_code(plain[i], s) << ");\n";
    for (int i = N - 1; i >= 0; --i) {
        const std::string x = "X[" + std::to_string(i) + "]";
        switch (cd_dependence(plain[i], s.vars[i])) {
        case Dependence::None:
            o << "    " << x << " = " << x << " + h2 * (" << cd_code(plain[i], s) << ");\n";
            break;
        case Dependence::Linear: {
            // f = sign*coef*x + rest  ->  x = (x + h2*rest) / (1 - sign*h2*coef)
            const LinearForm f = cd_linear_form(plain[i], s.vars[i]);
            const std::string coef = f.coef == "1" ? std::string("1") : cd_code(f.coef, s);
            const std::string den = std::string("(1 ") + (f.sign == "-" ? "+" : "-") + " h2" +
                                    (coef == "1" ? std::string(")") : " * " + coef + ")");
            if (f.rest.empty()) o << "    " << x << " = " << x << " / " << den << ";\n";
            else o << "    " << x << " = (" << x << " + h2 * (" << cd_code(f.rest, s) << ")) / " << den << ";\n";
            break;
        }
        case Dependence::Nonlinear: {
            const std::string tmp = "x" + std::to_string(i) + "_cd";
            const std::string code = cd_code(plain[i], s);
            o << "    numb " << tmp << " = " << x << ";\n";
            for (int it = 0; it < 4; ++it) o << "    " << x << " = " << tmp << " + h2 * (" << code << ");\n";
            break;
        }
        }
    }
    return o.str();
}

// ------------------------------------------------------ harmonic forcing --
// A cos / sin whose argument is linear in time, w * t + phi with w and phi
// free of the state and of t (periodic forcing), is an input of the fast
// modules' right-hand side: the kernels rotate its value from one stage
// time to the next instead of evaluating the function (the device source's
// uc_rhs_h, hc[k] / hs[k] = cos / sin(w[k] * t + phi[k])).
bool has_state_or_time(const P& n) {
    if (n->k == K::Var || n->k == K::Time) return true;
    if (n->a && has_state_or_time(n->a)) return true;
    if (n->b && has_state_or_time(n->b)) return true;
    for (auto& x : n->args)
        if (has_state_or_time(x)) return true;
    return false;
}
bool has_state(const P& n) {
    if (n->k == K::Var) return true;
    if (n->a && has_state(n->a)) return true;
    if (n->b && has_state(n->b)) return true;
    for (auto& x : n->args)
        if (has_state(x)) return true;
    return false;
}

// n = w * t + phi (phi null: 0) with w, phi free of the state and of t
bool linear_in_time(const P& n, P& w, P& phi) {
    P w0, p0;
    switch (n->k) {
    case K::Time: w = num(1.0); phi = nullptr; return true;
    case K::Neg:
        if (!linear_in_time(n->a, w0, p0)) return false;
        w = s_neg(w0);
        phi = p0 ? s_neg(p0) : nullptr;
        return true;
    case K::Mul:
    case K::Div: {
        const bool left = has_state_or_time(n->a), right = has_state_or_time(n->b);
        if (left == right) return false;
        if (n->k == K::Div && right) return false;  // t in a denominator
        const P& lin = left ? n->a : n->b;
        const P& c = left ? n->b : n->a;
        if (!linear_in_time(lin, w0, p0) || p0) return false;
        w = n->k == K::Div ? s_div(w0, c) : s_mul(w0, c);
        phi = nullptr;
        return true;
    }
    case K::Add:
    case K::Sub: {
        const bool left = has_state_or_time(n->a), right = has_state_or_time(n->b);
        if (left == right) return false;
        if (!linear_in_time(left ? n->a : n->b, w0, p0) || p0) return false;
        const P& c = left ? n->b : n->a;
        if (left) { w = w0; phi = n->k == K::Sub ? s_neg(c) : c; }
        else { w = n->k == K::Sub ? s_neg(w0) : w0; phi = c; }
        return true;
    }
    default: return false;
    }
}

struct Harmonics {
    std::vector<P> w, phi;
    std::map<std::string, int> index;
};

// n with every harmonic cos / sin replaced by hc[k] / hs[k]
P extract_harmonics(const P& n, Harmonics& h, const Style& st) {
    if (n->k == K::Call && (n->name == "cos" || n->name == "sin") && n->args.size() == 1 && !has_state(n->args[0])) {
        P w, phi;
        if (linear_in_time(n->args[0], w, phi)) {
            const std::string key = to_text(w, st) + "|" + (phi ? to_text(phi, st) : std::string());
            auto it = h.index.find(key);
            int k = (int)h.w.size();
            if (it == h.index.end()) {
                h.index[key] = k;
                h.w.push_back(w);
                h.phi.push_back(phi);
            } else {
                k = it->second;
            }
            auto r = std::make_shared<Node>();
            r->k = K::Raw;
            r->name = std::string(n->name == "cos" ? "hc[" : "hs[") + std::to_string(k) + "]";
            return r;
        }
    }
    switch (n->k) {
    case K::Neg: return neg(extract_harmonics(n->a, h, st));
    case K::Call: {
        std::vector<P> a;
        for (auto& x : n->args) a.push_back(extract_harmonics(x, h, st));
        return call(n->name, std::move(a));
    }
    case K::Add: case K::Sub: case K::Mul: case K::Div: case K::Pow:
        return bin(n->k, extract_harmonics(n->a, h, st), extract_harmonics(n->b, h, st));
    default: return n;
    }
}

}  // namespace

// ================================================================= public ==
std::string codegen_scheme(const System& s, Scheme sch) {
    switch (sch) {
    case Scheme::Euler: return krs_euler(s);
    case Scheme::EulerCromer: return krs_euler_cromer(s);
    case Scheme::ExplicitMidpoint: return krs_midpoint(s);
    case Scheme::RK4: return krs_rk4(s);
    case Scheme::DOPRI78: return krs_dopri78(s);
    case Scheme::CD: return krs_cd(s);
    }
    throw std::runtime_error("unknown scheme");
}

Scheme scheme_from_name(const std::string& name) {
    if (name == "Euler-Cromer") return Scheme::EulerCromer;
    if (name == "Explicit Midpoint") return Scheme::ExplicitMidpoint;
    if (name == "RK4") return Scheme::RK4;
    if (name == "DOPRI78") return Scheme::DOPRI78;
    if (name == "CD") return Scheme::CD;
    return Scheme::Euler;
}

std::string normalize_value(const std::string& value) {
    size_t a = value.find_first_not_of(" \t\n\r");
    if (a == std::string::npos) return "";
    size_t b = value.find_last_not_of(" \t\n\r");
    P t = Parser(value.substr(a, b - a + 1), false).parse();
    Style st;
    std::function<void(const P&)> check = [&](const P& n) {
        // values carry no names: every symbol (pi included) is an error, as in
        // the reference (it prints through an empty name map)
        if (n->k == K::Const) throw std::runtime_error("unknown symbol: " + n->name);
        if (n->a) check(n->a);
        if (n->b) check(n->b);
        for (auto& x : n->args) check(x);
    };
    check(t);
    return to_text(t, st);
}

std::string emit_system_source(const System& s, Dialect d) {
    const int N = (int)s.vars.size();
    const int M = 1 + (int)s.params.size();
    if (N <= 0) throw std::runtime_error("system has no state variables");
    auto trees = parse_system(s);

    Style st;
    st.typed_literals = true;
    st.c_functions = (d == Dialect::C);

    bool uses_time = false;
    std::function<void(const P&)> scan = [&](const P& n) {
        if (n->k == K::Time) uses_time = true;
        if (n->a) scan(n->a);
        if (n->b) scan(n->b);
        for (auto& x : n->args) scan(x);
    };
    for (auto& t : trees) scan(t);

    // Jacobian entries (row i, column j).
    std::vector<std::vector<P>> J(N, std::vector<P>(N));
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) J[i][j] = diff(trees[i], j);

    std::ostringstream o;
    o << "#define UC_N " << N << "\n#define UC_M " << M << "\n#define UC_USES_TIME " << (uses_time ? 1 : 0)
      << "\n#define UC_CUSTOM_KRS 0\n";
    const bool hip = (d == Dialect::Hip);
    const char* fn = hip ? "template <typename T>\n__device__ __forceinline__ void" : "static inline void";
    if (!hip) o << "typedef double T;\n";

    // right-hand side
    o << fn << " uc_rhs(const T* __restrict__ X, const T* __restrict__ a, const T t, T* __restrict__ dX) {\n";
    o << "    (void)a; (void)t;\n";
    for (int i = 0; i < N; ++i) o << "    dX[" << i << "] = " << to_text(trees[i], st) << ";\n";
    o << "}\n";

    if (hip) {
        // the right-hand side again with its harmonic forcing terms as inputs
        Harmonics hm;
        std::vector<P> forced;
        for (auto& t : trees) forced.push_back(extract_harmonics(t, hm, st));
        const int nh = (int)hm.w.size();
        o << "#define UC_HARMONICS " << nh << "\n";
        if (nh > 0) {
            o << "template <typename T>\n__device__ __forceinline__ void uc_harm_coef(const T* __restrict__ a, T* w, T* phi) "
                 "{\n    (void)a;\n";
            for (int k = 0; k < nh; ++k)
                o << "    w[" << k << "] = " << to_text(hm.w[k], st) << ";\n    phi[" << k
                  << "] = " << (hm.phi[k] ? to_text(hm.phi[k], st) : std::string("(T)0")) << ";\n";
            o << "}\n";
            o << "template <typename T>\n__device__ __forceinline__ void uc_rhs_h(const T* __restrict__ X, "
                 "const T* __restrict__ a, const T t, const T* hc, const T* hs, T* __restrict__ dX) {\n";
            o << "    (void)X; (void)a; (void)t; (void)hc; (void)hs;\n";
            for (int i = 0; i < N; ++i) o << "    dX[" << i << "] = " << to_text(forced[i], st) << ";\n";
            o << "}\n";
        }
    }

    // tangent: Jacobian entries once, then JV for each of the K vectors
    if (hip) o << "template <typename T, int K>\n__device__ __forceinline__ void";
    else o << "static inline void";
    o << " uc_tangent(const T* __restrict__ X, const T* __restrict__ a, const T t, const T* __restrict__ V, T* __restrict__ JV"
      << (hip ? "" : ", const int K") << ") {\n";
    o << "    (void)X; (void)a; (void)t;\n";
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            const P& e = J[i][j];
            if (is_zero(e) || e->k == K::Num || (e->k == K::Neg && e->a->k == K::Num)) continue;
            o << "    const T J" << i << "_" << j << " = " << to_text(e, st) << ";\n";
        }
    if (hip) o << "#pragma unroll\n";
    o << "    for (int k = 0; k < K; ++k) {\n";
    for (int i = 0; i < N; ++i) {
        std::ostringstream terms;
        bool first = true;
        for (int j = 0; j < N; ++j) {
            const P& e = J[i][j];
            if (is_zero(e)) continue;
            std::string v = "V[k * " + std::to_string(N) + " + " + std::to_string(j) + "]";
            std::string term;
            bool negative = false;
            if (is_num(e, 1.0)) term = v;
            else if (is_num(e, -1.0)) { term = v; negative = true; }
            else if (e->k == K::Num) term = "((T)" + format_number(e->num) + ") * " + v;
            else if (e->k == K::Neg && e->a->k == K::Num) term = "((T)" + format_number(-e->a->num) + ") * " + v;
            else term = "J" + std::to_string(i) + "_" + std::to_string(j) + " * " + v;
            if (first) terms << (negative ? "-" : "") << term;
            else terms << (negative ? " - " : " + ") << term;
            first = false;
        }
        o << "        JV[k * " << N << " + " << i << "] = " << (first ? std::string("(T)0") : terms.str()) << ";\n";
    }
    o << "    }\n}\n";
    return o.str();
}

std::string emit_custom_krs_source(const std::string& body, int n_vars, int n_values) {
    std::ostringstream o;
    o << "#define UC_N " << n_vars << "\n#define UC_M " << n_values
      << "\n#define UC_USES_TIME 0\n#define UC_CUSTOM_KRS 1\n#define UC_HARMONICS 0\n";
    // the step body in the reference prelude's environment (U-CUDA/krs_cpu.cpp:293-316):
    // numb, pi and euler of that type (a body may shadow them), min / max
    o << "template <typename numb>\n__device__ __forceinline__ void uc_krs_step(numb* X, const numb* a, const numb h) {\n"
         "    const numb pi = (numb)3.1415926535897932384626433832795;\n"
         "    const numb euler = (numb)2.7182818284590452353602874713527;\n"
         "    (void)pi; (void)euler; (void)a; (void)h;\n    {\n"
      << body << "\n    }\n}\n";
    // no right-hand side / Jacobian: the kernels that need them are rejected by the engine
    o << "template <typename T>\n__device__ __forceinline__ void uc_rhs(const T*, const T*, const T, T* dX) {\n"
         "    for (int i = 0; i < UC_N; ++i) dX[i] = (T)0;\n}\n";
    o << "template <typename T, int K>\n__device__ __forceinline__ void uc_tangent(const T*, const T*, const T, const T*, T* JV) {\n"
         "    for (int i = 0; i < K * UC_N; ++i) JV[i] = (T)0;\n}\n";
    return o.str();
}

// Constant subscripts of X[] / a[] must be in range (krs_cpu_check_indices,
// U-CUDA/krs_cpu.cpp:21-99): comments and literals are skipped, identifiers
// are taken whole, only integer-literal subscripts are checked.
std::string check_krs_indices(const std::string& b, int nx, int nv) {
    std::ostringstream diag;
    const size_t n = b.size();
    int line = 1;
    for (size_t i = 0; i < n;) {
        const char c = b[i];
        if (c == '\n') { ++line; ++i; continue; }
        if (c == '/' && i + 1 < n && b[i + 1] == '/') { while (i < n && b[i] != '\n') ++i; continue; }
        if (c == '/' && i + 1 < n && b[i + 1] == '*') {
            i += 2;
            while (i + 1 < n && !(b[i] == '*' && b[i + 1] == '/')) { if (b[i] == '\n') ++line; ++i; }
            i = (i + 1 < n) ? i + 2 : n;
            continue;
        }
        if (c == '"' || c == '\'') {
            const char q = c;
            ++i;
            while (i < n && b[i] != q) { if (b[i] == '\\') ++i; else if (b[i] == '\n') ++line; ++i; }
            ++i;
            continue;
        }
        if (!(std::isalpha((unsigned char)c) || c == '_')) { ++i; continue; }
        const size_t st = i;
        while (i < n && (std::isalnum((unsigned char)b[i]) || b[i] == '_')) ++i;
        const std::string id = b.substr(st, i - st);
        if (id != "X" && id != "a") continue;
        size_t j = i;
        while (j < n && (b[j] == ' ' || b[j] == '\t')) ++j;
        if (j >= n || b[j] != '[') continue;
        ++j;
        while (j < n && (b[j] == ' ' || b[j] == '\t')) ++j;
        const size_t ds = j;
        while (j < n && std::isdigit((unsigned char)b[j])) ++j;
        if (j == ds) continue;
        const size_t de = j;
        while (j < n && (b[j] == ' ' || b[j] == '\t')) ++j;
        if (j >= n || b[j] != ']') continue;
        const long idx = std::strtol(b.substr(ds, de - ds).c_str(), nullptr, 10);
        if (id == "X" && (idx < 0 || idx >= nx)) diag << "line " << line << ": X[" << idx << "] out of range (" << nx << " variables)\n";
        if (id == "a" && (idx < 0 || idx >= nv)) diag << "line " << line << ": a[" << idx << "] out of range (" << nv << " values)\n";
    }
    return diag.str();
}

OpCounts count_ops(const System& s) {
    OpCounts c;
    auto trees = parse_system(s);
    const int N = (int)s.vars.size();
    for (auto& t : trees) count(t, c.rhs_flops, c.rhs_calls);
    for (int i = 0; i < N; ++i) {
        int row_terms = 0;
        for (int j = 0; j < N; ++j) {
            P e = diff(trees[i], j);
            if (is_zero(e)) continue;
            ++c.jac_nonzeros;
            ++row_terms;
            if (!(e->k == K::Num || (e->k == K::Neg && e->a->k == K::Num))) count(e, c.jac_flops, c.jac_calls);
        }
        // per tangent vector: one multiply per nonzero (unit entries are free) plus adds
        (void)row_terms;
    }
    return c;
}

void eval_rhs(const System& s, const double* X, const double* a, double t, double* dX) {
    auto trees = parse_system(s);
    for (size_t i = 0; i < trees.size(); ++i) dX[i] = eval(trees[i], X, a, t);
}

}  // namespace ucuda
