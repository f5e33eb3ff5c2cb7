// qasm.cpp — OpenQASM 2.0 reader producing the zr_instr program of the
// compiler (the subset and rules of src/qasm.rs / docs/QASM.md: qreg/creg,
// the native gate table plus U/CX builtins, user gate definitions expanded
// inline, register broadcasting, measure/reset/barrier, if (creg == n),
// parameter expressions evaluated in f64 at parse time).  Errors are
// reported as "line L:C: message" through the library's last-error string.
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "host.hpp"
#include "zeno_hip.h"

using namespace zeno;

namespace {

constexpr uint32_t kMaxQubits = 48;  // front-end limit (docs/QASM.md)
constexpr int kMaxDepth = 128;       // gate-expansion nesting

struct QasmError : std::runtime_error {
    QasmError(int line, int col, const std::string &msg)
        : std::runtime_error("line " + std::to_string(line) + ":" + std::to_string(col) + ": " + msg) {}
};

// ------------------------------------------------------------------ lexer
enum class T { Ident, Int, Real, Str, Sym, End };

struct Tok {
    T kind;
    std::string text;  // identifier, symbol, string contents or number text
    int line, col;
};

std::vector<Tok> lex(const std::string &src) {
    std::vector<Tok> out;
    size_t i = 0;
    int line = 1, col = 1;
    auto adv = [&](size_t n = 1) {
        for (size_t k = 0; k < n && i < src.size(); ++k, ++i) {
            if (src[i] == '\n') ++line, col = 1;
            else ++col;
        }
    };
    while (i < src.size()) {
        const char c = src[i];
        if (c == ' ' || c == '\t' || c == '\r' || c == '\n') {
            adv();
            continue;
        }
        if (c == '/' && i + 1 < src.size() && src[i + 1] == '/') {
            while (i < src.size() && src[i] != '\n') adv();
            continue;
        }
        const int l = line, cl = col;
        if (std::isalpha((unsigned char)c) || c == '_') {
            size_t j = i;
            while (j < src.size() && (std::isalnum((unsigned char)src[j]) || src[j] == '_')) ++j;
            out.push_back({T::Ident, src.substr(i, j - i), l, cl});
            adv(j - i);
            continue;
        }
        if (std::isdigit((unsigned char)c) || (c == '.' && i + 1 < src.size() && std::isdigit((unsigned char)src[i + 1]))) {
            size_t j = i;
            bool real = false;
            while (j < src.size() && std::isdigit((unsigned char)src[j])) ++j;
            if (j < src.size() && src[j] == '.') {
                real = true;
                ++j;
                while (j < src.size() && std::isdigit((unsigned char)src[j])) ++j;
            }
            if (j < src.size() && (src[j] == 'e' || src[j] == 'E')) {
                size_t k = j + 1;
                if (k < src.size() && (src[k] == '+' || src[k] == '-')) ++k;
                if (k < src.size() && std::isdigit((unsigned char)src[k])) {
                    real = true;
                    j = k;
                    while (j < src.size() && std::isdigit((unsigned char)src[j])) ++j;
                } else {
                    throw QasmError(l, cl, "malformed exponent in number '" + src.substr(i, k - i) + "'");
                }
            }
            out.push_back({real ? T::Real : T::Int, src.substr(i, j - i), l, cl});
            adv(j - i);
            continue;
        }
        if (c == '"') {
            size_t j = i + 1;
            while (j < src.size() && src[j] != '"' && src[j] != '\n') ++j;
            if (j >= src.size() || src[j] != '"') throw QasmError(l, cl, "unterminated string");
            out.push_back({T::Str, src.substr(i + 1, j - i - 1), l, cl});
            adv(j + 1 - i);
            continue;
        }
        if (c == '-' && i + 1 < src.size() && src[i + 1] == '>') {
            out.push_back({T::Sym, "->", l, cl});
            adv(2);
            continue;
        }
        if (c == '=') {
            if (i + 1 < src.size() && src[i + 1] == '=') {
                out.push_back({T::Sym, "==", l, cl});
                adv(2);
                continue;
            }
            throw QasmError(l, cl, "found a single '='; the only equality operator is '==' (as in 'if (c == 1)')");
        }
        if (std::strchr(";,()[]{}+-*/^", c)) {
            out.push_back({T::Sym, std::string(1, c), l, cl});
            adv();
            continue;
        }
        throw QasmError(l, cl, std::string("unexpected character '") + c + "'");
    }
    out.push_back({T::End, "", line, col});
    return out;
}

// ------------------------------------------------------------ expressions
struct Expr {
    enum K { Num, Param, Neg, Add, Sub, Mul, Div, Pow, Fn } k = Num;
    double v = 0;
    int param = 0;
    std::string fn;
    std::unique_ptr<Expr> a, b;
    double eval(const std::vector<double> &p) const {
        switch (k) {
            case Num: return v;
            case Param: return p[(size_t)param];
            case Neg: return -a->eval(p);
            case Add: return a->eval(p) + b->eval(p);
            case Sub: return a->eval(p) - b->eval(p);
            case Mul: return a->eval(p) * b->eval(p);
            case Div: return a->eval(p) / b->eval(p);
            case Pow: return std::pow(a->eval(p), b->eval(p));
            case Fn: {
                const double x = a->eval(p);
                if (fn == "sin") return std::sin(x);
                if (fn == "cos") return std::cos(x);
                if (fn == "tan") return std::tan(x);
                if (fn == "exp") return std::exp(x);
                if (fn == "ln") return std::log(x);
                return std::sqrt(x);
            }
        }
        return 0;
    }
};
using ExprP = std::unique_ptr<Expr>;

ExprP mk_bin(Expr::K k, ExprP a, ExprP b) {
    auto e = std::make_unique<Expr>();
    e->k = k;
    e->a = std::move(a);
    e->b = std::move(b);
    return e;
}

// ---------------------------------------------------------------- parser
struct Reg {
    std::string name;
    uint32_t offset, size;
};

// An argument: a whole register (size > 0, index < 0) or one bit.
struct Arg {
    const Reg *reg;
    int64_t index;  // -1: whole register
    int line, col;
};

// A statement inside a gate body, kept unexpanded.
struct BodyStmt {
    bool barrier = false;
    std::string gate;
    std::vector<ExprP> params;
    std::vector<int> qargs;  // indices into the definition's formal qubits
    int line = 0, col = 0;
};

struct GateDef {
    std::vector<std::string> params, qargs;
    std::vector<BodyStmt> body;
};

const std::set<std::string> kReserved = {"OPENQASM", "include", "qreg", "creg",  "gate", "opaque", "measure",
                                         "reset",    "barrier", "if",   "pi",    "U",    "CX"};
const std::set<std::string> kFuncs = {"sin", "cos", "tan", "exp", "ln", "sqrt"};

class Parser {
  public:
    explicit Parser(std::vector<Tok> toks) : t_(std::move(toks)) {}

    void run() {
        header();
        while (peek().kind != T::End) statement();
    }

    std::vector<zr_instr> instrs;
    std::vector<Reg> qregs, cregs;

  private:
    std::vector<Tok> t_;
    size_t p_ = 0;
    std::map<std::string, GateDef> gates_;

    const Tok &peek(size_t k = 0) const { return t_[std::min(p_ + k, t_.size() - 1)]; }
    Tok next() {
        Tok t = peek();
        if (p_ < t_.size() - 1) ++p_;
        return t;
    }
    static std::string show(const Tok &t) { return t.kind == T::End ? "end of input" : "'" + t.text + "'"; }
    [[noreturn]] void fail(const Tok &t, const std::string &m) const { throw QasmError(t.line, t.col, m); }
    bool is_sym(const char *s, size_t k = 0) const { return peek(k).kind == T::Sym && peek(k).text == s; }
    void expect_sym(const char *s, const std::string &what) {
        if (!is_sym(s)) fail(peek(), "expected '" + std::string(s) + "' " + what + ", found " + show(peek()));
        next();
    }
    Tok expect_ident(const std::string &what) {
        if (peek().kind != T::Ident) fail(peek(), "expected " + what + ", found " + show(peek()));
        return next();
    }
    uint64_t expect_uint(uint64_t max, const std::string &what) {
        const Tok t = peek();
        if (t.kind != T::Int) fail(t, "expected " + what + " (an unsigned integer), found " + show(t));
        next();
        errno = 0;
        char *end = nullptr;
        const unsigned long long v = std::strtoull(t.text.c_str(), &end, 10);
        if (errno == ERANGE || v > max) fail(t, what + " " + t.text + " is too large (at most " + std::to_string(max) + ")");
        return v;
    }

    uint32_t total(const std::vector<Reg> &rs) const {
        uint32_t n = 0;
        for (const Reg &r : rs) n += r.size;
        return n;
    }
    const Reg *find(const std::vector<Reg> &rs, const std::string &n) const {
        for (const Reg &r : rs)
            if (r.name == n) return &r;
        return nullptr;
    }

    void header() {
        const Tok t = peek();
        if (t.kind != T::Ident || t.text != "OPENQASM") fail(t, "expected 'OPENQASM 2.0;' as the first statement, found " + show(t));
        next();
        const Tok v = next();
        if (!((v.kind == T::Real && (v.text == "2.0")) || (v.kind == T::Int && v.text == "2")))
            fail(v, "unsupported OpenQASM version " + show(v) + " (this reader handles 2.0)");
        expect_sym(";", "after the version");
    }

    void statement() {
        const Tok t = peek();
        if (t.kind != T::Ident) fail(t, "expected a statement, found " + show(t));
        if (t.text == "include") return include();
        if (t.text == "qreg" || t.text == "creg") return reg_decl();
        if (t.text == "gate") return gate_def();
        if (t.text == "opaque") fail(t, "'opaque' gates have no simulation semantics and are not supported");
        if (t.text == "barrier") return barrier();
        if (t.text == "if") return if_stmt();
        if (t.text == "OPENQASM") fail(t, "'OPENQASM' may appear only once, as the first statement");
        for (zr_instr &i : qop()) instrs.push_back(i);
    }

    void include() {
        next();
        const Tok s = next();
        if (s.kind != T::Str) fail(s, "expected a quoted file name after 'include', found " + show(s));
        if (s.text != "qelib1.inc") fail(s, "only \"qelib1.inc\" can be included (got \"" + s.text + "\")");
        expect_sym(";", "after the include");
    }

    void reg_decl() {
        const Tok kw = next();
        const Tok name = expect_ident("a register name");
        if (kReserved.count(name.text)) fail(name, "'" + name.text + "' is a reserved word");
        if (find(qregs, name.text) || find(cregs, name.text)) fail(name, "register '" + name.text + "' is already declared");
        expect_sym("[", "after the register name");
        const Tok st = peek();
        const uint64_t n = expect_uint(1u << 20, "register size");
        if (n == 0) fail(st, "register size must be at least 1");
        expect_sym("]", "after the register size");
        expect_sym(";", "after the register declaration");
        std::vector<Reg> &rs = kw.text == "qreg" ? qregs : cregs;
        const uint32_t off = total(rs);
        if (kw.text == "qreg" && off + n > kMaxQubits)
            fail(name, "too many qubits: " + std::to_string(off + n) + " (at most " + std::to_string(kMaxQubits) + ")");
        rs.push_back({name.text, off, (uint32_t)n});
    }

    Arg reg_arg(bool quantum) {
        const Tok name = expect_ident(quantum ? "a qubit argument" : "a classical argument");
        const Reg *r = find(quantum ? qregs : cregs, name.text);
        if (!r) {
            if (find(quantum ? cregs : qregs, name.text))
                fail(name, "'" + name.text + "' is a " + (quantum ? "classical" : "quantum") + " register, expected a " +
                               (quantum ? "quantum" : "classical") + " one");
            fail(name, "unknown " + std::string(quantum ? "quantum" : "classical") + " register '" + name.text + "'");
        }
        Arg a{r, -1, name.line, name.col};
        if (is_sym("[")) {
            next();
            const Tok it = peek();
            const uint64_t i = expect_uint(UINT32_MAX, "index");
            if (i >= r->size)
                fail(it, "index " + std::to_string(i) + " out of range for register '" + r->name + "' of size " +
                             std::to_string(r->size));
            expect_sym("]", "after the index");
            a.index = (int64_t)i;
        }
        return a;
    }

    std::string qubit_name(uint32_t q) const {
        for (const Reg &r : qregs)
            if (q >= r.offset && q < r.offset + r.size) return r.name + "[" + std::to_string(q - r.offset) + "]";
        return std::to_string(q);
    }

    // Broadcast length of a list of arguments (1 without whole registers).
    uint32_t bcast(const std::vector<Arg> &args, const Tok &at) const {
        const Arg *first = nullptr;
        for (const Arg &a : args)
            if (a.index < 0) {
                if (!first) first = &a;
                else if (a.reg->size != first->reg->size)
                    fail(at, "broadcast size mismatch: register '" + first->reg->name + "' has size " +
                                 std::to_string(first->reg->size) + " but register '" + a.reg->name + "' has size " +
                                 std::to_string(a.reg->size));
            }
        return first ? first->reg->size : 1;
    }
    static uint32_t at(const Arg &a, uint32_t i) { return a.reg->offset + (a.index < 0 ? i : (uint32_t)a.index); }

    // --- expressions (docs/QASM.md grammar)
    ExprP expr(const std::vector<std::string> &scope) {
        ExprP e = term(scope);
        while (is_sym("+") || is_sym("-")) {
            const bool add = next().text == "+";
            e = mk_bin(add ? Expr::Add : Expr::Sub, std::move(e), term(scope));
        }
        return e;
    }
    ExprP term(const std::vector<std::string> &scope) {
        ExprP e = unary(scope);
        while (is_sym("*") || is_sym("/")) {
            const bool mul = next().text == "*";
            e = mk_bin(mul ? Expr::Mul : Expr::Div, std::move(e), unary(scope));
        }
        return e;
    }
    ExprP unary(const std::vector<std::string> &scope) {
        if (is_sym("-")) {
            next();
            auto e = std::make_unique<Expr>();
            e->k = Expr::Neg;
            e->a = unary(scope);
            return e;
        }
        return power(scope);
    }
    ExprP power(const std::vector<std::string> &scope) {
        ExprP e = atom(scope);
        if (is_sym("^")) {
            next();
            e = mk_bin(Expr::Pow, std::move(e), unary(scope));
        }
        return e;
    }
    ExprP atom(const std::vector<std::string> &scope) {
        const Tok t = peek();
        auto e = std::make_unique<Expr>();
        if (t.kind == T::Int || t.kind == T::Real) {
            next();
            e->v = std::strtod(t.text.c_str(), nullptr);
            return e;
        }
        if (is_sym("(")) {
            next();
            ExprP in = expr(scope);
            expect_sym(")", "to close the parenthesis");
            return in;
        }
        if (t.kind == T::Ident) {
            next();
            if (t.text == "pi") {
                e->v = M_PI;
                return e;
            }
            if (kFuncs.count(t.text)) {
                expect_sym("(", "after function '" + t.text + "'");
                e->k = Expr::Fn;
                e->fn = t.text;
                e->a = expr(scope);
                expect_sym(")", "to close the function call");
                return e;
            }
            for (size_t i = 0; i < scope.size(); ++i)
                if (scope[i] == t.text) {
                    e->k = Expr::Param;
                    e->param = (int)i;
                    return e;
                }
            std::string in_scope = "pi, sin, cos, tan, exp, ln, sqrt";
            for (const auto &s : scope) in_scope += ", " + s;
            fail(t, "unknown identifier '" + t.text + "' in an expression (in scope: " + in_scope + ")");
        }
        fail(t, "expected an expression, found " + show(t));
    }

    std::vector<ExprP> param_list(const std::vector<std::string> &scope) {
        std::vector<ExprP> out;
        if (!is_sym("(")) return out;
        next();
        if (is_sym(")")) {
            next();
            return out;
        }
        out.push_back(expr(scope));
        while (is_sym(",")) {
            next();
            out.push_back(expr(scope));
        }
        expect_sym(")", "to close the parameter list");
        return out;
    }

    // Resolves a gate name: native (with U/CX/u builtins) or user-defined.
    // Returns arity and parameter count; user gates give def != nullptr.
    void resolve(const Tok &name, std::string *native, int *arity, int *np, const GateDef **def) const {
        *def = nullptr;
        std::string n = name.text;
        if (n == "U") n = "u3";
        else if (n == "CX") n = "cx";
        if (gate_lookup(n.c_str(), arity, np) == 0) {
            *native = n;
            return;
        }
        auto it = gates_.find(name.text);
        if (it != gates_.end()) {
            *def = &it->second;
            *arity = (int)it->second.qargs.size();
            *np = (int)it->second.params.size();
            return;
        }
        std::string lower = name.text;
        for (char &c : lower) c = (char)std::tolower((unsigned char)c);
        int a2, n2;
        if (lower != name.text && gate_lookup(lower.c_str(), &a2, &n2) == 0)
            fail(name, "unknown gate '" + name.text + "' (gate names are case-sensitive; did you mean '" + lower + "'?)");
        fail(name, "unknown gate '" + name.text + "'");
    }

    static zr_instr gate_instr(const std::string &name, const std::vector<double> &p, const std::vector<uint32_t> &q) {
        zr_instr r{};
        std::strncpy(r.name, name.c_str(), sizeof(r.name) - 1);
        r.kind = 0;
        r.nq = (int32_t)q.size();
        r.np = (int32_t)p.size();
        for (size_t i = 0; i < q.size(); ++i) r.q[i] = q[i];
        for (size_t i = 0; i < p.size(); ++i) r.p[i] = p[i];
        r.cond_creg = -1;
        return r;
    }
    static zr_instr barrier_instr() {
        zr_instr r{};
        r.kind = 3;
        r.cond_creg = -1;
        return r;
    }

    // Expands one application (native or user-defined) on concrete qubits.
    void apply(const std::string &native, const GateDef *def, const std::vector<double> &p,
               const std::vector<uint32_t> &q, std::vector<zr_instr> &out, int depth, const Tok &at) const {
        if (!def) {
            out.push_back(gate_instr(native, p, q));
            return;
        }
        if (depth >= kMaxDepth) fail(at, "gate expansion nested deeper than " + std::to_string(kMaxDepth));
        for (const BodyStmt &s : def->body) {
            if (s.barrier) {
                out.push_back(barrier_instr());
                continue;
            }
            std::vector<double> pv;
            for (const ExprP &e : s.params) pv.push_back(e->eval(p));
            std::vector<uint32_t> qv;
            for (int i : s.qargs) qv.push_back(q[(size_t)i]);
            std::string nat;
            int a, np;
            const GateDef *d;
            resolve(Tok{T::Ident, s.gate, s.line, s.col}, &nat, &a, &np, &d);
            apply(nat, d, pv, qv, out, depth + 1, at);
        }
    }

    // gate call, measure or reset (a qop), possibly broadcast.
    std::vector<zr_instr> qop() {
        const Tok t = peek();
        if (t.text == "measure") return measure();
        if (t.text == "reset") return reset();
        if (kReserved.count(t.text) && t.text != "U" && t.text != "CX")
            fail(t, "'" + t.text + "' cannot be used here");
        return gate_call();
    }

    std::vector<zr_instr> gate_call() {
        const Tok name = next();
        std::string native;
        int arity = 0, np = 0;
        const GateDef *def = nullptr;
        resolve(name, &native, &arity, &np, &def);
        std::vector<ExprP> pe = param_list({});
        if ((int)pe.size() != np)
            fail(name, "gate '" + name.text + "' takes " + std::to_string(np) + " parameter(s), got " +
                           std::to_string(pe.size()));
        std::vector<Arg> args;
        args.push_back(reg_arg(true));
        while (is_sym(",")) {
            next();
            args.push_back(reg_arg(true));
        }
        if ((int)args.size() != arity)
            fail(name, "gate '" + name.text + "' acts on " + std::to_string(arity) + " qubit(s), got " +
                           std::to_string(args.size()));
        expect_sym(";", "after the gate call");
        std::vector<double> pv;
        for (const ExprP &e : pe) pv.push_back(e->eval({}));
        const uint32_t n = bcast(args, name);
        std::vector<zr_instr> out;
        for (uint32_t i = 0; i < n; ++i) {
            std::vector<uint32_t> q;
            for (const Arg &a : args) q.push_back(at(a, i));
            for (size_t x = 0; x < q.size(); ++x)
                for (size_t y = x + 1; y < q.size(); ++y)
                    if (q[x] == q[y])
                        fail(name, "gate '" + name.text + "' applied to duplicate qubit '" + qubit_name(q[x]) +
                                       "' after broadcasting");
            apply(native, def, pv, q, out, 0, name);
        }
        return out;
    }

    std::vector<zr_instr> measure() {
        const Tok kw = next();
        const Arg q = reg_arg(true);
        expect_sym("->", "after the measured qubit");
        const Arg c = reg_arg(false);
        expect_sym(";", "after the measurement");
        const bool qw = q.index < 0 && q.reg->size > 1, cw = c.index < 0 && c.reg->size > 1;
        std::vector<zr_instr> out;
        auto push = [&](uint32_t qb, uint32_t cb) {
            zr_instr r{};
            r.kind = 1;
            r.nq = 1;
            r.q[0] = qb;
            r.clbit = cb;
            r.cond_creg = -1;
            out.push_back(r);
        };
        if (qw || cw) {
            if (!(qw && cw))
                fail(kw, "measure maps a register to a register of the same size, or a bit to a bit");
            if (q.reg->size != c.reg->size)
                fail(kw, "measure size mismatch: register '" + q.reg->name + "' has size " + std::to_string(q.reg->size) +
                             " but register '" + c.reg->name + "' has size " + std::to_string(c.reg->size));
            for (uint32_t i = 0; i < q.reg->size; ++i) push(q.reg->offset + i, c.reg->offset + i);
        } else {
            push(at(q, 0), at(c, 0));
        }
        return out;
    }

    std::vector<zr_instr> reset() {
        next();
        const Arg q = reg_arg(true);
        expect_sym(";", "after the reset");
        std::vector<zr_instr> out;
        const uint32_t n = q.index < 0 ? q.reg->size : 1;
        for (uint32_t i = 0; i < n; ++i) {
            zr_instr r{};
            r.kind = 2;
            r.nq = 1;
            r.q[0] = at(q, i);
            r.cond_creg = -1;
            out.push_back(r);
        }
        return out;
    }

    void barrier() {
        next();
        reg_arg(true);
        while (is_sym(",")) {
            next();
            reg_arg(true);
        }
        expect_sym(";", "after the barrier");
        instrs.push_back(barrier_instr());  // a full fence in the compiler whatever its qubits
    }

    void if_stmt() {
        const Tok kw = next();
        expect_sym("(", "after 'if'");
        const Tok name = expect_ident("a classical register");
        const Reg *r = find(cregs, name.text);
        if (!r) fail(name, "unknown classical register '" + name.text + "'");
        if (is_sym("[")) fail(peek(), "'if' compares a whole classical register; it cannot be indexed");
        expect_sym("==", "after the register in 'if'");
        const uint64_t v = expect_uint(UINT64_MAX, "comparison value");
        expect_sym(")", "to close the 'if' condition");
        const Tok op = peek();
        if (op.kind == T::Ident && (op.text == "barrier" || op.text == "if"))
            fail(op, "the body of 'if' must be a gate call, measure or reset");
        const int32_t idx = (int32_t)(r - cregs.data());
        for (zr_instr i : qop()) {
            if (i.kind == 3) fail(kw, "the body of 'if' must be a gate call, measure or reset");
            i.cond_creg = idx;
            i.cond_value = v;
            instrs.push_back(i);
        }
    }

    void gate_def() {
        next();
        const Tok name = expect_ident("a gate name");
        if (kReserved.count(name.text)) fail(name, "'" + name.text + "' is a reserved word");
        int a, np;
        if (gate_lookup(name.text.c_str(), &a, &np) == 0)
            fail(name, "gate '" + name.text + "' would shadow the native gate of that name");
        if (gates_.count(name.text)) fail(name, "gate '" + name.text + "' is already defined");
        GateDef def;
        if (is_sym("(")) {
            next();
            if (!is_sym(")")) {
                def.params.push_back(expect_ident("a parameter name").text);
                while (is_sym(",")) {
                    next();
                    def.params.push_back(expect_ident("a parameter name").text);
                }
            }
            expect_sym(")", "to close the parameter list");
        }
        def.qargs.push_back(expect_ident("a qubit argument name").text);
        while (is_sym(",")) {
            next();
            def.qargs.push_back(expect_ident("a qubit argument name").text);
        }
        expect_sym("{", "to open the gate body");
        while (!is_sym("}")) {
            const Tok t = peek();
            if (t.kind == T::End) fail(t, "unterminated body of gate '" + name.text + "'");
            if (t.kind != T::Ident) fail(t, "expected a gate call or 'barrier' in the body of '" + name.text + "', found " + show(t));
            BodyStmt s;
            s.line = t.line;
            s.col = t.col;
            if (t.text == "barrier") {
                next();
                s.barrier = true;
                formal(def, name.text);
                while (is_sym(",")) {
                    next();
                    formal(def, name.text);
                }
            } else {
                if (t.text == name.text) fail(t, "gate '" + name.text + "' cannot call itself (recursion is not allowed)");
                if (t.text == "measure" || t.text == "reset" || t.text == "if")
                    fail(t, "'" + t.text + "' is not allowed inside a gate body");
                next();
                std::string nat;
                int ar, npar;
                const GateDef *d;
                resolve(t, &nat, &ar, &npar, &d);
                s.gate = t.text;
                s.params = param_list(def.params);
                if ((int)s.params.size() != npar)
                    fail(t, "gate '" + t.text + "' takes " + std::to_string(npar) + " parameter(s), got " +
                                std::to_string(s.params.size()));
                s.qargs.push_back(formal(def, name.text));
                while (is_sym(",")) {
                    next();
                    s.qargs.push_back(formal(def, name.text));
                }
                if ((int)s.qargs.size() != ar)
                    fail(t, "gate '" + t.text + "' acts on " + std::to_string(ar) + " qubit(s), got " +
                                std::to_string(s.qargs.size()));
                for (size_t x = 0; x < s.qargs.size(); ++x)
                    for (size_t y = x + 1; y < s.qargs.size(); ++y)
                        if (s.qargs[x] == s.qargs[y])
                            fail(t, "gate '" + t.text + "' applied to duplicate qubit '" + def.qargs[(size_t)s.qargs[x]] + "'");
            }
            expect_sym(";", "after the statement in the gate body");
            def.body.push_back(std::move(s));
        }
        next();
        gates_[name.text] = std::move(def);
    }

    int formal(const GateDef &def, const std::string &gate) {
        const Tok t = expect_ident("a qubit argument of gate '" + gate + "'");
        for (size_t i = 0; i < def.qargs.size(); ++i)
            if (def.qargs[i] == t.text) {
                if (is_sym("[")) fail(peek(), "gate arguments are single qubits and cannot be indexed");
                return (int)i;
            }
        fail(t, "unknown qubit argument '" + t.text + "' in the body of gate '" + gate + "'");
    }
};

}  // namespace

extern "C" int64_t zr_parse_qasm(const char *src, zr_instr *out, int64_t cap, uint32_t *n_qubits, uint32_t *cregs,
                                 uint32_t *n_cregs, uint32_t cregs_cap) {
    try {
        Parser p(lex(src ? src : ""));
        p.run();
        *n_qubits = 0;
        for (const Reg &r : p.qregs) *n_qubits += r.size;
        *n_cregs = (uint32_t)p.cregs.size();
        for (size_t i = 0; i < p.cregs.size() && i < cregs_cap; ++i) cregs[i] = p.cregs[i].size;
        for (size_t i = 0; i < p.instrs.size() && (int64_t)i < cap; ++i) out[i] = p.instrs[i];
        return (int64_t)p.instrs.size();
    } catch (const std::exception &e) {
        set_error(e.what());
        return ZB_ERR_ARG;
    }
}
